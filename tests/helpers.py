"""Shared test helpers: the reference's variational loop
(src/solver/solver.h:338-397, 413-539) driven by a pluggable Hamiltonian +
Davidson backend, so the oracle and the GPU path run the same iterations."""
from __future__ import annotations

import numpy as np


def variation(host, backend, eps_vars, eps_schedule, target_error=5e-5, max_iters=None, n_states=1,
              energy_stop=False):
    """Runs extras (one iteration each) then each eps until converged.

    ``backend(up, dn, init, tol)`` builds/updates H over the det list and runs
    one Davidson call; returns (evals, evecs, converged, info).  Yields one
    record per iteration.  ``energy_stop``: an eps also ends when every state's
    energy changed by at most target_error * 0.001 since the previous iteration
    of that eps -- the check of src/solver/solver.h:513-517, which the revision
    that wrote examples/C/out applied and the current code never reaches (its
    ``i_state == n_states`` test sits inside the loop over i_state)."""
    coefs = np.zeros((n_states, 1))
    coefs[0, 0] = 1.0
    for s in range(1, n_states):
        coefs[s, 0] = 1e-16
    host.reset()
    it_sched = list(eps_schedule)
    eps_prev = float("inf")
    count = 0
    for eps in eps_vars:
        plan = [(e, False) for e in it_sched if e < eps_prev and e > eps] + [(eps, True)]
        for e, until in plan:
            dets_converged = False
            n_dets = host.n_dets
            e_prev = np.zeros(n_states)
            while True:
                if not dets_converged:
                    n_new = host.select(coefs, e)
                    up, dn = host.dets()
                    new = np.full((n_states, len(up) - coefs.shape[1]), 1e-16)
                    coefs = np.concatenate([coefs, new], axis=1)
                else:
                    n_new = 0
                tol = target_error / 500000 if until else target_error / 50
                evals, evecs, conv, info = backend(up, dn, coefs, tol)
                coefs = np.atleast_2d(evecs)
                rec = dict(eps=e, n_dets=len(up), n_new=n_new, evals=np.array(evals), converged=conv, **info)
                yield rec
                count += 1
                if max_iters is not None and count >= max_iters:
                    return
                if len(up) < n_dets * 1.001:
                    dets_converged = True
                n_dets = len(up)
                if energy_stop and until and np.all(np.abs(np.asarray(evals)[:n_states] - e_prev) <= target_error * 0.001):
                    break
                e_prev = np.asarray(evals, dtype=float)[:n_states].copy()
                if not until or (dets_converged and conv):
                    break
        eps_prev = eps
