"""Shared test helpers: the reference's variational loop
(src/solver/solver.h:338-397, 413-539) driven by a pluggable Hamiltonian +
Davidson backend, so the oracle and the GPU path run the same iterations."""
from __future__ import annotations

import numpy as np


def variation(host, backend, eps_vars, eps_schedule, target_error=5e-5, max_iters=None, n_states=1):
    """Runs extras (one iteration each) then each eps until converged.

    ``backend(up, dn, init, tol)`` builds/updates H over the det list and runs
    one Davidson call; returns (evals, evecs, converged, info).  Yields one
    record per iteration."""
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
                if not until or (dets_converged and conv):
                    break
        eps_prev = eps
