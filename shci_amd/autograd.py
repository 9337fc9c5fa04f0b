"""The variational energy as a differentiable function of the integrals.

``E(h, eri)`` is the lowest eigenvalue (or the state average over
``n_states`` roots) of the variational Hamiltonian over a fixed determinant
list.  Forward: GPU Hamiltonian build + Davidson (include/shci_hip.h).
Backward: the Hellmann-Feynman adjoint -- the converged vector contracted over
the same connection list into the 1-RDM and 2-RDM by the RDM kernel
(shci_rdm, reference RDM::get_1rdm / get_2rdm, src/chem/rdm.cc), then
dE/dh and dE/deri:

* ``h`` enters symmetrised, ``h_used = (h + h^T) / 2``, so
  ``dE/dh = (gamma + gamma^T) / 2``;
* ``eri`` is the packed vector of chemists' integrals, one per 8-fold class
  (key ``combine4``); since ``E = core + sum h gamma + 1/2 sum_pqrs D_pqrs (ps|qr)``
  (rdm.cc:731-745), ``dE/deri[K] = 1/2 sum_{combine4(p,s,q,r) = K} D_pqrs``.

The connection pattern is built with every excitation-<=2 pair kept
(SHCI_KEEP_ZEROS): a pair whose element is zero for these integrals still has
a nonzero derivative.  Later evaluations with new integrals recompute the
values on the fixed pattern (shci_update_existing_elems).
"""
from __future__ import annotations

import numpy as np
import torch

from . import integrals as si
from .hamiltonian import Davidson, Hamiltonian

__all__ = ["VariationalSolver", "VariationalEnergy", "variational_energy"]


class VariationalSolver:
    """Fixed determinant list + device Hamiltonian; warm-started solves."""

    def __init__(self, n_orb, up, dn, core=0.0, n_states=1, device=0, tol=1e-13, mode="fp64", max_calls=20,
                 res_tol=1e-9, max_rounds=40, layout="csr"):
        self.n = int(n_orb)
        self.up = np.ascontiguousarray(up, dtype=np.uint64)
        self.dn = np.ascontiguousarray(dn, dtype=np.uint64)
        self.core = float(core)
        self.n_states = n_states
        self.device = device
        self.tol = tol
        self.mode = mode
        self.max_calls = max_calls
        self.res_tol = res_tol
        self.max_rounds = max_rounds
        self.layout = layout
        self.residual = np.inf
        self.ham = None
        self.coefs = None
        npair = self.n * (self.n + 1) // 2
        self.n_eri = npair * (npair + 1) // 2
        self._keys = None

    def solve(self, h1: np.ndarray, eri: np.ndarray):
        h1 = 0.5 * (h1 + h1.T)
        if self.ham is None:
            self.ham = Hamiltonian(self.n, h1, eri, self.core, device=self.device,
                                   store="both" if self.mode == "mixed" else "fp64", keep_zeros=True,
                                   layout=self.layout)
            self.ham.update(self.up, self.dn)
        else:
            self.ham.set_integrals(self.n, h1, eri, self.core)
            self.ham.update_existing_elems()
        if self.coefs is None:
            init = np.zeros((self.n_states, len(self.up)))
            for s in range(self.n_states):
                init[s, s] = 1.0
        else:
            init = self.coefs
        # The gradient <c|dH|c> is first order in the eigenvector error, so
        # iterate warm-restarted Davidson calls until the residual
        # ||H c - E c|| of every root is below res_tol.
        d = Davidson(self.n_states)
        for _ in range(self.max_rounds):
            d.diagonalize(self.ham.matrix, init, self.tol, mode=self.mode, max_calls=self.max_calls)
            c = np.atleast_2d(d.get_lowest_eigenvectors())
            ev = np.array(d.get_lowest_eigenvalues(), dtype=np.float64)
            res = 0.0
            for s in range(c.shape[0]):
                r = self.ham.matrix.mul(c[s]) - ev[s] * c[s]
                res = max(res, float(np.linalg.norm(r)))
            init = c
            if res <= self.res_tol:
                break
        self.residual = res
        self.coefs = c.copy()
        self.converged = res <= self.res_tol
        return ev

    def rdm(self, coefs=None):
        return self.ham.rdm(self.coefs if coefs is None else coefs)

    def grads(self, weights=None, coefs=None):
        """(dE/dh, dE/deri) for E = sum_s w_s E_s (default: state average) at
        the eigenvectors ``coefs`` (default: those of the latest solve).  The
        connection pattern is fixed (every excitation-<=2 pair is kept), so
        the RDMs depend on the coefficients only, whatever integrals the
        solver holds now."""
        w = np.full(self.n_states, 1.0 / self.n_states) if weights is None else np.asarray(weights, float)
        if np.any(w < 0):
            raise ValueError("negative state weights are not supported")
        c = (self.coefs if coefs is None else np.atleast_2d(coefs)) * np.sqrt(np.abs(w))[:, None]
        r1, r2 = self.rdm(c)
        g_h = 0.5 * (r1 + r1.T)
        return g_h, self.eri_grad_from_rdm(r2)

    def eri_grad_from_rdm(self, r2: np.ndarray) -> np.ndarray:
        n = self.n
        if self._keys is None:
            idx = np.arange(n)
            p, q, r, s = np.meshgrid(idx, idx, idx, idx, indexing="ij")
            a = p * n + s
            b = q * n + r
            hi = np.maximum(a, b)
            lo = np.minimum(a, b)
            self._dpos = (hi * (hi + 1) // 2 + lo).ravel()   # packed 2-RDM index of D_pqrs
            self._keys = si.combine4(p, s, q, r).ravel()      # eri key of (ps|qr)
        D = r2[self._dpos]
        return 0.5 * np.bincount(self._keys, weights=D, minlength=self.n_eri)


class VariationalEnergy(torch.autograd.Function):
    """E(h, eri) with the RDM adjoint; ``solver`` carries the determinant list.

    The forward keeps its own eigenvectors on ``ctx``: one solver may serve
    several forward calls (E(h_a) + E(h_b), a closure re-evaluated before the
    backward), and each backward contracts the vectors of its own solve."""

    @staticmethod
    def forward(ctx, h1: torch.Tensor, eri: torch.Tensor, solver: VariationalSolver):
        e = solver.solve(h1.detach().cpu().double().numpy(), eri.detach().cpu().double().numpy())
        ctx.solver = solver
        ctx.coefs = solver.coefs.copy()
        ctx.dev = h1.device
        return torch.tensor(e.mean(), dtype=torch.float64, device=h1.device)

    @staticmethod
    def backward(ctx, grad_e: torch.Tensor):
        g_h, g_eri = ctx.solver.grads(coefs=ctx.coefs)
        g = grad_e.detach().cpu().double()
        return (torch.from_numpy(g_h).to(ctx.dev) * g.to(ctx.dev),
                torch.from_numpy(g_eri).to(ctx.dev) * g.to(ctx.dev), None)


def variational_energy(h1: torch.Tensor, eri: torch.Tensor, solver: VariationalSolver) -> torch.Tensor:
    return VariationalEnergy.apply(h1, eri, solver)
