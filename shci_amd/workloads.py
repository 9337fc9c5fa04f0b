"""Benchmark workloads: seeded synthetic integrals (shci_amd.integrals) and the
variational determinant list grown by the reference's selection loop
(src/solver/solver.h:338-539) -- heat-bath selection (on the device by
default, or the host's queue walk), GPU Hamiltonian and Davidson for the
coefficients."""
from __future__ import annotations

import sys
import time

import numpy as np

from . import integrals as si
from .hamiltonian import Davidson, Hamiltonian
from .host import HostSystem
from .selection import GpuSystem

__all__ = ["Workload", "grow_det_list", "det_hash", "SCHEDULES"]

# eps sweeps per config: every eps is applied once, the last one until the det
# count grows by less than 1% (at most 3 sweeps) or the target is reached.
# The stated eps of the workload text gives far fewer determinants than its
# stated size with this integral generator (e.g. cfg 2 at eps 3e-5: 9.3e4
# dets), so the sweeps continue to smaller eps until the size is reached.
SCHEDULES = {
    1: (5e-4, 2e-4, 1e-4),
    2: (1e-3, 1e-4, 1e-5, 3e-6, 1e-6, 6e-7, 4e-7),
    3: (1e-3, 1e-4, 1e-5, 3e-6, 1e-6, 5e-7, 3e-7, 2e-7, 1.6e-7),
    4: (1e-3, 1e-4, 1e-5, 3e-6, 1e-6, 5e-7),
    5: (1e-3, 1e-4, 1e-5, 3e-6, 1e-6, 8e-7),
}


class Workload:
    def __init__(self, cfg: int):
        self.cfg = cfg
        self.spec = si.CONFIGS[cfg]
        self.n, self.h1, self.eri, self.core = si.synthetic(cfg)
        self.up = self.dn = None
        self.log = []


def det_hash(up, dn) -> str:
    """Digest of a det list in its order (run-to-run reproducibility)."""
    import hashlib

    h = hashlib.sha256(np.ascontiguousarray(up, dtype=np.uint64).tobytes())
    h.update(np.ascontiguousarray(dn, dtype=np.uint64).tobytes())
    return h.hexdigest()[:16]


def grow_det_list(cfg: int, target: int | None = None, device: int = 0, verbose: bool = True,
                  schedule=None, max_final_sweeps: int = 3, selection: str = "device") -> Workload:
    """Grow the variational space of config ``cfg`` until ``target`` dets.
    ``selection``: ``"device"`` (shci_select) or ``"host"`` (the host's queue
    walk); both give the same det lists."""
    w = Workload(cfg)
    spec = w.spec
    target = target or spec.target_dets
    schedule = schedule or SCHEDULES[cfg]
    host = HostSystem.from_arrays(w.n, spec.n_up, spec.n_dn, w.h1, w.eri, w.core, n_states=spec.n_states)
    # The selection compares |H_ij c_i| with eps, so the coefficients that
    # drive it must not depend on the summation order of the products.  The
    # CSR sums rows in a fixed order; the ~1e7-det configs need the segment
    # layout's memory, whose products add member sums with atomics.  There
    # the Davidson runs in FP64 (values and vectors): the order changes the
    # coefficients by ~1e-16 relative, and no |H c| lies that close to eps
    # (with FP32 vectors, ~1e-7, a few borderline dets moved between runs).
    layout = "csr" if cfg in (1, 2, 4) else "segments"
    store, mode = ("fp32", "mixed") if layout == "csr" else ("fp64", "fp64")
    ham = Hamiltonian(w.n, w.h1, w.eri, w.core, device=device, store=store, layout=layout)
    if selection == "device":
        host = GpuSystem(ham, host.hf_up, host.hf_dn)
    dav = Davidson(spec.n_states)
    coefs = np.zeros((spec.n_states, 1))
    coefs[0, 0] = 1.0
    coefs[1:, 0] = 1e-16
    steps = [(e, 1) for e in schedule[:-1]] + [(schedule[-1], max_final_sweeps)]
    t_start = time.time()
    for eps, sweeps in steps:
        for _ in range(sweeps):
            t0 = time.time()
            n_before = host.n_dets
            n_new = host.select(coefs, eps)
            up, dn = host.dets()
            t_sel = time.time() - t0
            pad = np.full((coefs.shape[0], len(up) - coefs.shape[1]), 1e-16)
            coefs = np.concatenate([coefs, pad], axis=1)
            if spec.n_states > 1 and coefs.shape[1] > 1 and n_before == 1:
                # unit vectors for the excited states (solver.h:479-485)
                for s in range(1, spec.n_states):
                    coefs[s, s] = 1.0
            t0 = time.time()
            ham.update(up, dn)
            t_build = time.time() - t0
            t0 = time.time()
            dav.diagonalize(ham.matrix, coefs, 1e-6, mode=mode, max_calls=2)
            t_dav = time.time() - t0
            coefs = np.atleast_2d(dav.get_lowest_eigenvectors())
            _, _, nel = ham.matrix.counts()
            rec = dict(eps=eps, n_dets=len(up), n_new=n_new, nnz=nel, energy=float(dav.lowest_eigenvalues[0]),
                       t_select=t_sel, t_build=t_build, t_davidson=t_dav)
            w.log.append(rec)
            if verbose:
                print(f"[workload cfg{cfg}] eps {eps:.1e} dets {len(up):,} new {n_new:,} nnz {nel:,} "
                      f"E {rec['energy']:.8f} sel {t_sel:.2f}s build {t_build:.2f}s dav {t_dav:.2f}s", file=sys.stderr,
                      flush=True)
            if len(up) >= target or n_new < 0.01 * n_before:
                break
        if host.n_dets >= target:
            break
    w.up, w.dn = host.dets()
    w.coefs = coefs
    w.det_hash = det_hash(w.up, w.dn)
    # No clear: the context keeps the growth's device blocks, and the next
    # build of this list (on the same context) reuses them instead of asking
    # the HIP runtime for memory again.
    del ham
    if verbose:
        print(f"[workload cfg{cfg}] {len(w.up):,} dets in {time.time() - t_start:.1f}s", file=sys.stderr, flush=True)
    return w
