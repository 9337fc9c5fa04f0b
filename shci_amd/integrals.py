"""Integrals for the variational hot path: the seeded synthetic generator of
the benchmark workloads and an FCIDUMP writer.

Layout shared by every component (host library, GPU C-ABI, oracle):

* ``h1``  -- dense ``n x n`` float64, symmetric; ``h1[p, q]`` is the value the
  reference's ``Integrals::get_1b(p, q)`` returns (src/chem/integrals.cc:390-393).
* ``eri`` -- chemists' ``(pq|rs)`` packed by the reference key
  ``combine4(p, q, r, s) = combine2(combine2(p, q), combine2(r, s))``
  (src/chem/integrals.cc:410-422), i.e. one value per 8-fold symmetry class,
  length ``P(P+1)/2`` with ``P = n(n+1)/2``.
* ``core`` -- the core energy.

The reference drops ``|x| < 1e-9`` when reading an FCIDUMP
(src/chem/integrals.cc:89); :func:`synthetic` applies the same cut so that the
in-memory arrays equal what a round trip through an FCIDUMP would give.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

__all__ = [
    "SystemSpec", "CONFIGS", "synthetic", "write_fcidump", "combine2", "combine4",
    "pair_index_table", "unpack_eri", "pack_eri",
]


@dataclasses.dataclass(frozen=True)
class SystemSpec:
    """One benchmark workload: orbitals, electrons, variational eps schedule."""

    cfg: int
    n_orb: int
    n_up: int
    n_dn: int
    eps_vars: tuple
    eps_schedule: tuple
    n_states: int = 1
    target_dets: int = 0

    @property
    def seed(self) -> int:
        return 1000 + self.cfg


# The five workloads (synthetic seeded FCIDUMPs, C1 symmetry, core 0).
CONFIGS = {
    1: SystemSpec(1, 20, 5, 5, (1e-4,), (5e-4, 2e-4), target_dets=100_000),
    2: SystemSpec(2, 40, 8, 8, (3e-5,), (1e-3, 3e-4, 1e-4), target_dets=2_000_000),
    3: SystemSpec(3, 60, 10, 10, (2e-5,), (1e-3, 3e-4, 1e-4, 5e-5), target_dets=10_000_000),
    4: SystemSpec(4, 30, 7, 7, (3e-5,), (1e-3, 3e-4, 1e-4), target_dets=1_000_000),
    5: SystemSpec(5, 50, 9, 9, (2e-5,), (1e-3, 3e-4, 1e-4, 5e-5), n_states=4, target_dets=10_000_000),
}


def combine2(a, b):
    """src/chem/integrals.cc:410-416 (works on numpy arrays too)."""
    a = np.asarray(a, dtype=np.int64)
    b = np.asarray(b, dtype=np.int64)
    hi = np.maximum(a, b)
    lo = np.minimum(a, b)
    return hi * (hi + 1) // 2 + lo


def combine4(p, q, r, s):
    """src/chem/integrals.cc:418-422."""
    return combine2(combine2(p, q), combine2(r, s))


def pair_index_table(n: int) -> np.ndarray:
    """``T[p, q] = combine2(p, q)`` as an ``n x n`` int64 table."""
    idx = np.arange(n)
    return combine2(idx[:, None], idx[None, :])


def unpack_eri(eri: np.ndarray, n: int) -> np.ndarray:
    """Dense ``n^4`` chemists' tensor ``v[p, q, r, s] = (pq|rs)`` from the packed layout."""
    t = pair_index_table(n)
    return eri[combine2(t[:, :, None, None], t[None, None, :, :])]


def pack_eri(v: np.ndarray) -> np.ndarray:
    """Packed layout from a dense 8-fold symmetric tensor (takes the canonical entry)."""
    n = v.shape[0]
    npair = n * (n + 1) // 2
    p_of = np.empty(npair, dtype=np.int64)
    q_of = np.empty(npair, dtype=np.int64)
    k = 0
    for p in range(n):
        for q in range(p + 1):
            p_of[k], q_of[k] = p, q
            k += 1
    ab, cd = np.tril_indices(npair)
    return v[p_of[ab], q_of[ab], p_of[cd], q_of[cd]].copy()


def synthetic(cfg_or_n, n_up: int | None = None, seed: int | None = None):
    """Seeded synthetic integrals.

    ``h_pp = -2 + 0.25 p``; ``h_pq = h_qp ~ N(0, sd = 0.05 exp(-|p-q|/4))`` for
    ``p < q``; ``(pq|rs) = sum_{L=1..60} B^L_pq B^L_rs`` with ``B^L`` symmetric,
    ``B^L_pq ~ N(0, sd = 0.1 exp(-|p-q|/3))``.  Draw order: the ``h`` upper
    triangle row-major, then for each ``L`` the ``B^L`` upper triangle
    (including the diagonal) row-major, from ``numpy.random.default_rng(seed)``.
    Values with ``|x| < 1e-9`` are zeroed (the FCIDUMP read cut).

    Returns ``(n, h1, eri, core)``.
    """
    if isinstance(cfg_or_n, SystemSpec):
        spec = cfg_or_n
        n, seed = spec.n_orb, spec.seed
    elif n_up is None and seed is None:
        spec = CONFIGS[int(cfg_or_n)]
        n, seed = spec.n_orb, spec.seed
    else:
        n = int(cfg_or_n)
        if seed is None:
            raise ValueError("seed required with an explicit orbital count")
    rng = np.random.default_rng(seed)
    h1 = np.zeros((n, n))
    for p in range(n):
        h1[p, p] = -2.0 + 0.25 * p
    iu = np.triu_indices(n, 1)
    sd_h = 0.05 * np.exp(-np.abs(iu[0] - iu[1]) / 4.0)
    vals = rng.normal(0.0, 1.0, size=len(iu[0])) * sd_h
    h1[iu] = vals
    h1[(iu[1], iu[0])] = vals
    n_l = 60
    iu0 = np.triu_indices(n, 0)
    sd_b = 0.1 * np.exp(-np.abs(iu0[0] - iu0[1]) / 3.0)
    # Pair index of (p, q) in combine2 order for the upper-triangle draws.
    pidx = combine2(iu0[0], iu0[1])
    npair = n * (n + 1) // 2
    bmat = np.zeros((n_l, npair))
    for L in range(n_l):
        bmat[L, pidx] = rng.normal(0.0, 1.0, size=len(pidx)) * sd_b
    g = bmat.T @ bmat  # g[ab, cd] = sum_L B_ab B_cd
    ab, cd = np.tril_indices(npair)
    eri = g[ab, cd].copy()  # combine2(ab, cd) order with ab >= cd
    eri[np.abs(eri) < 1e-9] = 0.0
    h1[np.abs(h1) < 1e-9] = 0.0
    return n, h1, eri, 0.0


def write_fcidump(path, n, h1, eri, core, n_elec, ms2=0):
    """FCIDUMP with unique 8-fold entries (``%.17e``), C1 symmetry."""
    lines = [f" &FCI NORB={n},NELEC={n_elec},MS2={ms2},", "  ORBSYM=" + ",".join(["1"] * n) + ",",
             "  ISYM=1,", " &END"]
    npair = n * (n + 1) // 2
    p_of = np.empty(npair, dtype=np.int64)
    q_of = np.empty(npair, dtype=np.int64)
    k = 0
    for p in range(n):
        for q in range(p + 1):
            p_of[k], q_of[k] = p, q
            k += 1
    ab, cd = np.tril_indices(npair)
    nz = eri != 0.0
    out = []
    for v, a, c in zip(eri[nz], ab[nz], cd[nz]):
        out.append("%.17e %d %d %d %d" % (v, p_of[a] + 1, q_of[a] + 1, p_of[c] + 1, q_of[c] + 1))
    for p in range(n):
        for q in range(p + 1):
            if h1[p, q] != 0.0:
                out.append("%.17e %d %d 0 0" % (h1[p, q], p + 1, q + 1))
    out.append("%.17e 0 0 0 0" % core)
    with open(path, "w") as f:
        f.write("\n".join(lines + out) + "\n")


def hf_bits(n_up: int, n_dn: int):
    """Lowest-filled HF determinant (src/chem/integrals.cc:176-185)."""
    return (1 << n_up) - 1, (1 << n_dn) - 1


def n_dets_fci(n: int, n_up: int, n_dn: int) -> int:
    return math.comb(n, n_up) * math.comb(n, n_dn)
