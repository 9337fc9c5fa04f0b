"""The oracle's time-reversal-symmetric element (src/base_system.h:86-101)
pinned from first principles on a full space: with P mapping each canonical
det (up <= dn) to (|up,dn> + |dn,up>)/sqrt2, or |up,up>, the dense H_ts of
elem_time_sym equals P^T H P of the full dense H; its spectrum is the part of
H's spectrum whose eigenvectors are even under the spin flip."""
import itertools

import numpy as np

import oracle
from shci_amd import integrals as si


def full_space(n, ne):
    strs = [sum(1 << o for o in c) for c in itertools.combinations(range(n), ne)]
    return [(u, d) for u in strs for d in strs]


def ts_basis(space):
    return [(u, d) for u, d in space if u <= d]


def test_time_sym_element_is_projected_h():
    n, ne = 6, 2
    _, h1, eri, core = si.synthetic(n, seed=17)
    orc = oracle.System.from_arrays(n, ne, ne, h1, eri, core)
    space = full_space(n, ne)
    idx = {d: k for k, d in enumerate(space)}
    H = np.array([[orc.elem(a[0], a[1], b[0], b[1]) for b in space] for a in space])
    red = ts_basis(space)
    P = np.zeros((len(space), len(red)))
    for j, (u, d) in enumerate(red):
        if u == d:
            P[idx[(u, d)], j] = 1.0
        else:
            P[idx[(u, d)], j] = P[idx[(d, u)], j] = 1 / np.sqrt(2)
    Hts = np.array([[orc.elem_time_sym(a[0], a[1], b[0], b[1]) for b in red] for a in red])
    np.testing.assert_allclose(Hts, P.T @ H @ P, rtol=0, atol=1e-12)
    # spectrum: the spin-flip-even eigenvalues of H
    R = np.zeros_like(H)
    for k, (u, d) in enumerate(space):
        R[idx[(d, u)], k] = 1.0
    w, v = np.linalg.eigh(H)
    even = [w[k] for k in range(len(w)) if np.linalg.norm(R @ v[:, k] - v[:, k]) < 1e-6]
    # degenerate levels mix even and odd vectors: compare with eigh of H in the even subspace instead
    we = np.linalg.eigvalsh(P.T @ H @ P)
    assert np.allclose(np.sort(np.linalg.eigvalsh(Hts)), we, atol=1e-10)
    for x in even:
        assert np.min(np.abs(we - x)) < 1e-9
