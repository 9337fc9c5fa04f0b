"""Pins Davidson::diagonalize (src/solver/davidson.cc:8-196) with the
reference's own known-answer test DavidsonTest.HilbertSystem
(src/solver/davidson_test.cc:5-51) and scipy's dense eigensolver."""
import numpy as np
import scipy.linalg

import oracle

N = 1000


def _hilbert(i, j):
    return -1.0 / (2 * i + 1) if i == j else -1.0 / 10.0 / (i + j + 1)


def test_hilbert_known_answer():
    entries = [(i, j, _hilbert(i, j)) for i in range(N) for j in range(i, N)]
    init = np.zeros((1, N))
    init[0, 0] = 1.0
    ev, evec, conv, traj = oracle.sparse_davidson(N, entries, init, 1e-6)
    assert abs(ev[0] - (-1.00956719)) < 1e-6
    expected = [0.99292536, 0.08026708, 0.04720676, 0.03412438, 0.02694173]
    np.testing.assert_allclose(evec[0, :5], expected, atol=1e-4)
    # and against a dense solve of the same closed-form matrix
    H = np.fromfunction(lambda i, j: np.where(i == j, -1.0 / (2 * i + 1), -0.1 / (i + j + 1)), (N, N))
    w = scipy.linalg.eigh(H, eigvals_only=True, subset_by_index=[0, 0])
    assert abs(ev[0] - w[0]) < 1e-6


def test_spmv_matches_dense_symmetric_expansion():
    from shci_amd import integrals as si

    n, nu, nd = 10, 3, 3
    _, h1, eri, _ = si.synthetic(n, seed=4)
    from bruteforce import dets_all

    dets = dets_all(n, nu, nd)[:400]
    up = np.array([d[0] for d in dets], dtype=np.uint64)
    dn = np.array([d[1] for d in dets], dtype=np.uint64)
    orc = oracle.System.from_arrays(n, nu, nd, h1, eri, 0.0)
    orc.update(up, dn)
    rp, cols, vals = orc.export_upper()
    H = np.zeros((len(dets), len(dets)))
    rows = np.repeat(np.arange(len(dets)), np.diff(rp))
    H[rows, cols] = vals
    H = H + np.triu(H, 1).T
    x = np.random.default_rng(1).uniform(-1, 1, len(dets))
    np.testing.assert_allclose(orc.mul(x), H @ x, rtol=1e-12, atol=1e-12)
