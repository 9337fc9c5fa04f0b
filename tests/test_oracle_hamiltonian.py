"""Pins Hamiltonian<S>::update (src/solver/hamiltonian.h:87-598): the stored
connection pattern equals a brute-force O(N^2) enumeration of pairs with
excitation <= 2 and |H| >= 1e-12, and the append-only incremental update
gives the same matrix as a from-scratch build."""
import numpy as np
import pytest

import oracle
from bruteforce import dets_all
from shci_amd import integrals as si


def _pairs(rp, cols, vals, n):
    rows = np.repeat(np.arange(n), np.diff(rp))
    i = np.minimum(rows, cols)
    j = np.maximum(rows, cols)
    key = i * n + j
    o = np.argsort(key)
    return key[o], vals[o]


@pytest.mark.parametrize("seed", [0, 1])
def test_pattern_brute_force(seed):
    n, nu, nd = 10, 3, 3
    _, h1, eri, _ = si.synthetic(n, seed=30 + seed)
    rng = np.random.default_rng(seed)
    all_d = dets_all(n, nu, nd)
    pick = rng.choice(len(all_d), 1500, replace=False)
    dets = [all_d[k] for k in pick]
    up = np.array([d[0] for d in dets], dtype=np.uint64)
    dn = np.array([d[1] for d in dets], dtype=np.uint64)
    orc = oracle.System.from_arrays(n, nu, nd, h1, eri, 0.0)
    orc.update(up, dn)
    rp, cols, vals = orc.export_upper()
    N = len(dets)
    # brute force: all pairs i<j with popcount diff <= 4 total and per-spin <= 4
    xu = up[:, None] ^ up[None, :]
    xd = dn[:, None] ^ dn[None, :]
    pc = np.vectorize(lambda v: bin(int(v)).count("1"))(xu) + np.vectorize(lambda v: bin(int(v)).count("1"))(xd)
    ii, jj = np.nonzero(np.triu(pc <= 4, 1))
    keep = [abs(orc.elem(int(up[a]), int(dn[a]), int(up[b]), int(dn[b]), -1)) >= 1e-12 for a, b in zip(ii, jj)]
    bf = np.sort(ii[keep] * N + jj[keep])
    got, _ = _pairs(rp, cols, vals, N)
    diag = got[got % (N + 1) == 0]
    off = got[got % (N + 1) != 0]
    assert len(diag) == N
    assert np.array_equal(np.sort(off), bf)
    st = orc.stats()
    assert st["nnz"] == 2 * st["n_stored"] - N


def test_incremental_equals_scratch():
    n, nu, nd = 12, 3, 3
    _, h1, eri, _ = si.synthetic(n, seed=2)
    rng = np.random.default_rng(5)
    all_d = dets_all(n, nu, nd)
    pick = rng.choice(len(all_d), 3000, replace=False)
    up = np.array([all_d[k][0] for k in pick], dtype=np.uint64)
    dn = np.array([all_d[k][1] for k in pick], dtype=np.uint64)
    inc = oracle.System.from_arrays(n, nu, nd, h1, eri, 0.0)
    for m in (500, 1800, 3000):  # 1800 < 1.15 * 500? no -> partner-sorted; 3000 < 1.15*1800? no
        inc.update(up[:m], dn[:m])
    inc2 = oracle.System.from_arrays(n, nu, nd, h1, eri, 0.0)
    for m in (2000, 2200, 2400, 3000):  # growth < 15%: sort_by_det_id branch
        inc2.update(up[:m], dn[:m])
    scratch = oracle.System.from_arrays(n, nu, nd, h1, eri, 0.0)
    scratch.update(up, dn)
    N = len(up)
    ks, vs = _pairs(*scratch.export_upper(), N)
    for sysm in (inc, inc2):
        k, v = _pairs(*sysm.export_upper(), N)
        assert np.array_equal(k, ks)
        np.testing.assert_array_equal(v, vs)
    x = rng.uniform(-1, 1, N)
    np.testing.assert_allclose(inc.mul(x), scratch.mul(x), rtol=1e-12, atol=1e-12)


def test_update_existing_elems_matches_rebuild():
    n, nu, nd = 10, 3, 2
    _, h1, eri, _ = si.synthetic(n, seed=8)
    _, h1b, erib, _ = si.synthetic(n, seed=9)
    dets = dets_all(n, nu, nd)[:800]
    up = np.array([d[0] for d in dets], dtype=np.uint64)
    dn = np.array([d[1] for d in dets], dtype=np.uint64)
    a = oracle.System.from_arrays(n, nu, nd, h1, eri, 0.0)
    a.update(up, dn)
    a.set_integrals(h1b, erib, 0.0)
    a.update_existing_elems()
    rp, cols, vals = a.export_upper()
    for i in range(0, 800, 97):
        for k in range(rp[i], rp[i + 1]):
            j = cols[k]
            assert vals[k] == a.elem(int(up[i]), int(dn[i]), int(up[j]), int(dn[j]), -1 if i != j else 0)


def test_upper_digest_detects_plausible_mistakes():
    """oracle.upper_digest (the order-free comparison of large exports):
    equal under any within-row order, different after a one-ulp value
    change, a sign flip, a changed column, an entry moved to another row
    or a dropped entry."""
    rng = np.random.default_rng(5)
    n = 50
    rows = [np.sort(rng.choice(np.arange(i, n), size=min(n - i, 6), replace=False)) for i in range(n)]
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int64)
    cols = np.concatenate(rows).astype(np.int64)
    vals = rng.normal(size=len(cols))
    d0 = oracle.upper_digest(rp, cols, vals)
    perm = np.concatenate([rp[i] + rng.permutation(rp[i + 1] - rp[i]) for i in range(n)])
    assert oracle.upper_digest(rp, cols[perm], vals[perm]) == d0
    v1 = vals.copy()
    v1[17] = np.nextafter(v1[17], np.inf)
    assert oracle.upper_digest(rp, cols, v1) != d0
    v2 = vals.copy()
    v2[3] = -v2[3]
    assert oracle.upper_digest(rp, cols, v2) != d0
    c1 = cols.copy()
    c1[40] += 1
    assert oracle.upper_digest(rp, c1, vals) != d0
    rp1 = rp.copy()
    rp1[5] += 1  # the first entry of row 5 now belongs to row 4
    assert oracle.upper_digest(rp1, cols, vals) != d0
    rp2 = rp.copy()
    rp2[-1] -= 1
    assert oracle.upper_digest(rp2, cols[:-1], vals[:-1]) != d0
