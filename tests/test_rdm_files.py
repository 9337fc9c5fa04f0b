"""RDM file outputs (src/chem/rdm.cc:202-221 dump_1rdm, 271-331 dump_2rdm):
1rdm.csv, 2rdm.csv and the Dice-convention spatialRDM.txt written by the host
library from the shci_rdm arrays.  Pinned by hand-derived lines on a tiny
RDM (loops, thresholds, printf formats, orbital order) and by identities of
the parsed files on a full-space RDM: tr(gamma) = N and sum_pq D_pqqp =
N(N - 1), every element above the threshold present with its value."""
import numpy as np
import pytest

import oracle
from shci_amd import integrals as si
from shci_amd.host import dump_1rdm, dump_2rdm
from test_oracle_rdm import _fci


def _pack(D):
    """Packed 2-RDM of a dense D[p,q,r,s]: index a = p n + s, b = q n + r, a >= b."""
    n = D.shape[0]
    nn = n * n
    out = np.zeros(nn * (nn + 1) // 2)
    for p in range(n):
        for q in range(n):
            for r in range(n):
                for s in range(n):
                    a, b = p * n + s, q * n + r
                    if a >= b:
                        out[a * (a + 1) // 2 + b] = D[p, q, r, s]
    return out


def test_1rdm_csv_lines(tmp_path):
    # internal orbital i is original orbital order[i]
    order = [2, 0, 1]
    g = np.array([[1.5, 0.25, 1e-10], [0.25, 0.5, -0.125], [1e-10, -0.125, 2e-9]])
    f = tmp_path / "1rdm.csv"
    dump_1rdm(f, g, order)
    assert f.read_text().splitlines() == [
        "p,r,1rdm",
        "2,2,1.50000000000000",
        "2,0,0.250000000000000",
        "0,0,0.500000000000000",
        "0,1,-0.125000000000000",
        "1,1,2.00000000000000e-09",
    ]


def test_2rdm_csv_and_spatial_lines(tmp_path):
    n = 2
    D = np.zeros((n, n, n, n))
    # two elements and their images under the packing symmetry (p s) <-> (q r)
    D[0, 1, 1, 0] = D[1, 0, 0, 1] = 0.75
    D[1, 1, 0, 0] = D[0, 0, 1, 1] = -2.5e-7  # below the spatialRDM cut, above the csv cut
    order = [1, 0]
    r2 = _pack(D)
    f = tmp_path / "2rdm.csv"
    dump_2rdm(f, r2, n, order, csv=True)
    # p <= q; for p == q only s <= r; loops p, q, s, r (r fastest); original indices
    assert f.read_text().splitlines() == [
        "p,q,r,s,2rdm",
        "1,1,0,0,-2.50000000000000e-07",
        "1,0,0,1,0.750000000000000",
        "0,0,1,1,-2.50000000000000e-07",
    ]
    g = tmp_path / "spatialRDM.txt"
    dump_2rdm(g, r2, n, order, csv=False)
    # original orbitals printed as p, q, s, r with the value D at the internal
    # indices; |x| > 1e-6.  Internal (0,1,1,0) is original (1,0,0,1) and
    # internal (1,0,0,1) original (0,1,1,0).
    assert g.read_text().splitlines() == [
        "2",
        "  0     1     0     1         0.75",
        "  1     0     1     0         0.75",
    ]


def _parse_csv(path, k):
    lines = path.read_text().splitlines()[1:]
    return {tuple(int(t) for t in l.split(",")[:k]): float(l.split(",")[k]) for l in lines}


def test_full_space_files_identities(tmp_path):
    n, nu, nd = 6, 3, 2
    _, h1, eri, _ = si.synthetic(n, seed=17)
    orc, e, vec = _fci(n, nu, nd, h1, eri)
    r1, r2 = orc.rdm(vec)
    N = nu + nd
    order = np.array([3, 5, 0, 1, 4, 2])
    inv = np.argsort(order)
    dump_1rdm(tmp_path / "1rdm.csv", r1, order)
    dump_2rdm(tmp_path / "2rdm.csv", r2, n, order, csv=True)
    dump_2rdm(tmp_path / "spatialRDM.txt", r2, n, order, csv=False)
    g = _parse_csv(tmp_path / "1rdm.csv", 2)
    assert abs(sum(v for (p, r), v in g.items() if p == r) - N) < 1e-12
    D = oracle.unpack_2rdm(r2, n)
    # every element with |D| >= 1e-9 in the csv's index range, at its original indices
    c2 = _parse_csv(tmp_path / "2rdm.csv", 4)
    want = {}
    for p in range(n):
        for q in range(p, n):
            for s in range(n):
                for r in range(n):
                    if (p == q and s > r) or abs(D[p, q, r, s]) < 1e-9:
                        continue
                    want[(order[p], order[q], order[r], order[s])] = D[p, q, r, s]
    assert set(c2) == set(want)
    for k, v in want.items():
        assert abs(c2[k] - v) <= 1e-14 * max(1.0, abs(v))
    # spatialRDM: sum over its lines of D_pqqp (Dice p q s r with q == s... read back as p,q,s,r)
    lines = (tmp_path / "spatialRDM.txt").read_text().splitlines()
    assert int(lines[0]) == n
    tr = 0.0
    for l in lines[1:]:
        p, q, s, r, v = l.split()
        p, q, s, r = int(p), int(q), int(s), int(r)
        assert abs(float(v) - D[inv[p], inv[q], inv[r], inv[s]]) <= 5e-8 * max(1.0, abs(float(v)))
        if q == r and p == s:
            tr += float(v)
    full_tr = sum(D[p, q, q, p] for p in range(n) for q in range(n) if abs(D[p, q, q, p]) > 1e-6)
    assert abs(tr - full_tr) < 1e-6
    assert abs(np.einsum("pqqp->", D) - N * (N - 1)) < 1e-10


@pytest.mark.gpu
def test_gpu_rdm_files_match_oracle(tmp_path):
    """The files written from the GPU RDM list the oracle RDM's elements with
    the same values (selected space, internal = original order)."""
    from shci_amd.hamiltonian import Hamiltonian

    from bruteforce import dets_all

    n, nu, nd = 6, 3, 2
    _, h1, eri, _ = si.synthetic(n, seed=17)
    orc, e, vec = _fci(n, nu, nd, h1, eri)
    dets = dets_all(n, nu, nd)  # the det order of _fci: HF first
    hf = ((1 << nu) - 1, (1 << nd) - 1)
    dets.remove(hf)
    dets = [hf] + dets
    up = np.array([d[0] for d in dets], dtype=np.uint64)
    dn = np.array([d[1] for d in dets], dtype=np.uint64)
    ham = Hamiltonian(n, h1, eri, 0.0, store="fp64", layout="segments")
    ham.update(up, dn)
    g1, g2 = ham.rdm(vec)
    r1, r2 = orc.rdm(vec)
    for name, a, b in (("gpu", g1, g2), ("orc", r1, r2)):
        dump_1rdm(tmp_path / f"{name}_1.csv", a)
        dump_2rdm(tmp_path / f"{name}_2.csv", b, n)
    for k in (1, 2):
        ga = _parse_csv(tmp_path / f"gpu_{k}.csv", 2 * k)
        oa = _parse_csv(tmp_path / f"orc_{k}.csv", 2 * k)
        common = set(ga) & set(oa)
        for key in set(ga) ^ set(oa):  # only elements at the 1e-9 cut may differ
            v = ga.get(key, oa.get(key))
            assert abs(v) < 1.1e-9
        for key in common:
            assert abs(ga[key] - oa[key]) <= 1e-11 * max(1.0, abs(oa[key]))
