"""The oracle's selection sweep (src/solver/solver.h:435-495 with
src/chem/chem_system.cc:115-402, restated without the sorted queues) pinned
by what fixes it independently of its own formulas:

* the reference's own example run (examples/C log): per-iteration det counts
  with this selection driving the variational loop, and the same det lists
  as the host selection (which walks the sorted queues);
* the criteria's closed form for one det of coefficient c: the doubles and
  singles it adds are exactly the dets one or two excitations away with
  |H_ij| >= eps_var / |c| (S(p, r) bounds |H_ij| of a single from above, and
  a double's queue value is |H_ij|), from brute-force Slater-Condon elements;
* eps_tried_prev: a later sweep walks the doubles with eps_var / |c| <= |H|
  < 1.000000001 eps_var / |c_prev| only (singles again in full), and skips a
  det whose |c| dropped;
* several states: eps_min uses the largest |c| over the states."""
import itertools
import json

import numpy as np
import pytest

import oracle
from helpers import variation
from shci_amd import integrals as si


class OracleSystem:
    """The variational det list with the oracle's selection (the interface
    of shci_amd.host.HostSystem used by the variational loop)."""

    def __init__(self, orc, hf_up, hf_dn):
        self.orc, self.hf = orc, (hf_up, hf_dn)
        self.reset()

    def reset(self):
        self.up = np.array([self.hf[0]], dtype=np.uint64)
        self.dn = np.array([self.hf[1]], dtype=np.uint64)
        self.eps = np.full(1, 1e100)

    @property
    def n_dets(self):
        return len(self.up)

    def dets(self):
        return self.up.copy(), self.dn.copy()

    def select(self, coefs, eps_var):
        nu, nd = self.orc.select(self.up, self.dn, coefs, eps_var, self.eps)
        self.up = np.concatenate([self.up, nu])
        self.dn = np.concatenate([self.dn, nd])
        self.eps = np.concatenate([self.eps, np.full(len(nu), 1e100)])
        return len(nu)


def test_examples_C_oracle_selection(golden_dir):
    """The reference's examples/C run: det counts per iteration with the
    oracle's selection, and the same det lists as the host's queue walk."""
    from shci_amd.host import HostSystem

    gold = json.loads((golden_dir / "examples_C.json").read_text())
    host = HostSystem.from_fcidump(golden_dir / "C_FCIDUMP", 4, 2)
    orc = oracle.System.from_fcidump(golden_dir / "C_FCIDUMP", 4, 2)
    osys = OracleSystem(orc, orc.hf_up, orc.hf_dn)
    builder = oracle.System.from_fcidump(golden_dir / "C_FCIDUMP", 4, 2)
    lists = []

    def backend(up, dn, init, tol):
        builder.update(up, dn)
        ev, evec, conv, traj, _ = builder.davidson(init, tol)
        lists.append((up.copy(), dn.copy()))
        return ev, evec, conv, dict(nnz=builder.stats()["nnz"])

    recs = list(variation(osys, backend, [5e-5], [2e-3, 1e-3, 5e-4, 2e-4, 1e-4], max_iters=8))
    for k, (rec, g) in enumerate(zip(recs, gold["iterations"])):
        assert rec["n_dets"] == g["n_dets"], (k, rec["n_dets"], g["n_dets"])
        if "nnz" in g:
            assert rec["nnz"] == g["nnz"]
    # the host's queue walk gives the same lists, det for det
    builder.clear()
    hl = []

    def hbackend(up, dn, init, tol):
        builder.update(up, dn)
        ev, evec, conv, traj, _ = builder.davidson(init, tol)
        hl.append((up.copy(), dn.copy()))
        return ev, evec, conv, {}

    list(variation(host, hbackend, [5e-5], [2e-3, 1e-3, 5e-4, 2e-4, 1e-4], max_iters=8))
    assert len(hl) == len(lists)
    for (a, b), (c, d) in zip(lists, hl):
        assert np.array_equal(a, c) and np.array_equal(b, d)


def _all_dets(n, nu, nd):
    ups = [sum(1 << o for o in c) for c in itertools.combinations(range(n), nu)]
    dns = [sum(1 << o for o in c) for c in itertools.combinations(range(n), nd)]
    return [(u, d) for u in ups for d in dns]


def _exc(a, b):
    return bin(a[0] ^ b[0]).count("1") // 2 + bin(a[1] ^ b[1]).count("1") // 2


@pytest.mark.parametrize("n,nu,nd,seed", [(8, 2, 2, 3), (9, 3, 2, 5), (10, 3, 3, 11)])
def test_selection_closed_form(n, nu, nd, seed):
    n_, h1, eri, core = si.synthetic(n, seed=seed)
    orc = oracle.System.from_arrays(n, nu, nd, h1, eri, core)
    space = _all_dets(n, nu, nd)
    rng = np.random.default_rng(seed)
    det = space[int(rng.integers(len(space)))]
    hij = {}
    for c in space:
        k = _exc(det, c)
        if 1 <= k <= 2:
            hij[c] = abs(orc.elem(det[0], det[1], c[0], c[1], k))
    vals = np.array(sorted(v for v in hij.values() if v > 0))
    for q in (0.0, 0.3, 0.7, 0.95):
        eps = float(np.quantile(vals, q)) * (1 - 1e-9) if q > 0 else 1e-14
        coef = 0.5
        want = sorted(c for c, v in hij.items() if v >= eps / coef)
        eps_tried = np.full(1, 1e100)
        nu_, nd_ = orc.select([det[0]], [det[1]], np.array([[coef]]), eps, eps_tried)
        got = sorted(zip(nu_.tolist(), nd_.tolist()))
        assert got == want, (q, len(got), len(want))
        assert list(zip(nu_.tolist(), nd_.tolist())) == got  # (up, dn) ascending
        assert eps_tried[0] == 1.000000001 * (eps / coef)
    # eps_tried_prev = 1.000000001 eps_min: the same coefficient walks the det
    # again (singles have no upper cut) but adds only doubles in
    # [eps_min, eps_tried_prev); a smaller one skips the det; a larger one
    # walks the doubles below the previous eps_min and every single above the
    # new one, minus the dets already selected
    eps = float(np.quantile(vals, 0.8))
    eps_tried = np.full(1, 1e100)
    first = orc.select([det[0]], [det[1]], np.array([[1.0]]), eps, eps_tried)
    prev = eps_tried[0]
    up = np.concatenate([[det[0]], first[0]]).astype(np.uint64)
    dn = np.concatenate([[det[1]], first[1]]).astype(np.uint64)
    et = np.concatenate([eps_tried, np.full(len(first[0]), 1e100)])
    sel1 = set(zip(up.tolist(), dn.tolist()))
    for c0, walked in ((1.0, True), (0.5, False)):
        coefs = np.zeros((1, len(up)))
        coefs[0, 0] = c0
        e2 = et.copy()
        again = orc.select(up, dn, coefs, eps, e2)
        want = sorted(c for c, v in hij.items() if walked and c not in sel1 and eps / c0 <= v < prev)
        assert sorted(zip(again[0].tolist(), again[1].tolist())) == want
        assert e2[0] == (1.000000001 * eps / c0 if walked else prev)
    coefs = np.zeros((1, len(up)))
    coefs[0, 0] = 4.0
    second = orc.select(up, dn, coefs, eps, et)
    want = sorted(c for c, v in hij.items() if c not in sel1 and v >= eps / 4 and
                  (_exc(det, c) == 1 or v < prev))
    assert sorted(zip(second[0].tolist(), second[1].tolist())) == want
    assert et[0] == 1.000000001 * eps / 4 and np.all(et[1:] == 1e100)  # zero coefficients: not walked


def test_selection_max_coefficient_over_states():
    n, nu, nd = 8, 2, 2
    _, h1, eri, core = si.synthetic(n, seed=21)
    orc = oracle.System.from_arrays(n, nu, nd, h1, eri, core)
    det = (0b11, 0b11)
    eps = 2e-3
    c = np.array([[0.1], [-0.5]])
    a = orc.select([det[0]], [det[1]], c, eps, np.full(1, 1e100))
    b = orc.select([det[0]], [det[1]], np.array([[0.5]]), eps, np.full(1, 1e100))
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_selection_time_sym_rules():
    """time_sym (solver.h:447-455) from the plain sweep: one canonical det at
    a time, eps_min times sqrt2 when up != dn, candidates in canonical form,
    minus the space."""
    n, ne = 8, 2
    _, h1, eri, core = si.synthetic(n, seed=23)
    orc = oracle.System.from_arrays(n, ne, ne, h1, eri, core)
    dets = [(0b0011, 0b0011), (0b0101, 0b0110), (0b0011, 0b1001), (0b1001, 0b1100)]
    up = np.array([d[0] for d in dets], dtype=np.uint64)
    dn = np.array([d[1] for d in dets], dtype=np.uint64)
    c = np.array([[0.9, -0.3, 0.05, 0.01]])
    eps = 2e-3
    got = orc.select(up, dn, c, eps, np.full(len(dets), 1e100), time_sym=True)
    got = set(zip(got[0].tolist(), got[1].tolist()))
    want = set()
    for k, (u, d) in enumerate(dets):
        e = eps * (1.4142135623730951 if u != d else 1.0)
        nu, nd = orc.select(up[k:k + 1], dn[k:k + 1], c[:, k:k + 1], e, np.full(1, 1e100))
        for a, b in zip(nu.tolist(), nd.tolist()):
            want.add((a, b) if a <= b else (b, a))
    want -= set(dets)
    assert got == want and len(got) > 0
