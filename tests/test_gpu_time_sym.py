"""GPU time-reversal symmetry (SHCI_TIME_SYM) against the oracle's H_ts
(base_system.h:86-101, pinned against P^T H P in test_oracle_time_sym.py) on
selected canonical det lists: products (1e-12), the diagonal (1e-12), the
lowest eigenvalue (FP64 1e-9, mixed 1 uHa), the RDMs of the unpacked
wavefunction (base_system.h:108-122) against the oracle's (1e-12), append-only
updates, and the canonical-form check."""
import numpy as np
import pytest

import oracle
from shci_amd import integrals as si


def _ts_dets(n, ne, h1, eri, eps_list):
    """Host selection (or, for ("orbs", list), every det over those orbitals,
    the lowest ones occupied first), then the canonical form (up <= dn),
    first occurrence kept."""
    if isinstance(eps_list, tuple):
        import itertools

        orbs = sorted(eps_list[1])
        strs = [sum(1 << k for k in c) for c in itertools.combinations(orbs, ne)]
        hf = sum(1 << k for k in orbs[:ne])
        dets = [(hf, hf)] + [(u, d) for u in strs for d in strs]
        seen, out = set(), []
        for u, d in dets:
            c = (u, d) if u <= d else (d, u)
            if c not in seen:
                seen.add(c)
                out.append(c)
        return np.array([c[0] for c in out], dtype=np.uint64), np.array([c[1] for c in out], dtype=np.uint64)
    from shci_amd.host import HostSystem

    host = HostSystem.from_arrays(n, ne, ne, h1, eri, 0.0)
    coefs = np.ones((1, 1))
    for eps in eps_list:
        host.select(coefs, eps)
        m = host.n_dets
        coefs = np.concatenate([[1.0], np.full(m - 1, 1e-3)])[None]
    up, dn = host.dets()
    seen, out = set(), []
    for u, d in zip(up.tolist(), dn.tolist()):
        c = (u, d) if u <= d else (d, u)
        if c not in seen:
            seen.add(c)
            out.append(c)
    return np.array([c[0] for c in out], dtype=np.uint64), np.array([c[1] for c in out], dtype=np.uint64)


def _dense_ts(orc, up, dn):
    m = len(up)
    H = np.zeros((m, m))
    for i in range(m):
        for j in range(m):
            H[i, j] = orc.elem_time_sym(int(up[i]), int(dn[i]), int(up[j]), int(dn[j]))
    return H


@pytest.mark.gpu
@pytest.mark.parametrize("layout", ["csr", "segments"])
@pytest.mark.parametrize("n,ne,seed,eps", [(12, 3, 7, [1e-2, 3e-3]), (14, 4, 11, [2e-2, 8e-3]),
                                            # every 3+3 det over the top seven of 64 orbitals
                                            (64, 3, 29, ("orbs", [57, 58, 59, 60, 61, 62, 63]))])
def test_gpu_time_sym_vs_oracle(layout, n, ne, seed, eps):
    from shci_amd.hamiltonian import Davidson, Hamiltonian

    _, h1, eri, core = si.synthetic(n, seed=seed)
    up, dn = _ts_dets(n, ne, h1, eri, eps)
    m = len(up)
    orc = oracle.System.from_arrays(n, ne, ne, h1, eri, core)
    H = _dense_ts(orc, up, dn)
    print(f"[ts] n={n} {ne}+{ne} {layout}: {m} symmetrised dets, {int(np.sum(up != dn))} pairs", flush=True)
    ham = Hamiltonian(n, h1, eri, core, store="both", layout=layout, time_sym=True)
    ham.update(up, dn)
    assert ham.matrix.count_n_rows() == m
    # the stored matrix is over the dets and their reverses: its entry
    # counts are not H_ts's, so they are refused rather than misreported
    from shci_amd.hamiltonian import ShciError

    with pytest.raises(ShciError, match="time_sym"):
        ham.matrix.count_n_elems()
    np.testing.assert_allclose(ham.matrix.get_diag(), np.diag(H), rtol=0, atol=1e-12)
    x = np.random.default_rng(seed).uniform(-1, 1, m)
    np.testing.assert_allclose(ham.matrix.mul(x), H @ x, rtol=0, atol=1e-11)
    w = np.linalg.eigvalsh(H)
    init = np.zeros(m)
    init[0] = 1.0
    d = Davidson(1)
    d.diagonalize(ham.matrix, init, 1e-11, max_calls=40)
    assert abs(d.lowest_eigenvalues[0] - w[0]) < 1e-9, (d.lowest_eigenvalues[0], w[0])
    c = np.atleast_2d(d.get_lowest_eigenvectors())[0]
    # the reduced eigenvector: unit norm, Rayleigh quotient at the eigenvalue
    # (second order in the vector error), small residual (first order; the
    # Davidson stops on the eigenvalue change)
    assert abs(c @ c - 1.0) < 1e-8
    assert abs(c @ H @ c - w[0]) < 1e-9
    assert np.linalg.norm(H @ c - w[0] * c) < 1e-3
    dm = Davidson(1)
    dm.diagonalize(ham.matrix, init, 1e-10, mode="mixed", max_calls=40)
    assert abs(dm.lowest_eigenvalues[0] - w[0]) < 1e-6
    # three roots: the mixed mode's block products, each projected
    init3 = np.zeros((3, m))
    for k in range(3):
        init3[k, k] = 1.0
    d3 = Davidson(3)
    d3.diagonalize(ham.matrix, init3, 1e-10, mode="mixed", max_calls=100)
    np.testing.assert_allclose(d3.lowest_eigenvalues, w[:3], rtol=0, atol=1e-6)
    # RDMs of the unpacked wavefunction (each pair's coefficient / sqrt2 on
    # both dets, the reference's SQRT2_INV)
    r1, r2 = ham.rdm(c[None])
    eu, ed, ec = [], [], []
    for u, dd, ci in zip(up.tolist(), dn.tolist(), c.tolist()):
        if u == dd:
            eu.append(u); ed.append(dd); ec.append(ci)
        else:
            eu += [u, dd]; ed += [dd, u]; ec += [ci * 0.7071067811865475] * 2
    orc.update(np.array(eu, dtype=np.uint64), np.array(ed, dtype=np.uint64))
    o1, o2 = orc.rdm(np.array(ec)[None])
    np.testing.assert_allclose(r1, o1, rtol=0, atol=1e-12)
    np.testing.assert_allclose(r2, o2, rtol=0, atol=1e-12)


def _b(*orbs):
    return sum(1 << k for k in orbs)


@pytest.mark.gpu
@pytest.mark.parametrize("layout", ["csr", "segments"])
@pytest.mark.parametrize("label,ups,dns", [
    # one det that is its own reverse: H_ts = H
    ("self", [_b(0, 1, 2)], [_b(0, 1, 2)]),
    # one det and its reverse: a one-dimensional H_ts from a two-det list
    ("pair", [_b(0, 1, 2)], [_b(0, 1, 5)]),
    # the two together: the self det couples to the pair state with a sqrt2
    ("self_and_pair", [_b(0, 1, 2), _b(0, 1, 2)], [_b(0, 1, 2), _b(0, 1, 5)]),
])
def test_gpu_time_sym_tiny_lists(layout, label, ups, dns):
    """The smallest symmetrised lists against the oracle's dense H_ts: the
    diagonal, the product, FP64 and mixed eigenvalues, and the RDMs of the
    unpacked wavefunction."""
    from shci_amd.hamiltonian import Davidson, Hamiltonian

    n, ne = 12, 3
    _, h1, eri, core = si.synthetic(n, seed=7)
    up, dn = np.array(ups, dtype=np.uint64), np.array(dns, dtype=np.uint64)
    m = len(up)
    orc = oracle.System.from_arrays(n, ne, ne, h1, eri, core)
    H = _dense_ts(orc, up, dn)
    print(f"[ts tiny] {label} {layout}: {m} symmetrised dets", flush=True)
    ham = Hamiltonian(n, h1, eri, core, store="both", layout=layout, time_sym=True)
    ham.update(up, dn)
    assert ham.matrix.count_n_rows() == m
    np.testing.assert_allclose(ham.matrix.get_diag(), np.diag(H), rtol=0, atol=1e-12)
    x = np.linspace(0.5, -1.0, m) + 0.1
    np.testing.assert_allclose(ham.matrix.mul(x), H @ x, rtol=0, atol=1e-11)
    w = np.linalg.eigvalsh(H)
    init = np.zeros(m)
    init[0] = 1.0
    d = Davidson(1)
    d.diagonalize(ham.matrix, init, 1e-11, max_calls=40)
    assert abs(d.lowest_eigenvalues[0] - w[0]) < 1e-9, (d.lowest_eigenvalues[0], w[0])
    dm = Davidson(1)
    dm.diagonalize(ham.matrix, init, 1e-10, mode="mixed", max_calls=40)
    assert abs(dm.lowest_eigenvalues[0] - w[0]) < 1e-6
    c = np.atleast_2d(d.get_lowest_eigenvectors())[0]
    assert abs(c @ c - 1.0) < 1e-8
    r1, r2 = ham.rdm(c[None])
    eu, ed, ec = [], [], []
    for u, dd, ci in zip(up.tolist(), dn.tolist(), c.tolist()):
        if u == dd:
            eu.append(u); ed.append(dd); ec.append(ci)
        else:
            eu += [u, dd]; ed += [dd, u]; ec += [ci * 0.7071067811865475] * 2
    orc.update(np.array(eu, dtype=np.uint64), np.array(ed, dtype=np.uint64))
    o1, o2 = orc.rdm(np.array(ec)[None])
    np.testing.assert_allclose(r1, o1, rtol=0, atol=1e-12)
    np.testing.assert_allclose(r2, o2, rtol=0, atol=1e-12)


@pytest.mark.gpu
def test_gpu_time_sym_append_and_checks():
    from shci_amd.hamiltonian import Hamiltonian, ShciError

    n, ne = 12, 3
    _, h1, eri, core = si.synthetic(n, seed=7)
    up, dn = _ts_dets(n, ne, h1, eri, [1e-2, 3e-3])
    m = len(up)
    ham = Hamiltonian(n, h1, eri, core, store="fp64", layout="segments", time_sym=True)
    ham.update(up[: m // 2], dn[: m // 2])
    ham.update(up, dn)  # append-only growth (hamiltonian.h:87-112)
    fresh = Hamiltonian(n, h1, eri, core, store="fp64", layout="csr", time_sym=True)
    x = np.random.default_rng(1).uniform(-1, 1, m)
    y1 = ham.matrix.mul(x)
    fresh.update(up, dn)
    np.testing.assert_allclose(y1, fresh.matrix.mul(x), rtol=0, atol=1e-11)
    bad = Hamiltonian(n, h1, eri, core, store="fp64", time_sym=True)
    with pytest.raises(ShciError):
        bad.update(np.array([0b1110], dtype=np.uint64), np.array([0b0111], dtype=np.uint64))  # up > dn


@pytest.mark.gpu
@pytest.mark.parametrize("n,ne,seed,dets", [(14, 4, 11, [2e-2, 5e-3]),
                                             # every canonical 2+2 det over the top six of 64 orbitals
                                             (64, 2, 9, ("orbs", [58, 59, 60, 61, 62, 63]))])
def test_gpu_select_time_sym_vs_oracle(n, ne, seed, dets):
    from shci_amd.hamiltonian import Hamiltonian
    from shci_amd.selection import Selector

    _, h1, eri, core = si.synthetic(n, seed=seed)
    up, dn = _ts_dets(n, ne, h1, eri, dets)
    rng = np.random.default_rng(5)
    c = rng.choice([-1.0, 1.0], (2, len(up))) * 10.0 ** rng.uniform(-5, 0, (2, len(up)))
    c[0, 0] = 0.95
    orc = oracle.System.from_arrays(n, ne, ne, h1, eri, core)
    sel = Selector(Hamiltonian(n, h1, eri, core, time_sym=True))
    eps_tried = np.full(len(up), 1e100)
    for eps in (3e-3, 1e-3, 1e-3):
        eg, eo = eps_tried.copy(), eps_tried.copy()
        gu, gd = sel.select(up, dn, c, eps, eg)
        ou, od = orc.select(up, dn, c, eps, eo, time_sym=True)
        print(f"[ts] selection eps {eps}: {len(gu)} new", flush=True)
        assert np.array_equal(gu, ou) and np.array_equal(gd, od) and np.array_equal(eg, eo)
        assert np.all(gu <= gd)
        eps_tried = eg
        c = c * 2.0


@pytest.mark.gpu
def test_gpu_time_sym_variational_loop():
    """The variational loop with time_sym throughout on the device
    (selection rules, symmetrised build, Davidson): the final energy is the
    lowest eigenvalue of the oracle's dense H_ts over the final dets."""
    from helpers import variation
    from shci_amd.hamiltonian import Davidson, Hamiltonian
    from shci_amd.selection import GpuSystem

    n, ne = 12, 3
    _, h1, eri, core = si.synthetic(n, seed=7)
    ham = Hamiltonian(n, h1, eri, core, store="fp64", layout="segments", time_sym=True)
    hf = (1 << ne) - 1
    gsys = GpuSystem(ham, hf, hf)
    dav = Davidson(1)
    last = {}

    def backend(up, dn, init, tol):
        ham.update(up, dn)
        ev = dav.diagonalize(ham.matrix, init, tol)
        last.update(up=up.copy(), dn=dn.copy(), e=float(ev[0]))
        return ev, dav.get_lowest_eigenvectors(), dav.converged, {}

    recs = list(variation(gsys, backend, [2e-3], [1e-2, 5e-3], target_error=1e-6))
    up, dn = last["up"], last["dn"]
    assert np.all(up <= dn) and len(set(zip(up.tolist(), dn.tolist()))) == len(up)
    orc = oracle.System.from_arrays(n, ne, ne, h1, eri, core)
    H = _dense_ts(orc, up, dn)
    w0 = np.linalg.eigvalsh(H)[0]
    print(f"[ts] loop: {len(recs)} iterations, {len(up)} dets, E {last['e']:.10f} (dense {w0:.10f})", flush=True)
    assert abs(last["e"] - w0) < 1e-9
