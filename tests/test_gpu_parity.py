"""GPU Hamiltonian / SpMV / Davidson / RDM vs the CPU oracle on the same seeded
inputs.  The bar: connection pattern, signs and FP64 values bit-exact (the
kernel follows the reference's operation order); FP64 energies within 1e-9 Ha;
mixed precision within 1 uHa of the FP64 energy."""
import numpy as np
import pytest

import oracle
from shci_amd import integrals as si


def _dets_from_selection(n, n_up, n_dn, h1, eri, eps_list):
    """CPU heat-bath det list with oracle Davidson coefficients."""
    from shci_amd.host import HostSystem

    host = HostSystem.from_arrays(n, n_up, n_dn, h1, eri, 0.0)
    orc = oracle.System.from_arrays(n, n_up, n_dn, h1, eri, 0.0)
    coefs = np.ones((1, 1))
    for eps in eps_list:
        host.select(coefs, eps)
        up, dn = host.dets()
        orc.update(up, dn)
        init = np.full(len(up), 1e-16)
        init[: coefs.shape[1]] = coefs[0]
        ev, evec, conv, traj, _ = orc.davidson(init[None], 1e-8)
        coefs = evec
    return up, dn, orc, coefs


def _canon_upper(rp, cols, vals, n):
    rows = np.repeat(np.arange(n), np.diff(rp))
    key = rows * n + cols
    order = np.argsort(key, kind="stable")
    return key[order], vals[order]


def _full_space(n, n_up, n_dn, h1, eri, orbs=None):
    """Every determinant over the orbitals ``orbs`` (all n by default; the
    lowest ones occupied first) with oracle Davidson coefficients."""
    import itertools

    orbs = list(range(n)) if orbs is None else sorted(orbs)
    ups = [sum(1 << k for k in c) for c in itertools.combinations(orbs, n_up)]
    dns = [sum(1 << k for k in c) for c in itertools.combinations(orbs, n_dn)]
    hf = (sum(1 << k for k in orbs[:n_up]), sum(1 << k for k in orbs[:n_dn]))
    dets = [hf] + [(u, d) for u in ups for d in dns if (u, d) != hf]
    up = np.array([d[0] for d in dets], dtype=np.uint64)
    dn = np.array([d[1] for d in dets], dtype=np.uint64)
    orc = oracle.System.from_arrays(n, n_up, n_dn, h1, eri, 0.0)
    orc.update(up, dn)
    init = np.zeros(len(up))
    init[0] = 1.0
    vec = init[None]
    for _ in range(20):
        _, vec, conv, _, _ = orc.davidson(vec, 1e-10)
        if conv:
            break
    return up, dn, orc, vec


SYSTEMS = [
    # (label, n_orb, n_up, n_dn, seed, eps schedule or None for the full space)
    ("n12_3_3", 12, 3, 3, 7, [1e-3, 1e-4, 1e-5]),
    ("n16_4_3", 16, 4, 3, 11, [1e-3, 3e-4]),
    ("cfg1_n20_5_5", 20, 5, 5, 1001, [5e-4, 2e-4]),
    # one alpha electron (A cores are minus-one strings), two beta electrons
    ("n14_1_2", 14, 1, 2, 13, [1e-3, 1e-5]),
    # full space: mixed cores of up to 81 members (one-wavefront segments)
    ("fci_n11_2_2", 11, 2, 2, 17, None),
    # full space at the largest orbital count (one uint64 per spin, orbital
    # 63 occupied): one mixed core of all 4,096 dets, the largest segment
    # the layout allows
    ("fci_n64_1_1", 64, 1, 1, 19, None),
    # every 3+3 det over the top eight of 64 orbitals (56-63): 3,136 dets,
    # orbital 63 in 21 of the 56 strings of each spin, two-electron A and B
    # cores
    ("sub_n64_3_3", 64, 3, 3, 29, ("orbs", [56, 57, 58, 59, 60, 61, 62, 63])),
]


@pytest.fixture(scope="module", params=SYSTEMS, ids=[s[0] for s in SYSTEMS])
def system(request):
    label, n, nu, nd, seed, eps = request.param
    _, h1, eri, _ = si.synthetic(n, seed=seed)
    if eps is None:
        up, dn, orc, coefs = _full_space(n, nu, nd, h1, eri)
    elif isinstance(eps, tuple):
        up, dn, orc, coefs = _full_space(n, nu, nd, h1, eri, orbs=eps[1])
    else:
        up, dn, orc, coefs = _dets_from_selection(n, nu, nd, h1, eri, eps)
    return dict(label=label, n=n, nu=nu, nd=nd, h1=h1, eri=eri, up=up, dn=dn, orc=orc, coefs=coefs)


@pytest.mark.gpu
def test_pattern_and_values_bit_exact(system):
    from shci_amd.hamiltonian import Hamiltonian

    s = system
    ham = Hamiltonian(s["n"], s["h1"], s["eri"], 0.0, store="both")
    ham.update(s["up"], s["dn"])
    n = len(s["up"])
    print(f"{s['label']}: {n} dets", flush=True)
    g_rp, g_cols, g_vals = ham.matrix.export_upper()
    o_rp, o_cols, o_vals = s["orc"].export_upper()
    gk, gv = _canon_upper(g_rp, g_cols, g_vals, n)
    ok, ov = _canon_upper(o_rp, o_cols, o_vals, n)
    assert len(gk) == len(ok)
    assert np.array_equal(gk, ok)
    assert np.array_equal(gv, ov)  # bit-exact FP64 values, signs included
    np.testing.assert_array_equal(ham.matrix.get_diag(), s["orc"].diag())
    assert ham.matrix.count_n_elems() == s["orc"].stats()["nnz"]


@pytest.mark.gpu
@pytest.mark.parametrize("layout", ["csr", "segments"])
def test_spmv_fp64_fp32(system, layout):
    from shci_amd.hamiltonian import Hamiltonian

    s = system
    ham = Hamiltonian(s["n"], s["h1"], s["eri"], 0.0, store="both", layout=layout)
    ham.update(s["up"], s["dn"])
    rng = np.random.default_rng(7)
    x = rng.uniform(-1, 1, len(s["up"]))
    y_ref = s["orc"].mul(x)
    y64 = ham.matrix.mul(x)
    np.testing.assert_allclose(y64, y_ref, rtol=1e-12, atol=1e-11)
    y32 = ham.matrix.mul(x, fp32=True)
    np.testing.assert_allclose(y32, y_ref, rtol=2e-5, atol=2e-5 * np.abs(y_ref).max())
    # the FP32 block product (k = 3 and 4 vectors per pass) against the
    # oracle's product of each vector
    for k in (3, 4):
        X = np.stack([x, 0.5 * x] + [rng.uniform(-1, 1, len(x)) for _ in range(k - 2)])
        Y = ham.matrix.mul_block(X, fp32=True)
        for v in range(k):
            yo = y_ref if v == 0 else 0.5 * y_ref if v == 1 else s["orc"].mul(X[v])
            assert np.linalg.norm(Y[v] - yo) <= 1e-5 * np.linalg.norm(yo), (k, v)


@pytest.mark.gpu
@pytest.mark.parametrize("layout", ["csr", "segments"])
def test_davidson_fp64_and_mixed(system, layout):
    from shci_amd.hamiltonian import Davidson, Hamiltonian

    s = system
    n = len(s["up"])
    ham = Hamiltonian(s["n"], s["h1"], s["eri"], 0.0, store="both", layout=layout)
    ham.update(s["up"], s["dn"])
    init = np.zeros(n)
    init[0] = 1.0
    ev_o, _, _, traj_o, _ = s["orc"].davidson(init[None], 1e-10)
    d = Davidson(1)
    d.diagonalize(ham.matrix, init, 1e-10, mode="fp64")
    np.testing.assert_allclose(d.trajectory, traj_o, rtol=0, atol=1e-10)
    assert abs(d.lowest_eigenvalues[0] - ev_o[0]) < 1e-9
    # converge tightly (repeated warm restarts) for the reference energy
    d.diagonalize(ham.matrix, init, 1e-11, mode="fp64", max_calls=50)
    e64 = d.lowest_eigenvalues[0]
    dm = Davidson(1)
    dm.diagonalize(ham.matrix, init, 1e-9, mode="mixed", max_calls=50)
    print(f"{s['label']}: E64 {e64:.12f} mixed {dm.lowest_eigenvalues[0]:.12f}", flush=True)
    assert abs(dm.lowest_eigenvalues[0] - e64) <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("layout", ["csr", "segments"])
def test_rdm_vs_oracle(system, layout):
    from shci_amd.hamiltonian import Hamiltonian

    s = system
    ham = Hamiltonian(s["n"], s["h1"], s["eri"], 0.0, store="fp64", layout=layout)
    ham.update(s["up"], s["dn"])
    c = s["coefs"][:1]
    r1g, r2g = ham.rdm(c)
    r1o, r2o = s["orc"].rdm(c)
    np.testing.assert_allclose(r1g, r1o, rtol=0, atol=1e-12)
    np.testing.assert_allclose(r2g, r2o, rtol=0, atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("keep_zeros", [False, True])
def test_segments_counts_and_update_existing(system, keep_zeros):
    """The segment layout stores the same pairs (count) as the CSR and
    recomputes values on its fixed pattern like the CSR does (one context
    holds one matrix: CSR first, then segments)."""
    from shci_amd.hamiltonian import Hamiltonian

    s = system
    _, h1b, erib, _ = si.synthetic(s["n"], seed=999)
    x = np.random.default_rng(3).uniform(-1, 1, len(s["up"]))
    ys = {}
    for layout in ("csr", "segments"):
        ham = Hamiltonian(s["n"], s["h1"], s["eri"], 0.0, store="both", layout=layout, keep_zeros=keep_zeros)
        ham.update(s["up"], s["dn"])
        ys[layout + "_nnz"] = ham.matrix.count_n_elems()
        ham.set_integrals(s["n"], h1b, erib, 0.0)
        ham.update_existing_elems()
        ys[layout] = ham.matrix.mul(x)
        ys[layout + "32"] = ham.matrix.mul(x, fp32=True)
        if layout == "segments":
            st = ham.matrix.layout_stats()
            rg = ham.matrix.segment_ranges()
            # the ranges partition the segments, memberships and slots; a
            # segment of m members keeps m(m-1)/2 slots and sits in its class
            assert sum(r["segments"] for r in rg) == st["segments"]
            assert sum(r["memberships"] for r in rg) == st["memberships"]
            assert sum(r["slots"] for r in rg) == st["value_slots"]
            # the kept pairs of the ranges are the stored pairs (the product
            # and the oracle's count agree on those), each in a slot of its range
            assert 2 * sum(r["kept_pairs"] for r in rg) + len(s["up"]) == ys[layout + "_nnz"]
            assert all(r["kept_pairs"] <= r["slots"] for r in rg)
            lo = [1025, 257, 65, 33, 17, 9, 5, 3, 2]
            hi = [4096, 1024, 256, 64, 32, 16, 8, 4, 2]
            for r in rg:
                c, ns, nm = r["class"], r["segments"], r["memberships"]
                assert lo[c] * ns <= nm <= hi[c] * ns
                assert (nm * (lo[c] - 1)) // 2 <= r["slots"] <= (nm * (hi[c] - 1)) // 2
        ham.clear()
    if not keep_zeros:
        assert ys["csr_nnz"] == s["orc"].stats()["nnz"]
        # the oracle's pattern for the first integrals, values recomputed with the second
        s["orc"].set_integrals(h1b, erib, 0.0)
        s["orc"].update_existing_elems()
        y_orc = s["orc"].mul(x)
        s["orc"].set_integrals(s["h1"], s["eri"], 0.0)
        s["orc"].update_existing_elems()
        np.testing.assert_allclose(ys["csr"], y_orc, rtol=1e-12, atol=1e-11)
    assert ys["segments_nnz"] == ys["csr_nnz"]
    np.testing.assert_allclose(ys["segments"], ys["csr"], rtol=1e-12, atol=1e-11)
    np.testing.assert_allclose(ys["segments32"], ys["csr"], rtol=0, atol=2e-4 * np.abs(ys["csr"]).max())


@pytest.mark.gpu
@pytest.mark.parametrize("n,ne,cls", [(34, 1, 0), (20, 2, 1)])
def test_large_segments_vs_csr(n, ne, cls):
    """Full spaces whose C segments exceed 1024 members (34 orbitals, 1+1
    electrons: one segment of 1156) and 256 members (20 orbitals, 2+2: 361):
    the one-wavefront kernel with each segment's steps split over several
    workgroups.  FP64 and FP32 products and the FP64 Davidson energy against
    the CSR layout."""
    import itertools

    from shci_amd.hamiltonian import Davidson, Hamiltonian

    _, h1, eri, _ = si.synthetic(n, seed=31 + n)
    strs = [sum(1 << k for k in c) for c in itertools.combinations(range(n), ne)]
    hf = (1 << ne) - 1
    dets = [(hf, hf)] + [(u, d) for u in strs for d in strs if (u, d) != (hf, hf)]
    up = np.array([d[0] for d in dets], dtype=np.uint64)
    dn = np.array([d[1] for d in dets], dtype=np.uint64)
    x = np.random.default_rng(9).uniform(-1, 1, len(up))
    init = np.zeros(len(up))
    init[0] = 1.0
    res = {}
    for layout in ("csr", "segments"):
        ham = Hamiltonian(n, h1, eri, 0.0, store="both", layout=layout)
        ham.update(up, dn)
        if layout == "segments":
            k = sum(r["segments"] for r in ham.matrix.segment_ranges() if r["class"] == cls)
            print(f"n={n} {ne}+{ne}: {len(up)} dets, {k} segments of class {cls}", flush=True)
            assert k > 0
        res[layout] = ham.matrix.mul(x)
        res[layout + "32"] = ham.matrix.mul(x, fp32=True)
        d = Davidson(1)
        d.diagonalize(ham.matrix, init, 1e-10, max_calls=40)
        res[layout + "_e"] = d.lowest_eigenvalues[0]
        ham.clear()
    np.testing.assert_allclose(res["segments"], res["csr"], rtol=1e-12, atol=1e-11)
    np.testing.assert_allclose(res["segments32"], res["csr"], rtol=0, atol=2e-4 * np.abs(res["csr"]).max())
    assert abs(res["segments_e"] - res["csr_e"]) < 1e-9


@pytest.mark.gpu
def test_segments_with_split_tiles_vs_csr():
    """Full space of 12 orbitals, 3+3 electrons (48,400 dets): the spin groups
    hold thousands of memberships, so the segment SpMV runs on groups cut into
    several tiles.  FP64 and FP32 products, and the FP64 Davidson energy,
    against the CSR layout (bit-exact against the oracle on the smaller
    systems)."""
    import itertools

    from shci_amd.hamiltonian import Davidson, Hamiltonian

    n, nu, nd = 12, 3, 3
    _, h1, eri, _ = si.synthetic(n, seed=21)
    strs = [sum(1 << k for k in c) for c in itertools.combinations(range(n), nu)]
    hf = (1 << nu) - 1
    dets = [(hf, hf)] + [(u, d) for u in strs for d in strs if (u, d) != (hf, hf)]
    up = np.array([d[0] for d in dets], dtype=np.uint64)
    dn = np.array([d[1] for d in dets], dtype=np.uint64)
    x = np.random.default_rng(5).uniform(-1, 1, len(up))
    init = np.zeros(len(up))
    init[0] = 1.0
    res = {}
    for layout in ("csr", "segments"):
        ham = Hamiltonian(n, h1, eri, 0.0, store="both", layout=layout)
        ham.update(up, dn)
        print(f"{layout}: {len(up)} dets, stats {ham.matrix.layout_stats()}", flush=True)
        res[layout] = ham.matrix.mul(x)
        res[layout + "32"] = ham.matrix.mul(x, fp32=True)
        res[layout + "_nnz"] = ham.matrix.count_n_elems()
        d = Davidson(1)
        d.diagonalize(ham.matrix, init, 1e-10, max_calls=40)
        res[layout + "_e"] = d.lowest_eigenvalues[0]
        if layout == "segments":
            st = ham.matrix.layout_stats()
            groups = st["alpha_groups"] + st["beta_groups"]
            assert st["tiles"] > groups  # some groups span several tiles
        ham.clear()
    assert res["segments_nnz"] == res["csr_nnz"]
    np.testing.assert_allclose(res["segments"], res["csr"], rtol=1e-12, atol=1e-11)
    np.testing.assert_allclose(res["segments32"], res["csr"], rtol=0, atol=2e-4 * np.abs(res["csr"]).max())
    assert abs(res["segments_e"] - res["csr_e"]) < 1e-9


_LOWEST3 = {}


def _lowest_three(s):
    """The three lowest eigenvalues of the oracle's matrix: numpy eigh of the
    dense matrix up to 4,000 dets, scipy's Lanczos (eigsh, smallest
    algebraic, tol 1e-13) on the sparse symmetric matrix above."""
    if s["label"] not in _LOWEST3:
        import scipy.sparse as sp
        import scipy.sparse.linalg as sla

        n = len(s["up"])
        rp, cols, vals = s["orc"].export_upper()
        rows = np.repeat(np.arange(n), np.diff(rp))
        if n <= 4000:
            H = np.zeros((n, n))
            H[rows, cols] = vals
            H = H + H.T - np.diag(np.diag(H))
            w = np.linalg.eigvalsh(H)[:3]
        else:
            U = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
            H = (U + U.T - sp.diags(U.diagonal())).tocsr()
            w = np.sort(sla.eigsh(H, k=3, which="SA", tol=1e-13, v0=np.ones(n), return_eigenvectors=False))
        _LOWEST3[s["label"]] = w
    return _LOWEST3[s["label"]]


@pytest.mark.gpu
@pytest.mark.parametrize("layout", ["csr", "segments"])
def test_multiroot_davidson_vs_exact_roots(system, layout):
    """Three roots (the multi-root path of cfg 5): GPU FP64 and mixed Davidson
    against the oracle matrix's three lowest eigenvalues (dense eigh, or
    Lanczos above 4,000 dets), and the state-summed RDM against the oracle."""
    from shci_amd.hamiltonian import Davidson, Hamiltonian

    s = system
    n = len(s["up"])
    w = _lowest_three(s)
    ham = Hamiltonian(s["n"], s["h1"], s["eri"], 0.0, store="both", layout=layout)
    ham.update(s["up"], s["dn"])
    init = np.zeros((3, n))
    for k in range(3):
        init[k, k] = 1.0
    d = Davidson(3)
    d.diagonalize(ham.matrix, init, 1e-12, mode="fp64", max_calls=100)
    print(f"{s['label']}: dense {w} fp64 {d.lowest_eigenvalues}", flush=True)
    np.testing.assert_allclose(d.lowest_eigenvalues, w, rtol=0, atol=1e-9)
    dm = Davidson(3)
    dm.diagonalize(ham.matrix, init, 1e-10, mode="mixed", max_calls=100)
    np.testing.assert_allclose(dm.lowest_eigenvalues, w, rtol=0, atol=1e-6)
    c = d.get_lowest_eigenvectors()
    r1g, r2g = ham.rdm(c)
    r1o, r2o = s["orc"].rdm(c)
    np.testing.assert_allclose(r1g, r1o, rtol=0, atol=1e-11)
    np.testing.assert_allclose(r2g, r2o, rtol=0, atol=1e-11)


@pytest.mark.gpu
@pytest.mark.parametrize("layout", ["csr", "segments"])
def test_mixed_multiroot_degenerate_pair(layout):
    """Every 3+3 det over orbitals 0, 1, 30 and 59-63 of 64 (3,136 dets): the
    second and third roots are degenerate to 2.5e-10, the states of two
    5-det clusters with the same spectrum that couple to each other at
    1.8e-9 Ha and to the rest at most 8.2e-5 Ha.  The start vectors (dets
    0-2) lie outside the partner's cluster.  With FP32 products the
    eigenvalue-change test (davidson.cc:174-186) is met by stagnation with
    the second root ~5e-3 Ha high (a numpy restatement of the reference
    Davidson with FP32 products stops there too); the mixed mode's residual
    guard restarts from the Ritz vectors and reaches all three roots."""
    from shci_amd.hamiltonian import Davidson, Hamiltonian

    n_orb, orbs = 64, [0, 1, 30, 59, 60, 61, 62, 63]
    _, h1, eri, _ = si.synthetic(n_orb, seed=29)
    up, dn, orc, _ = _full_space(n_orb, 3, 3, h1, eri, orbs=orbs)
    n = len(up)
    w = _lowest_three(dict(label="degenerate_n64_3_3", up=up, dn=dn, orc=orc))
    assert w[2] - w[1] < 1e-9  # the pair is degenerate
    ham = Hamiltonian(n_orb, h1, eri, 0.0, store="both", layout=layout)
    ham.update(up, dn)
    print(f"degenerate pair {layout}: {n} dets, roots {w}", flush=True)
    init = np.zeros((3, n))
    for k in range(3):
        init[k, k] = 1.0
    d = Davidson(3)
    d.diagonalize(ham.matrix, init, 1e-12, mode="fp64", max_calls=100)
    np.testing.assert_allclose(d.lowest_eigenvalues, w, rtol=0, atol=1e-9)
    dm = Davidson(3)
    dm.diagonalize(ham.matrix, init, 1e-10, mode="mixed", max_calls=100)
    print(f"  mixed {dm.lowest_eigenvalues} converged {dm.converged}", flush=True)
    assert dm.converged
    np.testing.assert_allclose(dm.lowest_eigenvalues, w, rtol=0, atol=1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("layout", ["csr", "segments"])
def test_ritz_capture_fp64_energies(system, layout):
    """The captured Ritz vectors (shci_capture_ritz) and their FP64 energies
    (shci_ritz_energies): in FP64 mode each trajectory entry is a Rayleigh
    quotient of the FP64 matrix, so the captured vectors' FP64 energies
    reproduce the trajectory; in mixed mode they are variational (never below
    the dense ground state) and the last lies within 1 uHa of it."""
    from shci_amd.hamiltonian import Davidson, Hamiltonian

    s = system
    n = len(s["up"])
    ham = Hamiltonian(s["n"], s["h1"], s["eri"], 0.0, store="both", layout=layout)
    ham.update(s["up"], s["dn"])
    init = np.zeros(n)
    init[0] = 1.0
    e0, vec0, _, _, _ = s["orc"].davidson(init[None], 1e-12)
    for _ in range(30):
        e1, vec0, conv, _, _ = s["orc"].davidson(vec0, 1e-12)
        if abs(e1[0] - e0[0]) < 1e-13:
            break
        e0 = e1
    cap = 2048
    ham.matrix.capture_ritz(cap)
    d = Davidson(1)
    d.diagonalize(ham.matrix, init, 1e-10, mode="fp64", max_calls=100)
    e = ham.matrix.ritz_energies()
    assert d.converged and len(d.trajectory) <= cap
    assert len(e) == len(d.trajectory)
    np.testing.assert_allclose(e, d.trajectory, rtol=0, atol=1e-10)
    dm = Davidson(1)
    dm.diagonalize(ham.matrix, init, 1e-10, mode="mixed", max_calls=100)
    em = ham.matrix.ritz_energies()
    ham.matrix.capture_ritz(0)
    print(f"{s['label']} {layout}: {len(e)} FP64 and {len(em)} mixed Ritz vectors; last mixed "
          f"{em[-1]:.12f} vs the oracle's {e1[0]:.12f}", flush=True)
    assert len(em) == len(dm.trajectory) - 1  # the final re-evaluation is not a Ritz step
    assert np.all(em >= e1[0] - 1e-9)
    assert abs(em[-1] - e1[0]) <= 1e-6
    np.testing.assert_allclose(em, dm.trajectory[: len(em)], rtol=0, atol=2e-5 * max(1.0, abs(e1[0])))
