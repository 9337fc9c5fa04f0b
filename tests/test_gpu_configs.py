"""The HIP path on the workloads of every benchmark configuration, against the
oracle.  Each test grows its configuration's det list with that config's own
integral generator (``si.synthetic(cfg)``, seed 1000 + cfg) and its selection
schedule (``workloads.grow_det_list``), then compares at a size the oracle
finishes in seconds (a prefix of the grown list is itself an append-only
det list):

* cfg 1 (n_orb 20, 5+5): the full schedule to eps 1e-4;
* cfg 2 (n_orb 40, 8+8): >= 1e5 dets; CSR pattern, signs and FP64 values
  bit-exact (src/solver/hamiltonian.h:415-598), FP64 energy 1e-9;
* cfg 3 (n_orb 60, 10+10): >= 5e4 dets; FP64 energy 1e-9, mixed 1 uHa, both
  layouts;
* cfg 4 (n_orb 30, 7+7): the full ~1.4e6-det space: the RDM-adjoint
  gradient against 20 central finite differences, and the GPU RDMs against
  the oracle's on a prefix;
* cfg 5 (n_orb 50, 9+9, 4 roots): 4-root FP64 and mixed eigenvalues, the
  k = 4 block product vector by vector and the state-summed RDM.

Bars: bit-exact for the pattern, signs and FP64 values; FP64 energies 1e-9 Ha;
mixed precision 1 uHa; FP64 products 1e-12 relative to the vector norm; FP32
products 1e-5 (FP32 values and accumulation)."""
import numpy as np
import pytest

import oracle

from test_gpu_parity import _canon_upper


def _grown(cfg, prefix=None, target=None):
    """The config's grown list (to ``target`` dets, default the config's own
    size) and its first ``prefix`` dets (default all)."""
    from shci_amd import workloads

    w = workloads.grow_det_list(cfg, target=target, verbose=True)
    prefix = prefix or len(w.up)
    m = min(prefix, len(w.up))
    print(f"[cfg{cfg}] grown list {len(w.up):,} dets (digest {w.det_hash}); testing the first {m:,}", flush=True)
    return w, w.up[:m].copy(), w.dn[:m].copy()


def _oracle(w, up, dn):
    spec = w.spec
    orc = oracle.System.from_arrays(w.n, spec.n_up, spec.n_dn, w.h1, w.eri, w.core)
    orc.update(up, dn)
    return orc


def _oracle_eigen(orc, n_dets, n_states, tol=1e-12, calls=60):
    """The oracle's Davidson from unit vectors on the first n_states dets,
    warm-restarted until every eigenvalue changes by less than 1e-12."""
    init = np.zeros((n_states, n_dets))
    for s in range(n_states):
        init[s, s] = 1.0
    ev_prev, vec = None, init
    for _ in range(calls):
        ev, vec, conv, _, _ = orc.davidson(vec, tol)
        if ev_prev is not None and np.max(np.abs(ev - ev_prev)) < 1e-12:
            break
        ev_prev = ev
    return ev, vec, init


def _gpu_eigen(ham, init, mode, tol=1e-12, calls=100):
    from shci_amd.hamiltonian import Davidson

    d = Davidson(init.shape[0])
    d.diagonalize(ham.matrix, init, tol, mode=mode, max_calls=calls)
    return np.asarray(d.lowest_eigenvalues, dtype=float), np.atleast_2d(d.get_lowest_eigenvectors())


def _assert_csr_bit_exact(ham, orc, n):
    g = ham.matrix.export_upper()
    o = orc.export_upper()
    np.testing.assert_array_equal(g[0], o[0])  # stored entries per row
    if len(g[1]) <= 30_000_000:
        gk, gv = _canon_upper(*g, n)
        ok, ov = _canon_upper(*o, n)
        assert np.array_equal(gk, ok)
        assert np.array_equal(gv, ov)  # FP64 values and signs, bit for bit
    else:
        assert oracle.upper_digest(*g) == oracle.upper_digest(*o)
    np.testing.assert_array_equal(ham.matrix.get_diag(), orc.diag())
    assert ham.matrix.count_n_elems() == orc.stats()["nnz"]


def _assert_products(ham, orc, n, seed):
    x = np.random.default_rng(seed).uniform(-1, 1, n)
    yo = orc.mul(x)
    nrm = np.linalg.norm(yo)
    assert np.linalg.norm(ham.matrix.mul(x) - yo) <= 1e-12 * nrm
    assert np.linalg.norm(ham.matrix.mul(x, fp32=True) - yo) <= 1e-5 * nrm


@pytest.mark.gpu
def test_cfg1_full_schedule():
    """cfg 1 at its whole eps schedule (5e-4, 2e-4, 1e-4): pattern and values
    bit-exact, FP64 energy 1e-9 and mixed 1 uHa against the oracle."""
    from shci_amd.hamiltonian import Hamiltonian

    w, up, dn = _grown(1)
    n = len(up)
    assert w.log[-1]["eps"] == 1e-4 and n > 30_000
    orc = _oracle(w, up, dn)
    ev_o, _, init = _oracle_eigen(orc, n, 1)
    ham = Hamiltonian(w.n, w.h1, w.eri, w.core, store="both", layout="csr")
    ham.update(up, dn)
    _assert_csr_bit_exact(ham, orc, n)
    _assert_products(ham, orc, n, 1)
    e64, _ = _gpu_eigen(ham, init, "fp64")
    em, _ = _gpu_eigen(ham, init, "mixed", tol=1e-9)
    print(f"[cfg1] {n} dets: oracle {ev_o[0]:.12f} fp64 {e64[0]:.12f} mixed {em[0]:.12f}", flush=True)
    assert abs(e64[0] - ev_o[0]) <= 1e-9
    assert abs(em[0] - ev_o[0]) <= 1e-6


@pytest.mark.gpu
def test_cfg2_prefix_bit_exact():
    """cfg 2 on the first 120,000 dets of its grown list: the FP64 drop-in
    (CSR pattern, signs and values bit-exact; FP64 energy 1e-9) and the
    segment layout's products and energy."""
    from shci_amd.hamiltonian import Hamiltonian

    w, up, dn = _grown(2, 120_000, target=120_000)
    n = len(up)
    assert n >= 100_000
    orc = _oracle(w, up, dn)
    ev_o, _, init = _oracle_eigen(orc, n, 1)
    ham = Hamiltonian(w.n, w.h1, w.eri, w.core, store="fp64", layout="csr")
    ham.update(up, dn)
    _assert_csr_bit_exact(ham, orc, n)
    e64, _ = _gpu_eigen(ham, init, "fp64")
    ham.clear()
    seg = Hamiltonian(w.n, w.h1, w.eri, w.core, store="both", layout="segments")
    seg.update(up, dn)
    assert seg.matrix.count_n_elems() == orc.stats()["nnz"]
    _assert_products(seg, orc, n, 2)
    e64s, _ = _gpu_eigen(seg, init, "fp64")
    ems, _ = _gpu_eigen(seg, init, "mixed", tol=1e-9)
    seg.clear()
    print(f"[cfg2] {n} dets: oracle {ev_o[0]:.12f} fp64 csr {e64[0]:.12f} segments {e64s[0]:.12f} "
          f"mixed {ems[0]:.12f}", flush=True)
    assert abs(e64[0] - ev_o[0]) <= 1e-9
    assert abs(e64s[0] - ev_o[0]) <= 1e-9
    assert abs(ems[0] - ev_o[0]) <= 1e-6


@pytest.fixture(scope="module")
def cfg3():
    w, up, dn = _grown(3, 60_000, target=60_000)
    orc = _oracle(w, up, dn)
    ev_o, _, init = _oracle_eigen(orc, len(up), 1)
    return w, up, dn, orc, ev_o, init


@pytest.mark.gpu
@pytest.mark.parametrize("layout", ["csr", "segments"])
def test_cfg3_prefix(cfg3, layout):
    """cfg 3 (n_orb 60, 10+10) on the first 60,000 dets of its grown list:
    FP64 energy 1e-9 and mixed 1 uHa against the oracle; products; the CSR
    pattern and values bit-exact."""
    from shci_amd.hamiltonian import Hamiltonian

    w, up, dn, orc, ev_o, init = cfg3
    n = len(up)
    assert n >= 50_000
    ham = Hamiltonian(w.n, w.h1, w.eri, w.core, store="both", layout=layout)
    ham.update(up, dn)
    if layout == "csr":
        _assert_csr_bit_exact(ham, orc, n)
    else:
        assert ham.matrix.count_n_elems() == orc.stats()["nnz"]
    _assert_products(ham, orc, n, 3)
    e64, _ = _gpu_eigen(ham, init, "fp64")
    em, _ = _gpu_eigen(ham, init, "mixed", tol=1e-9)
    ham.clear()
    print(f"[cfg3 {layout}] {n} dets: oracle {ev_o[0]:.12f} fp64 {e64[0]:.12f} mixed {em[0]:.12f} "
          f"({abs(em[0] - ev_o[0]) * 1e6:.4f} uHa)", flush=True)
    assert abs(e64[0] - ev_o[0]) <= 1e-9
    assert abs(em[0] - ev_o[0]) <= 1e-6


@pytest.mark.gpu
def test_cfg4_full_size_gradient():
    """cfg 4 (n_orb 30, 7+7) at its full grown size: the E(h, eri) gradient
    of the RDM adjoint against central finite differences along 20 seeded
    symmetric unit directions (step 1e-5, 1e-6 relative), and the GPU RDMs of
    the converged vector against the oracle's on a 40,000-det prefix."""
    from shci_amd.autograd import VariationalSolver
    from shci_amd.hamiltonian import Hamiltonian
    from test_gpu_autograd import _fd_check

    w, up, dn = _grown(4)
    assert len(up) > 1_000_000
    solver = VariationalSolver(w.n, up, dn, w.core, layout="segments")
    worst = _fd_check(solver, w.h1, w.eri, n_dirs=20, seed=2024)
    print(f"[cfg4] {len(up):,} dets: worst FD relative error {worst:.2e}", flush=True)
    solver.ham.clear()
    m = 40_000
    orc = _oracle(w, up[:m], dn[:m])
    _, vec, _ = _oracle_eigen(orc, m, 1)
    ham = Hamiltonian(w.n, w.h1, w.eri, w.core, store="fp64", layout="segments")
    ham.update(up[:m], dn[:m])
    r1g, r2g = ham.rdm(vec)
    r1o, r2o = orc.rdm(vec)
    ham.clear()
    np.testing.assert_allclose(r1g, r1o, rtol=0, atol=1e-11)
    np.testing.assert_allclose(r2g, r2o, rtol=0, atol=1e-11)


@pytest.mark.gpu
def test_cfg5_four_roots():
    """cfg 5 (n_orb 50, 9+9, 4 roots) on the first 40,000 dets of its grown
    list: the four FP64 eigenvalues (1e-9) and mixed eigenvalues (1 uHa)
    against the oracle's Davidson; the k = 4 FP32 block product against the
    oracle's product of each vector; the state-summed RDM of the four roots."""
    from shci_amd.hamiltonian import Hamiltonian

    w, up, dn = _grown(5, 40_000, target=40_000)
    n = len(up)
    assert w.spec.n_states == 4 and n >= 30_000
    orc = _oracle(w, up, dn)
    ev_o, vec_o, init = _oracle_eigen(orc, n, 4)
    ham = Hamiltonian(w.n, w.h1, w.eri, w.core, store="both", layout="segments")
    ham.update(up, dn)
    assert ham.matrix.count_n_elems() == orc.stats()["nnz"]
    X = np.random.default_rng(5).uniform(-1, 1, (4, n))
    Y = ham.matrix.mul_block(X, fp32=True)
    for k in range(4):
        yo = orc.mul(X[k])
        assert np.linalg.norm(Y[k] - yo) <= 1e-5 * np.linalg.norm(yo), k
    e64, c64 = _gpu_eigen(ham, init, "fp64")
    em, _ = _gpu_eigen(ham, init, "mixed", tol=1e-9)
    print(f"[cfg5] {n} dets: oracle {ev_o} fp64 {e64} mixed {em}", flush=True)
    np.testing.assert_allclose(e64, ev_o, rtol=0, atol=1e-9)
    np.testing.assert_allclose(em, ev_o, rtol=0, atol=1e-6)
    r1g, r2g = ham.rdm(c64)
    r1o, r2o = orc.rdm(c64)
    ham.clear()
    np.testing.assert_allclose(r1g, r1o, rtol=0, atol=1e-11)
    np.testing.assert_allclose(r2g, r2o, rtol=0, atol=1e-11)
    assert abs(np.trace(r1g) - 4 * 18) < 1e-8  # 4 normalised states of 18 electrons
