"""The smallest inputs at the boundary, GPU against the oracle: one det (the
reference Davidson's one-dimensional branch, davidson.cc:20-26), dets that
no pair of which connects (a diagonal H: no segment is kept, and the
Davidson correction vanishes, davidson.cc:106-128), and two dets one alpha
single apart (the pair is stored once and is a zero slot in the other
2-member segments it shares; src/solver/hamiltonian.h:415-598 for the
pattern).  An empty list is refused.  Same bar as the larger parity
systems: pattern and FP64 values bit-exact, FP64 energies within 1e-9 Ha,
mixed within 1 uHa."""
import numpy as np
import pytest

import oracle
from shci_amd import integrals as si


def _bits(*orbs):
    return sum(1 << k for k in orbs)


CASES = [
    # (label, n_up, n_dn, alpha strings, beta strings)
    ("one_det", 3, 3, [_bits(0, 1, 2)], [_bits(0, 1, 2)]),
    # every pair four excitations apart
    ("disconnected", 2, 2, [_bits(0, 1), _bits(2, 3), _bits(4, 5), _bits(6, 7)],
     [_bits(0, 1), _bits(2, 3), _bits(4, 5), _bits(6, 7)]),
    # alpha 2 -> 5: three mixed and two alpha-alpha cores hold the pair
    ("one_pair", 3, 3, [_bits(0, 1, 2), _bits(0, 1, 5)], [_bits(0, 1, 2), _bits(0, 1, 2)]),
]
N_ORB = 12


@pytest.fixture(scope="module", params=CASES, ids=[c[0] for c in CASES])
def case(request):
    label, nu, nd, ups, dns = request.param
    _, h1, eri, _ = si.synthetic(N_ORB, seed=5)
    up = np.array(ups, dtype=np.uint64)
    dn = np.array(dns, dtype=np.uint64)
    orc = oracle.System.from_arrays(N_ORB, nu, nd, h1, eri, 0.0)
    orc.update(up, dn)
    return dict(label=label, h1=h1, eri=eri, up=up, dn=dn, orc=orc)


def _ham(c, layout, store="both"):
    from shci_amd.hamiltonian import Hamiltonian

    ham = Hamiltonian(N_ORB, c["h1"], c["eri"], 0.0, store=store, layout=layout)
    ham.update(c["up"], c["dn"])
    print(f"{c['label']} {layout}: {len(c['up'])} dets", flush=True)
    return ham


@pytest.mark.gpu
def test_tiny_pattern_and_values(case):
    c = case
    ham = _ham(c, "csr")
    n = len(c["up"])
    keys = []
    for rp, cols, vals in (ham.matrix.export_upper(), c["orc"].export_upper()):
        rows = np.repeat(np.arange(n), np.diff(rp))
        order = np.argsort(rows * n + cols, kind="stable")
        keys.append(((rows * n + cols)[order], vals[order]))
    assert np.array_equal(keys[0][0], keys[1][0])
    assert np.array_equal(keys[0][1], keys[1][1])  # bit-exact FP64 values
    np.testing.assert_array_equal(ham.matrix.get_diag(), c["orc"].diag())
    assert ham.matrix.count_n_elems() == c["orc"].stats()["nnz"]


@pytest.mark.gpu
@pytest.mark.parametrize("layout", ["csr", "segments"])
def test_tiny_products(case, layout):
    c = case
    ham = _ham(c, layout)
    n = len(c["up"])
    x = np.linspace(-1.0, 1.5, n) + 0.25
    y_ref = c["orc"].mul(x)
    np.testing.assert_allclose(ham.matrix.mul(x), y_ref, rtol=1e-12, atol=1e-11)
    np.testing.assert_allclose(ham.matrix.mul(x, fp32=True), y_ref, rtol=2e-5, atol=2e-5 * np.abs(y_ref).max())
    X = np.stack([x, -x, 2 * x, np.ones(n)])
    Y = ham.matrix.mul_block(X, fp32=True)
    for v in range(4):
        yo = c["orc"].mul(X[v])
        assert np.linalg.norm(Y[v] - yo) <= 1e-5 * np.linalg.norm(yo), v


@pytest.mark.gpu
@pytest.mark.parametrize("layout", ["csr", "segments"])
def test_tiny_davidson(case, layout):
    from shci_amd.hamiltonian import Davidson

    c = case
    ham = _ham(c, layout)
    n = len(c["up"])
    init = np.zeros(n)
    init[0] = 1.0
    ev_o, _, conv_o, traj_o, _ = c["orc"].davidson(init[None], 1e-10)
    d = Davidson(1)
    d.diagonalize(ham.matrix, init, 1e-10, mode="fp64")
    assert d.converged and conv_o
    if n == 1:  # the one-dimensional branch records no Rayleigh-Ritz step
        assert len(d.trajectory) == len(traj_o) == 0
    else:
        # once the basis spans the space, the next correction's norm is
        # rounding noise near the 1e-12 stopping test (davidson.cc:125), so
        # either side may take one more step: compare the common prefix
        k = min(len(d.trajectory), len(traj_o))
        assert k >= 1 and abs(len(d.trajectory) - len(traj_o)) <= 1
        np.testing.assert_allclose(d.trajectory[:k], traj_o[:k], rtol=0, atol=1e-10)
    assert abs(d.lowest_eigenvalues[0] - ev_o[0]) < 1e-9
    dm = Davidson(1)
    dm.diagonalize(ham.matrix, init, 1e-10, mode="mixed")
    assert abs(dm.lowest_eigenvalues[0] - ev_o[0]) <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("layout,keep_zeros", [("csr", False), ("csr", True), ("segments", False)])
def test_tiny_rdm(case, layout, keep_zeros):
    from shci_amd.hamiltonian import Hamiltonian

    c = case
    ham = Hamiltonian(N_ORB, c["h1"], c["eri"], 0.0, store="fp64", layout=layout, keep_zeros=keep_zeros)
    ham.update(c["up"], c["dn"])
    n = len(c["up"])
    coefs = np.linspace(1.0, 0.3, n)[None] / np.linalg.norm(np.linspace(1.0, 0.3, n))
    r1g, r2g = ham.rdm(coefs)
    r1o, r2o = c["orc"].rdm(coefs)
    np.testing.assert_allclose(r1g, r1o, rtol=0, atol=1e-12)
    np.testing.assert_allclose(r2g, r2o, rtol=0, atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("layout", ["csr", "segments"])
def test_empty_list_refused(layout):
    from shci_amd.hamiltonian import Hamiltonian, ShciError

    _, h1, eri, _ = si.synthetic(N_ORB, seed=5)
    ham = Hamiltonian(N_ORB, h1, eri, 0.0, store="fp64", layout=layout)
    empty = np.zeros(0, dtype=np.uint64)
    with pytest.raises(ShciError, match="empty"):
        ham.update(empty, empty)
