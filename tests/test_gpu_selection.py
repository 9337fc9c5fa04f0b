"""GPU heat-bath selection (shci_select) against the oracle's brute-force
sweep and the host's queue walk on the same seeded inputs.  The bar: the same
new dets in the same order ((up, dn) ascending) and bit-identical
eps_tried_prev, sweep after sweep; and the reference's examples/C iterations
with the selection on the device."""
import json
import os

import numpy as np
import pytest

import oracle
from helpers import variation
from shci_amd import integrals as si


def _host_dets(n, nu, nd, h1, eri, eps_list):
    """The host selection's dets at eps_list, or for ("orbs", list) every det
    over those orbitals (the lowest ones occupied first)."""
    if isinstance(eps_list, tuple):
        import itertools

        orbs = sorted(eps_list[1])
        ups = [sum(1 << k for k in c) for c in itertools.combinations(orbs, nu)]
        dns = [sum(1 << k for k in c) for c in itertools.combinations(orbs, nd)]
        hf = (sum(1 << k for k in orbs[:nu]), sum(1 << k for k in orbs[:nd]))
        dets = [hf] + [(u, d) for u in ups for d in dns if (u, d) != hf]
        return (np.array([d[0] for d in dets], dtype=np.uint64), np.array([d[1] for d in dets], dtype=np.uint64))
    from shci_amd.host import HostSystem

    host = HostSystem.from_arrays(n, nu, nd, h1, eri, 0.0)
    coefs = np.ones((1, 1))
    for eps in eps_list:
        host.select(coefs, eps)
        m = host.n_dets
        rng = np.random.default_rng(m)
        coefs = np.concatenate([[1.0], rng.choice([-1, 1], m - 1) * 10.0 ** rng.uniform(-5, -1, m - 1)])[None]
    return host.dets()


def _random_state(m, n_states, seed):
    rng = np.random.default_rng(seed)
    c = rng.choice([-1.0, 1.0], (n_states, m)) * 10.0 ** rng.uniform(-6, 0, (n_states, m))
    c[:, rng.random(m) < 0.05] = 0.0  # never walked (eps_min = inf)
    c[0, 0] = 0.9
    eps = np.full(m, 1e100)
    tried = rng.random(m) < 0.5  # walked before, with a larger eps_min, or with a smaller one
    eps[tried] = 1.000000001 * 1e-4 / np.abs(c[0, tried] + 1e-300) * 10.0 ** rng.uniform(-1, 1, tried.sum())
    return c, eps


@pytest.mark.gpu
@pytest.mark.parametrize("n,nu,nd,seed,eps_list,n_states", [
    (12, 3, 3, 7, [1e-2, 1e-3], 1),
    (20, 5, 5, 1001, [1e-2, 2e-3], 2),
    (33, 4, 3, 5, [3e-2, 1e-2], 1),
    (40, 8, 8, 1002, [3e-2], 3),
    (64, 2, 2, 9, [1e-2, 1e-3], 1),
    # every 2+2 det over the top six of 64 orbitals
    (64, 2, 2, 9, ("orbs", [58, 59, 60, 61, 62, 63]), 1),
])
def test_gpu_select_vs_oracle(n, nu, nd, seed, eps_list, n_states):
    from shci_amd.hamiltonian import Hamiltonian
    from shci_amd.selection import Selector

    _, h1, eri, core = si.synthetic(n, seed=seed)
    up, dn = _host_dets(n, nu, nd, h1, eri, eps_list)
    if len(up) > 3000:
        up, dn = up[:3000], dn[:3000]
    coefs, eps_tried = _random_state(len(up), n_states, seed)
    orc = oracle.System.from_arrays(n, nu, nd, h1, eri, core)
    ham = Hamiltonian(n, h1, eri, core)
    sel = Selector(ham)
    for k, eps_var in enumerate((3e-3, 1e-3, 2e-4)):
        print(f"[select] n={n} {nu}+{nd} dets {len(up)} states {n_states} eps {eps_var}", flush=True)
        e_gpu, e_orc = eps_tried.copy(), eps_tried.copy()
        gu, gd = sel.select(up, dn, coefs, eps_var, e_gpu)
        ou, od = orc.select(up, dn, coefs, eps_var, e_orc)
        assert len(gu) == len(ou), (len(gu), len(ou))
        assert np.array_equal(gu, ou) and np.array_equal(gd, od)
        assert np.array_equal(e_gpu, e_orc)
        eps_tried = e_gpu  # the next sweep starts from this one's eps_tried_prev
        coefs = coefs * 3.0


@pytest.mark.gpu
def test_gpu_select_lockstep_with_host():
    """Sweeps of the host's queue walk and of the device, in lockstep from
    HF with the same coefficients (cfg-2 integrals: n_orb = 40, 8 + 8)."""
    from shci_amd.hamiltonian import Hamiltonian
    from shci_amd.host import HostSystem
    from shci_amd.selection import GpuSystem

    n, nu, nd = 40, 8, 8
    _, h1, eri, core = si.synthetic(n, seed=1002)
    host = HostSystem.from_arrays(n, nu, nd, h1, eri, core)
    ham = Hamiltonian(n, h1, eri, core)
    gsys = GpuSystem(ham, host.hf_up, host.hf_dn)
    rng = np.random.default_rng(3)
    coefs = np.ones((1, 1))
    for eps in (2e-2, 5e-3, 1e-3, 1e-3, 5e-4):
        a = host.select(coefs, eps)
        b = gsys.select(coefs, eps)
        hu, hd = host.dets()
        gu, gd = gsys.dets()
        print(f"[select] lockstep eps {eps}: host +{a}, device +{b}, {len(gu)} dets "
              f"({gsys.selector.last_ms:.1f} ms)", flush=True)
        assert a == b and np.array_equal(hu, gu) and np.array_equal(hd, gd)
        m = len(gu)
        c = rng.choice([-1, 1], m) * 10.0 ** rng.uniform(-5, -0.5, m)
        c[0] = 0.95
        c[: coefs.shape[1]] = np.where(rng.random(coefs.shape[1]) < 0.7, coefs[0], c[: coefs.shape[1]])
        coefs = c[None]


@pytest.mark.gpu
def test_gpu_select_fallbacks_equal_single_pass():
    """The single pass, its repeat with an exact buffer (first buffer too
    small) and the count + chunked fill passes (budget too small: chunks
    merged and made unique again) give the same result."""
    from shci_amd.hamiltonian import Hamiltonian
    from shci_amd.selection import Selector

    n, nu, nd = 20, 5, 5
    _, h1, eri, core = si.synthetic(n, seed=1001)
    up, dn = _host_dets(n, nu, nd, h1, eri, [1e-2, 2e-3])
    coefs, eps_tried = _random_state(len(up), 1, 4)
    sel = Selector(Hamiltonian(n, h1, eri, core))
    e1 = eps_tried.copy()
    a = sel.select(up, dn, coefs, 2e-4, e1)
    print(f"[select] single pass: {len(a[0])} new dets", flush=True)
    assert len(a[0]) > 2000
    for var, val in (("SHCI_SELECT_FIRST", "1000"), ("SHCI_SELECT_CAP", "2000")):
        os.environ[var] = val
        try:
            e2 = eps_tried.copy()
            b = sel.select(up, dn, coefs, 2e-4, e2)
        finally:
            del os.environ[var]
        print(f"[select] {var}={val}: {len(b[0])} new dets", flush=True)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(e1, e2)


@pytest.mark.gpu
def test_gpu_select_rejects_bad_dets():
    from shci_amd.hamiltonian import Hamiltonian, ShciError
    from shci_amd.selection import Selector

    n = 12
    _, h1, eri, core = si.synthetic(n, seed=7)
    sel = Selector(Hamiltonian(n, h1, eri, core))
    up = np.array([0b111, 0b1011], dtype=np.uint64)
    dn = np.array([0b111, 0b11], dtype=np.uint64)  # 6 and 5 electrons
    with pytest.raises(ShciError):
        sel.select(up, dn, np.ones((1, 2)), 1e-3, np.full(2, 1e100))
    up = np.array([0b111, 1 << 12 | 0b11], dtype=np.uint64)  # orbital 12 >= n_orb
    dn = np.array([0b111, 0b111], dtype=np.uint64)
    with pytest.raises(ShciError):
        sel.select(up, dn, np.ones((1, 2)), 1e-3, np.full(2, 1e100))
    # nothing to walk: every coefficient zero
    e = np.full(2, 1e100)
    nu_, nd_ = sel.select(np.array([0b111, 0b1011], dtype=np.uint64), np.array([0b111, 0b111], dtype=np.uint64),
                          np.zeros((1, 2)), 1e-3, e)
    assert len(nu_) == 0 and np.all(e == 1e100)


@pytest.mark.gpu
def test_examples_C_gpu_selection(golden_dir):
    """The reference's examples/C run with every step on the device:
    selection, build and Davidson; per-iteration det counts, nnz and
    Davidson trajectories against the reference's log."""
    from shci_amd.hamiltonian import Davidson, Hamiltonian
    from shci_amd.host import HostSystem
    from shci_amd.selection import GpuSystem

    gold = json.loads((golden_dir / "examples_C.json").read_text())
    host = HostSystem.from_fcidump(golden_dir / "C_FCIDUMP", 4, 2)
    h1, eri, _ = host.integrals()
    ham = Hamiltonian(host.n_orbs, h1, eri, host.core, store="fp64", layout="segments")
    gsys = GpuSystem(ham, host.hf_up, host.hf_dn)
    dav = Davidson(1)

    def backend(up, dn, init, tol):
        ham.update(up, dn)
        ev = dav.diagonalize(ham.matrix, init, tol)
        _, _, nel = ham.matrix.counts()
        return ev, dav.get_lowest_eigenvectors(), dav.converged, dict(nnz=nel, traj=dav.trajectory.copy())

    recs = list(variation(gsys, backend, [5e-5], [2e-3, 1e-3, 5e-4, 2e-4, 1e-4], max_iters=8))
    assert len(recs) == 8
    for k, (rec, g) in enumerate(zip(recs, gold["iterations"])):
        print(f"iteration {k + 1}: dets {rec['n_dets']} nnz {rec['nnz']} E {rec['evals'][0]:.10f}", flush=True)
        assert rec["n_dets"] == g["n_dets"]
        if "nnz" in g:
            assert rec["nnz"] == g["nnz"]
        m = min(len(g["davidson"]), len(rec["traj"]))
        np.testing.assert_allclose(rec["traj"][:m], g["davidson"][:m], rtol=0, atol=2e-10 if k < 6 else 2e-9)
    assert abs(recs[-1]["evals"][0] - gold["energy_var"]["5.00e-05"]) < 1e-8


@pytest.mark.gpu
def test_examples_C_cpp_loop_selection_modes(golden_dir):
    """The reference's loop in C++ through the drop-in adapter
    (shci_host_run_variation), examples/C to eps 5e-5: the host queue walk and
    the adapter's device Selector give the same det list and the reference's
    final energy."""
    from shci_amd.host import HostSystem

    gold = json.loads((golden_dir / "examples_C.json").read_text())
    lists, energies = [], []
    for gpu_select in (False, True):
        host = HostSystem.from_fcidump(golden_dir / "C_FCIDUMP", 4, 2)
        e = host.run_variation([5e-5], [2e-3, 1e-3, 5e-4, 2e-4, 1e-4], gpu_select=gpu_select)
        lists.append(host.dets())
        energies.append(e[0])
        print(f"[select] C++ loop, device selection {gpu_select}: {host.n_dets} dets, E {e[0]:.10f}", flush=True)
    assert np.array_equal(lists[0][0], lists[1][0]) and np.array_equal(lists[0][1], lists[1][1])
    for e in energies:
        assert abs(e - gold["energy_var"]["5.00e-05"]) < 1e-8
