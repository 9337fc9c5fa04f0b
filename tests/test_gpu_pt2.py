"""GPU deterministic PT2 (shci_pt_dtm) against the oracle's brute-force
restatement on the same seeded inputs (same number of PT dets; the energy to
1e-13 relative: each hc_sum adds its terms in the same order, only the final
sum over the PT dets is ordered differently), with and without hash batches;
and the reference's examples/C log with every step on the device."""
import json

import numpy as np
import pytest

import oracle
from helpers import variation
from shci_amd import integrals as si


def _host_dets(n, nu, nd, h1, eri, eps_list):
    from shci_amd.host import HostSystem

    host = HostSystem.from_arrays(n, nu, nd, h1, eri, 0.0)
    coefs = np.ones((1, 1))
    for eps in eps_list:
        host.select(coefs, eps)
        m = host.n_dets
        rng = np.random.default_rng(m)
        coefs = np.concatenate([[1.0], rng.choice([-1, 1], m - 1) * 10.0 ** rng.uniform(-5, -1, m - 1)])[None]
    return host.dets()


@pytest.mark.gpu
@pytest.mark.parametrize("n,nu,nd,seed,eps_list", [
    (12, 3, 3, 7, [1e-2, 1e-3]),
    (20, 5, 5, 1001, [1e-2, 2e-3]),
    (33, 4, 3, 5, [3e-2, 1e-2]),
    (64, 2, 2, 9, [1e-2, 1e-3]),
])
def test_gpu_pt_dtm_vs_oracle(n, nu, nd, seed, eps_list):
    from shci_amd.hamiltonian import Hamiltonian
    from shci_amd.selection import pt_dtm

    _, h1, eri, core = si.synthetic(n, seed=seed)
    up, dn = _host_dets(n, nu, nd, h1, eri, eps_list)
    up, dn = up[:2000], dn[:2000]
    rng = np.random.default_rng(seed)
    c = rng.choice([-1.0, 1.0], len(up)) * 10.0 ** rng.uniform(-5, 0, len(up))
    c[rng.random(len(up)) < 0.05] = 0.0
    c[0] = 0.9
    orc = oracle.System.from_arrays(n, nu, nd, h1, eri, core)
    ham = Hamiltonian(n, h1, eri, core)
    e_var = float(orc.elem(int(up[0]), int(dn[0]), int(up[0]), int(dn[0]), 0)) - 0.05
    for eps in (1e-3, 1e-4, 1e-5):
        eo, no = orc.pt_dtm(up, dn, c, eps, e_var)
        eg, ng, ms = pt_dtm(ham, up, dn, c, eps, e_var)
        print(f"[pt] n={n} {nu}+{nd} dets {len(up)} eps {eps}: {ng} PT dets, E {eg:.15e} ({ms:.1f} ms)", flush=True)
        assert ng == no
        assert abs(eg - eo) <= 1e-13 * abs(eo) + 1e-300
        eb, nb, _ = pt_dtm(ham, up, dn, c, eps, e_var, n_batches=3)
        assert nb == no and abs(eb - eo) <= 1e-13 * abs(eo) + 1e-300


@pytest.mark.gpu
def test_examples_C_gpu_pt_dtm(golden_dir):
    """examples/C to eps_var 5e-5 with selection, build, Davidson and PT2 on
    the device: the log's 52,762 PT dets and -0.0000088982 Ha."""
    from shci_amd.hamiltonian import Davidson, Hamiltonian
    from shci_amd.host import HostSystem
    from shci_amd.selection import GpuSystem, pt_dtm

    gold = json.loads((golden_dir / "examples_C.json").read_text())
    g = [r for r in gold["pt_dtm"] if r["eps_var"] == 5e-5][0]
    host = HostSystem.from_fcidump(golden_dir / "C_FCIDUMP", 4, 2)
    h1, eri, _ = host.integrals()
    ham = Hamiltonian(host.n_orbs, h1, eri, host.core, store="fp64", layout="segments")
    gsys = GpuSystem(ham, host.hf_up, host.hf_dn)
    dav = Davidson(1)
    last = {}

    def backend(up, dn, init, tol):
        ham.update(up, dn)
        ev = dav.diagonalize(ham.matrix, init, tol)
        vec = np.atleast_2d(dav.get_lowest_eigenvectors())
        last.update(up=up.copy(), dn=dn.copy(), c=vec[0].copy(), e=float(ev[0]))
        return ev, vec, dav.converged, {}

    list(variation(gsys, backend, [5e-5], [2e-3, 1e-3, 5e-4, 2e-4, 1e-4]))
    assert len(last["up"]) == 24087
    e, npt, ms = pt_dtm(ham, last["up"], last["dn"], last["c"], g["eps_pt_dtm"], last["e"])
    print(f"[pt] examples/C: {npt} PT dets, correction {e:.12f} (log {g['correction']:.10f}) in {ms:.1f} ms",
          flush=True)
    assert npt == g["n_pt_dets"]
    assert abs(e - g["correction"]) < 1e-10
