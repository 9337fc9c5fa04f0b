"""The oracle's deterministic PT2 (src/solver/solver.h:637-764, one batch,
restated without the queues) pinned by the reference's own example run: at
eps_var = 5e-5 (the 24,087-det wavefunction of the examples/C loop) with
eps_pt_dtm = 5e-6 the log prints "Number of dtm pt dets: 52,762" and
"PT dtm correction ... -0.0000088982 Ha"; and by the definition on a full
space: with every det variational there is no PT det, and on a tiny space
the correction equals the brute-force sum over all external dets."""
import itertools
import json

import numpy as np

import oracle
from helpers import variation
from shci_amd import integrals as si


def examples_C_wavefunction(golden_dir):
    """The examples/C variational wavefunction at eps_var = 5e-5 (oracle
    build + Davidson, oracle selection), as the reference's loop makes it."""
    from test_oracle_selection import OracleSystem

    orc = oracle.System.from_fcidump(golden_dir / "C_FCIDUMP", 4, 2)
    osys = OracleSystem(orc, orc.hf_up, orc.hf_dn)
    builder = oracle.System.from_fcidump(golden_dir / "C_FCIDUMP", 4, 2)
    last = {}

    def backend(up, dn, init, tol):
        builder.update(up, dn)
        ev, evec, conv, traj, _ = builder.davidson(init, tol)
        last.update(up=up.copy(), dn=dn.copy(), coefs=np.atleast_2d(evec)[0].copy(), energy=float(ev[0]))
        return ev, evec, conv, {}

    list(variation(osys, backend, [5e-5], [2e-3, 1e-3, 5e-4, 2e-4, 1e-4]))
    return orc, last


def test_examples_C_pt_dtm_oracle(golden_dir):
    gold = json.loads((golden_dir / "examples_C.json").read_text())
    g = [r for r in gold["pt_dtm"] if r["eps_var"] == 5e-5][0]
    orc, wf = examples_C_wavefunction(golden_dir)
    assert len(wf["up"]) == 24087
    assert abs(wf["energy"] - g["energy_var"]) < 1e-9
    e, npt = orc.pt_dtm(wf["up"], wf["dn"], wf["coefs"], g["eps_pt_dtm"], wf["energy"])
    print(f"PT dtm: {npt} dets, correction {e:.12f} (log {g['correction']:.10f}, {g['n_pt_dets']} dets)")
    assert npt == g["n_pt_dets"]
    assert abs(e - g["correction"]) < 1e-10


def test_pt_dtm_full_space_and_brute_force():
    n, nu, nd = 7, 2, 2
    _, h1, eri, core = si.synthetic(n, seed=13)
    orc = oracle.System.from_arrays(n, nu, nd, h1, eri, core)
    ups = [sum(1 << o for o in c) for c in itertools.combinations(range(n), nu)]
    dns = [sum(1 << o for o in c) for c in itertools.combinations(range(n), nd)]
    space = [(u, d) for u in ups for d in dns]
    # every det variational: no external det
    up = np.array([s[0] for s in space], dtype=np.uint64)
    dn = np.array([s[1] for s in space], dtype=np.uint64)
    rng = np.random.default_rng(2)
    c = rng.uniform(-1, 1, len(space))
    e, npt = orc.pt_dtm(up, dn, c, 1e-12, -10.0)
    assert npt == 0 and e == 0.0
    # a small variational space, eps_pt_dtm tiny: every connected external det
    # a contributes (sum_i H_ai c_i)^2 / (E - H_aa), brute force over the space
    var = space[:12]
    cv = rng.uniform(-1, 1, len(var))
    ev = -9.5
    e, npt = orc.pt_dtm([v[0] for v in var], [v[1] for v in var], cv, 1e-13, ev)
    want, cnt = 0.0, 0
    for a in space[12:]:
        s = 0.0
        hit = False
        for (u, d), ci in zip(var, cv):
            k = bin(u ^ a[0]).count("1") // 2 + bin(d ^ a[1]).count("1") // 2
            if 1 <= k <= 2:
                h = orc.elem(u, d, a[0], a[1], k)
                if abs(h * ci) >= 1e-13:
                    s += h * ci
                    hit = True
        if hit:
            cnt += 1
            want += s * s / (ev - orc.elem(a[0], a[1], a[0], a[1], 0))
    assert npt == cnt
    assert abs(e - want) < 1e-12 * max(1.0, abs(want))
