"""GPU path vs the reference's own example run (examples/C/out, result.json):
per-iteration determinant counts, "Number of nonzero elems" and every
"Davidson #k" energy printed by the reference, for the first variational
iterations (CPU selection feeding the GPU build + Davidson).

The log was written by an older revision of the reference whose final-eps
Davidson stopped at an eigenvalue change of ~2e-9; the current code
(src/solver/solver.h:500-501) uses target_error/500000 = 1e-10, which this
path follows, so final-eps trajectories are compared over the logged prefix
and the final energy to 1e-8 Ha."""
import json

import numpy as np
import pytest

from helpers import variation


@pytest.mark.gpu
@pytest.mark.parametrize("layout", ["csr", "segments"])
def test_examples_C_gpu_iterations(golden_dir, layout):
    from shci_amd.hamiltonian import Davidson, Hamiltonian
    from shci_amd.host import HostSystem

    gold = json.loads((golden_dir / "examples_C.json").read_text())
    host = HostSystem.from_fcidump(golden_dir / "C_FCIDUMP", 4, 2)
    h1, eri, _ = host.integrals()
    ham = Hamiltonian(host.n_orbs, h1, eri, host.core, store="fp64", layout=layout)
    dav = Davidson(1)

    def backend(up, dn, init, tol):
        ham.update(up, dn)
        ev = dav.diagonalize(ham.matrix, init, tol)
        _, _, nel = ham.matrix.counts()
        return ev, dav.get_lowest_eigenvectors(), dav.converged, dict(nnz=nel, traj=dav.trajectory.copy())

    recs = list(variation(host, backend, [5e-5], [2e-3, 1e-3, 5e-4, 2e-4, 1e-4], max_iters=8))
    assert len(recs) == 8
    for k, (rec, g) in enumerate(zip(recs, gold["iterations"])):
        print(f"iteration {k + 1}: dets {rec['n_dets']} nnz {rec['nnz']} E {rec['evals'][0]:.10f}", flush=True)
        assert rec["n_dets"] == g["n_dets"]
        if "nnz" in g:
            assert rec["nnz"] == g["nnz"]
        m = len(g["davidson"])
        if g["eps"] > 5e-5:  # schedule extras: same tolerance as the log (target_error / 50)
            assert len(rec["traj"]) == m
        m = min(m, len(rec["traj"]))
        np.testing.assert_allclose(rec["traj"][:m], g["davidson"][:m], rtol=0, atol=2e-10 if k < 6 else 2e-9)
    assert abs(recs[-1]["evals"][0] - gold["energy_var"]["5.00e-05"]) < 1e-8
