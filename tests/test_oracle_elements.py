"""Pins the Slater-Condon elements (src/chem/chem_system.cc:409-609) against a
Jordan-Wigner second-quantised Hamiltonian built independently (bruteforce.py),
element by element over the full determinant space of small systems, and the
lowest eigenvalue against dense diagonalisation."""
import numpy as np
import pytest

import oracle
from bruteforce import dets_all, hamiltonian_dense
from shci_amd import integrals as si


@pytest.mark.parametrize("n,nu,nd,seed", [(5, 2, 2, 3), (6, 2, 1, 5), (4, 2, 2, 9)])
def test_elements_match_second_quantisation(n, nu, nd, seed):
    _, h1, eri, _ = si.synthetic(n, seed=seed)
    core = 0.37
    v = si.unpack_eri(eri, n)
    dets = dets_all(n, nu, nd)
    Hb = hamiltonian_dense(n, h1, v, core, dets)
    orc = oracle.System.from_arrays(n, nu, nd, h1, eri, core)
    Ho = np.zeros_like(Hb)
    for i, (ui, di) in enumerate(dets):
        for j, (uj, dj) in enumerate(dets):
            Ho[i, j] = orc.elem(ui, di, uj, dj, 0 if i == j else -1)
    np.testing.assert_allclose(Ho, Hb, rtol=0, atol=1e-12)
    np.testing.assert_allclose(Ho, Ho.T, rtol=0, atol=0)  # H_ij == H_ji bit for bit


def test_full_space_davidson_equals_dense_fci():
    n, nu, nd = 6, 3, 2
    _, h1, eri, _ = si.synthetic(n, seed=21)
    dets = dets_all(n, nu, nd)
    # put the HF det first (lowest-filled)
    hf = ((1 << nu) - 1, (1 << nd) - 1)
    dets.remove(hf)
    dets = [hf] + dets
    orc = oracle.System.from_arrays(n, nu, nd, h1, eri, 0.0)
    up = np.array([d[0] for d in dets], dtype=np.uint64)
    dn = np.array([d[1] for d in dets], dtype=np.uint64)
    orc.update(up, dn)
    init = np.zeros(len(dets))
    init[0] = 1.0
    vec = init[None]
    for _ in range(50):
        ev, vec, conv, _, _ = orc.davidson(vec, 1e-12)
        if conv:
            break
    Hb = hamiltonian_dense(n, h1, si.unpack_eri(eri, n), 0.0, dets)
    assert abs(ev[0] - np.linalg.eigvalsh(Hb)[0]) < 1e-9
