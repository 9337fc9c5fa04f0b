"""Pins the integral handling (src/chem/integrals.cc:37-422): the synthetic
generator's symmetries, the FCIDUMP round trip through both readers (oracle and
host library), combine4 keys, and the examples/C HF energy / orbital order."""
import json

import numpy as np

import oracle
from shci_amd import integrals as si


def test_generator_symmetries_and_determinism():
    n, h1, eri, core = si.synthetic(8, seed=5)
    assert np.array_equal(h1, h1.T)
    v = si.unpack_eri(eri, n)
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1), (3, 2, 1, 0)]:
        assert np.array_equal(v, v.transpose(perm))
    n2, h12, eri2, _ = si.synthetic(8, seed=5)
    assert np.array_equal(eri, eri2) and np.array_equal(h1, h12)
    # positive semidefinite Coulomb structure: (pq|pq) >= 0 supermatrix diag
    assert (np.diagonal(v.reshape(n * n, n * n)) >= -1e-12).all()
    assert np.array_equal(si.pack_eri(v), eri)


def test_config_generator_seed_and_shapes():
    spec = si.CONFIGS[1]
    n, h1, eri, core = si.synthetic(1)
    assert n == spec.n_orb == 20 and core == 0.0
    assert h1[3, 3] == -2 + 0.25 * 3
    assert eri.shape == ((20 * 21 // 2) * (20 * 21 // 2 + 1) // 2,)


def test_fcidump_round_trip(tmp_path):
    from shci_amd.host import HostSystem

    n, h1, eri, core = si.synthetic(7, seed=13)
    path = tmp_path / "FCIDUMP"
    si.write_fcidump(path, n, h1, eri, core, n_elec=6)
    orc = oracle.System.from_fcidump(path, 3, 3, reorder=0)
    h1o, erio, _ = orc.integrals()
    np.testing.assert_array_equal(h1o, h1)
    np.testing.assert_array_equal(erio, eri)
    host = HostSystem.from_fcidump(path, 3, 3, reorder=0)
    h1h, erih, _ = host.integrals()
    np.testing.assert_array_equal(h1h, h1)
    np.testing.assert_array_equal(erih, eri)


def test_read_cut_and_duplicate_reducer(tmp_path):
    path = tmp_path / "FCIDUMP"
    path.write_text(" &FCI NORB=2,NELEC=2,MS2=0,\n ORBSYM=1,1,\n ISYM=1,\n &END\n"
                    "0.5 1 1 1 1\n-0.7 1 1 1 1\n5e-10 2 2 1 1\n0.3 2 1 1 1\n-1.0 1 1 0 0\n-0.5 2 2 0 0\n"
                    "0.1 2 1 0 0\n2.0 0 0 0 0\n")
    orc = oracle.System.from_fcidump(path, 1, 1, reorder=0)
    h1, eri, _ = orc.integrals()
    assert orc.core == 2.0
    assert eri[si.combine4(0, 0, 0, 0)] == -0.7  # larger |x| kept
    assert eri[si.combine4(1, 1, 0, 0)] == 0.0  # |x| < 1e-9 dropped
    assert eri[si.combine4(0, 1, 0, 0)] == 0.3 and eri[si.combine4(1, 0, 0, 0)] == 0.3
    assert h1[1, 0] == h1[0, 1] == 0.1


def test_examples_C_hf_energy_and_reorder(golden_dir):
    from shci_amd.host import HostSystem

    gold = json.loads((golden_dir / "examples_C.json").read_text())
    orc = oracle.System.from_fcidump(golden_dir / "C_FCIDUMP", 4, 2)
    assert abs(orc.elem(orc.hf_up, orc.hf_dn, orc.hf_up, orc.hf_dn, 0) - gold["energy_hf"]) < 1e-12
    # orbital order printed by the reference (examples/C/out "origin #")
    expected = [0, 1, 2, 3, 18, 17, 13, 4, 21, 6, 7, 5, 12, 11, 9, 8, 10, 14, 20, 19, 15, 16]
    _, _, order = orc.integrals()
    assert list(order) == expected
    host = HostSystem.from_fcidump(golden_dir / "C_FCIDUMP", 4, 2)
    assert list(host.integrals()[2]) == expected
    assert abs(host.energy_hf - gold["energy_hf"]) < 1e-12
