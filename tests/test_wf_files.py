"""Wavefunction files wf_eps1_<eps>.dat (Solver::save_variation_result /
load_variation_result, src/solver/solver.h:1179-1224; fields of
BaseSystem::serialize, src/base_system.h:125-133) through
shci_host_save_wf / shci_host_load_wf.

The checkout holds no file written by the reference and its serializer (hps)
is an absent submodule, so the bytes are pinned by hps's published encoding:
the base-128 varint examples (protobuf's documentation uses the same
encoding: 1 -> 01, 300 -> ac 02) and a small file assembled byte by byte here
from the field list, plus round trips against the oracle's restatement
(oracle/wf_format.py) and the variational loop's resume from its own files."""
import json
import struct

import numpy as np
import pytest

from oracle import wf_format
from shci_amd import host as H


@pytest.mark.parametrize("v,hexs", [
    (0, "00"), (1, "01"), (127, "7f"), (128, "8001"), (300, "ac02"), (16384, "808001"),
    (2**32 - 1, "ffffffff0f"), (2**63, "80808080808080808001"), (2**64 - 1, "ffffffffffffffffff01"),
])
def test_varint_known_answers(v, hexs):
    assert wf_format.varint(v).hex() == hexs
    assert wf_format.read_varint(bytes.fromhex(hexs), 0) == (v, len(hexs) // 2)


def _hand_file():
    """n_up 2, n_dn 1, dets (0b11, 0b1) and (0b101, 0b10), one state
    (0.9, -0.1), energy_hf -1.5, energy_var (-1.75,), time_sym false."""
    b = bytes([2, 1, 2])  # n_up, n_dn, det count
    b += bytes([3, 0, 1, 0])  # det 0: up chunks 0b11, 0; dn chunks 0b1, 0
    b += bytes([5, 0, 2, 0])  # det 1
    b += bytes([1, 2]) + struct.pack("<d", 0.9) + struct.pack("<d", -0.1)  # coefs: 1 state of 2
    b += bytes.fromhex("000000000000f8bf")  # -1.5
    b += bytes([1]) + bytes.fromhex("000000000000fcbf")  # energy_var: (-1.75,)
    b += bytes([0])  # time_sym
    return b


def test_save_matches_hand_assembled_file(tmp_path):
    p = tmp_path / "wf.dat"
    H.save_wf(p, 2, 1, [0b11, 0b101], [0b1, 0b10], [[0.9, -0.1]], -1.5, [-1.75])
    assert p.read_bytes() == _hand_file()
    assert wf_format.encode(2, 1, [(0b11, 0b1), (0b101, 0b10)], [[0.9, -0.1]], -1.5, [-1.75]) == _hand_file()
    w = H.load_wf(p)
    assert (w["n_up"], w["n_dn"], w["time_sym"]) == (2, 1, False)
    assert w["up"].tolist() == [3, 5] and w["dn"].tolist() == [1, 2]
    assert w["coefs"].tolist() == [[0.9, -0.1]] and w["energy_hf"] == -1.5 and w["energy_var"].tolist() == [-1.75]


def test_round_trips_with_the_oracle(tmp_path):
    rng = np.random.default_rng(5)
    n, ns = 2000, 3
    up = rng.integers(0, 2**63, n, dtype=np.uint64) | (rng.integers(0, 2, n, dtype=np.uint64) << np.uint64(63))
    dn = rng.integers(0, 2**40, n, dtype=np.uint64)
    coefs = rng.standard_normal((ns, n)) * 10.0 ** rng.uniform(-12, 0, (ns, n))
    ev = [-75.123456789, -74.9, -74.5]
    p = tmp_path / "a.dat"
    H.save_wf(p, 7, 6, up, dn, coefs, -74.0, ev, time_sym=True)
    want = wf_format.encode(7, 6, list(zip(up.tolist(), dn.tolist())), coefs.tolist(), -74.0, ev, True)
    assert p.read_bytes() == want
    d = wf_format.decode(p.read_bytes())
    assert d["dets"] == list(zip(up.tolist(), dn.tolist())) and d["time_sym"]
    # oracle-written file -> the library's reader
    q = tmp_path / "b.dat"
    q.write_bytes(wf_format.encode(4, 4, [(9, 12), (2**64 - 1, 3)], [[0.5, -0.25]], 1.25, [-2.0]))
    w = H.load_wf(q)
    assert w["up"].tolist() == [9, 2**64 - 1] and w["dn"].tolist() == [12, 3]
    assert w["coefs"].tolist() == [[0.5, -0.25]] and not w["time_sym"]
    w = H.load_wf(p)
    assert np.array_equal(w["up"], up) and np.array_equal(w["dn"], dn)
    assert np.array_equal(w["coefs"], coefs) and w["energy_var"].tolist() == ev and w["time_sym"]


def test_rejections(tmp_path):
    good = _hand_file()
    cases = {
        "truncated": good[:-3],
        "trailing": good + b"\x00",
        "chunk1": wf_format.encode(2, 1, [(1 << 70 | 3, 1)], [[1.0]], 0.0, [0.0]),
        "time_sym_2": good[:-1] + b"\x02",
        "coef_count": wf_format.encode(2, 1, [(3, 1)], [[1.0, 2.0]], 0.0, [0.0]),
        "energies": wf_format.encode(2, 1, [(3, 1)], [[1.0]], 0.0, [0.0, 1.0]),
        "huge_count": bytes([2, 1]) + wf_format.varint(10**12),
    }
    for name, b in cases.items():
        p = tmp_path / f"{name}.dat"
        p.write_bytes(b)
        with pytest.raises(RuntimeError):
            H.load_wf(p)
    with pytest.raises(RuntimeError):
        H.load_wf(tmp_path / "missing.dat")


def test_wf_filename():
    assert H.wf_filename(5e-5) == "wf_eps1_5.00e-05.dat"
    assert H.wf_filename(1e-3) == "wf_eps1_1.00e-03.dat"


@pytest.mark.gpu
def test_variation_resumes_from_wf_files(golden_dir, tmp_path, monkeypatch):
    """examples/C through the C++ loop with wavefunction files: the first run
    writes one file per eps_var with its dets, coefficients and energies; a
    run with every file present loads them and runs nothing; a run with the
    last file removed resumes from the first and reproduces it, det for det;
    force_var runs everything again."""
    gold = json.loads((golden_dir / "examples_C.json").read_text())
    monkeypatch.chdir(tmp_path)
    eps_vars, sched = [1e-3, 2e-4], [2e-3]
    names = [H.wf_filename(e) for e in eps_vars]

    def run(**kw):
        host = H.HostSystem.from_fcidump(golden_dir / "C_FCIDUMP", 4, 2)
        e = host.run_variation(eps_vars, sched, wf_files=True, **kw)
        return host, e

    h1, e1 = run()
    up, dn = h1.dets()
    w = H.load_wf(names[1])
    assert np.array_equal(w["up"], up) and np.array_equal(w["dn"], dn)
    assert (w["n_up"], w["n_dn"], w["time_sym"]) == (4, 2, False)
    assert w["energy_var"][0] == e1[0] and abs(w["energy_hf"] - gold["energy_hf"]) < 1e-8
    c = w["coefs"][0]
    assert abs(np.dot(c, c) - 1.0) < 1e-9 and c[0] > 0
    first = H.load_wf(names[0])
    assert len(first["up"]) < len(up) and np.array_equal(first["up"], up[: len(first["up"])])
    mtime = (tmp_path / names[1]).stat().st_mtime_ns
    h2, e2 = run()  # both files load
    assert e2[0] == e1[0] and np.array_equal(h2.dets()[0], up)
    assert (tmp_path / names[1]).stat().st_mtime_ns == mtime
    (tmp_path / names[1]).unlink()
    h3, e3 = run()  # resumes from the first file
    print(f"[wf] full run E {e1[0]:.12f}, resumed E {e3[0]:.12f}, {len(up)} dets", flush=True)
    assert np.array_equal(h3.dets()[0], up) and np.array_equal(h3.dets()[1], dn)
    assert abs(e3[0] - e1[0]) < 1e-10
    h4, e4 = run(force_var=True)
    assert np.array_equal(h4.dets()[0], up) and abs(e4[0] - e1[0]) < 1e-10
