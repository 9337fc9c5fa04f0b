"""Wavefunction file format of the reference -- TEST INFRASTRUCTURE ONLY (the
checker of ``shci_host_save_wf`` / ``shci_host_load_wf``).

A wf_eps1_<eps>.dat file (Solver::save_variation_result /
load_variation_result, src/solver/solver.h:1179-1224) is the serialized
BaseSystem, its fields in the order of BaseSystem::serialize
(src/base_system.h:125-133): n_up, n_dn, dets, coefs, energy_hf, energy_var,
time_sym.  A Det is up then dn (src/det/det.h:31-34); a HalfDet is its
N_CHUNKS = 2 uint64 chunks (src/det/half_det.h:10,58-66), orbital p in
chunk p // 64.

The byte encoding is the hps serializer's (the submodule is absent from the
reference checkout; restated from its published format): unsigned integers
(``unsigned``, ``size_t``, ``uint64_t``, ``bool``) as base-128 varints, the
low 7-bit group first and bit 7 set on every byte but the last; ``double`` as
its 8 IEEE-754 bytes, little-endian; a ``std::vector`` as its size (varint)
followed by its elements.  Pinned in tests/test_wf_files.py by the published
varint examples and a file assembled byte by byte from that description; no
file written by the reference exists in its checkout, so agreement with the
reference's own bytes is not pinned beyond the published format."""
from __future__ import annotations

import struct


def varint(v: int) -> bytes:
    if v < 0:
        raise ValueError("varints encode unsigned integers")
    out = bytearray()
    while v >= 128:
        out.append((v & 127) | 128)
        v >>= 7
    out.append(v)
    return bytes(out)


def read_varint(b: bytes, pos: int):
    v, shift = 0, 0
    while True:
        c = b[pos]
        pos += 1
        v |= (c & 127) << shift
        shift += 7
        if not c & 128:
            return v, pos


def encode(n_up, n_dn, dets, coefs, energy_hf, energy_var, time_sym=False) -> bytes:
    """dets: list of (up, dn) ints of up to 128 orbitals; coefs: list of
    per-state lists."""
    out = bytearray(varint(n_up) + varint(n_dn) + varint(len(dets)))
    mask = (1 << 64) - 1
    for up, dn in dets:
        for h in (up, dn):
            out += varint(h & mask) + varint(h >> 64)
    out += varint(len(coefs))
    for c in coefs:
        out += varint(len(c)) + b"".join(struct.pack("<d", x) for x in c)
    out += struct.pack("<d", energy_hf)
    out += varint(len(energy_var)) + b"".join(struct.pack("<d", x) for x in energy_var)
    out += varint(1 if time_sym else 0)
    return bytes(out)


def decode(b: bytes) -> dict:
    pos = 0
    n_up, pos = read_varint(b, pos)
    n_dn, pos = read_varint(b, pos)
    nd, pos = read_varint(b, pos)
    dets = []
    for _ in range(nd):
        h = []
        for _ in range(2):
            lo, pos = read_varint(b, pos)
            hi, pos = read_varint(b, pos)
            h.append(lo | hi << 64)
        dets.append(tuple(h))
    ns, pos = read_varint(b, pos)
    coefs = []
    for _ in range(ns):
        m, pos = read_varint(b, pos)
        coefs.append(list(struct.unpack_from(f"<{m}d", b, pos)))
        pos += 8 * m
    (energy_hf,) = struct.unpack_from("<d", b, pos)
    pos += 8
    ne, pos = read_varint(b, pos)
    energy_var = list(struct.unpack_from(f"<{ne}d", b, pos))
    pos += 8 * ne
    ts, pos = read_varint(b, pos)
    if pos != len(b):
        raise ValueError("trailing bytes")
    return dict(n_up=n_up, n_dn=n_dn, dets=dets, coefs=coefs, energy_hf=energy_hf, energy_var=energy_var,
                time_sym=bool(ts))
