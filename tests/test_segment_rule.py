"""The segment layout's ownership rule, checked by brute force on the CPU.

The GPU segment layout (shci_amd/csrc/shci_hip.hip, SHCI_LAYOUT_SEGMENTS)
stores each connected pair of determinants once, in the one segment that
owns it:

* C segments group dets by their (alpha - 1, beta - 1) cores and own the pairs
  that differ in both spins;
* A segments group dets by (alpha core, beta string) where the core drops
  min(n_up, 2) electrons; they own alpha doubles and those alpha singles whose
  lowest common orbital lies outside the core;
* B segments are A with the spins exchanged.

This test enumerates every segment of small determinant spaces (full spaces
and random subsets, including one-electron spins) and checks that every pair
with excitation <= 2 (the reference's connection criterion,
src/solver/hamiltonian.h:430-532) is owned exactly once and that no other pair
is owned.  The GPU tests compare the resulting stored-pair counts and products
with the CSR build and the oracle.
"""
import itertools

import numpy as np
import pytest


def _bits(s):
    return [k for k in range(64) if (s >> k) & 1]


def _cores(s, rm):
    e = _bits(s)
    return [s & ~sum(1 << b for b in c) for c in itertools.combinations(e, rm)]


def _segments(up, dn):
    segs = {}
    nu, nd = bin(up[0]).count("1"), bin(dn[0]).count("1")
    ra, rb = min(nu, 2), min(nd, 2)
    for i, (u, d) in enumerate(zip(up, dn)):
        if nu and nd:
            for a in _cores(u, 1):
                for b in _cores(d, 1):
                    segs.setdefault(("C", a, b), []).append(i)
        if nu:
            for a in _cores(u, ra):
                segs.setdefault(("A", a, d), []).append(i)
        if nd:
            for b in _cores(d, rb):
                segs.setdefault(("B", b, u), []).append(i)
    return segs, ra, rb


def _owns(kind, core, n_rm, uj, dj, ul, dl):
    if kind == "C":
        return uj != ul and dj != dl
    sj, sl = (uj, ul) if kind == "A" else (dj, dl)
    if bin(sj ^ sl).count("1") != 2 or n_rm < 2:
        return True
    c = sj & sl
    return ((c & -c) & ~core) != 0


def _check(up, dn):
    segs, ra, rb = _segments(up, dn)
    owned = {}
    for (kind, core, _), mem in segs.items():
        n_rm = ra if kind == "A" else rb
        for j, l in itertools.combinations(mem, 2):
            if _owns(kind, core, n_rm, up[j], dn[j], up[l], dn[l]):
                owned[(j, l)] = owned.get((j, l), 0) + 1
    assert all(v == 1 for v in owned.values()), "a pair is owned by two segments"
    n = len(up)
    connected = {(a, b) for a in range(n) for b in range(a + 1, n)
                 if bin(up[a] ^ up[b]).count("1") + bin(dn[a] ^ dn[b]).count("1") <= 4}
    assert set(owned) == connected
    return len(connected)


def _space(n, nu, nd):
    ups = [sum(1 << k for k in c) for c in itertools.combinations(range(n), nu)]
    dns = [sum(1 << k for k in c) for c in itertools.combinations(range(n), nd)]
    return [u for u in ups for d in dns], [d for u in ups for d in dns]


@pytest.mark.parametrize("n,nu,nd", [(7, 2, 2), (7, 1, 2), (7, 2, 1), (6, 1, 1), (8, 3, 2), (7, 3, 3)])
def test_every_connected_pair_owned_once(n, nu, nd):
    up, dn = _space(n, nu, nd)
    assert _check(up, dn) > 0
    # random subsets: segments of every size, most of them partial
    rng = np.random.default_rng(n * 100 + nu * 10 + nd)
    for frac in (0.5, 0.15):
        sel = np.sort(rng.choice(len(up), max(2, int(frac * len(up))), replace=False))
        _check([up[i] for i in sel], [dn[i] for i in sel])
