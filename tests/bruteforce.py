"""Independent checks that do not share code with the oracle or the GPU path:
second-quantised Hamiltonian built by applying creation/annihilation operators
to occupation bitstrings (Jordan-Wigner signs), with the determinant convention
|a, b> = (prod_{p in a, ascending} a+_{p,up}) (prod_{q in b, ascending} a+_{q,dn}) |0>
(alpha modes 0..n-1 precede beta modes n..2n-1)."""
from __future__ import annotations

import itertools

import numpy as np


def dets_all(n, n_up, n_dn):
    ups = [sum(1 << p for p in c) for c in itertools.combinations(range(n), n_up)]
    dns = [sum(1 << p for p in c) for c in itertools.combinations(range(n), n_dn)]
    return [(u, d) for u in ups for d in dns]


def _annihilate(state, k):
    if not (state >> k) & 1:
        return None, 0
    sign = -1 if bin(state & ((1 << k) - 1)).count("1") % 2 else 1
    return state & ~(1 << k), sign


def _create(state, k):
    if (state >> k) & 1:
        return None, 0
    sign = -1 if bin(state & ((1 << k) - 1)).count("1") % 2 else 1
    return state | (1 << k), sign


def apply_ops(state, ops):
    """ops: list of (kind, mode) applied right-to-left as written (last first)."""
    sign = 1
    for kind, k in reversed(ops):
        if kind == "a":
            state, s = _annihilate(state, k)
        else:
            state, s = _create(state, k)
        if state is None:
            return None, 0
        sign *= s
    return state, sign


def hamiltonian_dense(n, h1, v, core, dets):
    """<i|H|j> with H = sum_pq h_pq a+_p a_q + 1/2 sum (pq|rs) a+_p a+_r a_s a_q (spin-summed)."""
    idx = {}
    modes = []
    for i, (u, d) in enumerate(dets):
        st = u | (d << n)
        idx[st] = i
        modes.append(st)
    N = len(dets)
    H = np.zeros((N, N))
    for j, st in enumerate(modes):
        H[j, j] += core
        for p in range(n):
            for q in range(n):
                if h1[p, q] == 0.0:
                    continue
                for s in (0, n):
                    out, sg = apply_ops(st, [("c", p + s), ("a", q + s)])
                    if out is not None and out in idx:
                        H[idx[out], j] += h1[p, q] * sg
        for p in range(n):
            for q in range(n):
                for r in range(n):
                    for s in range(n):
                        val = v[p, q, r, s]
                        if val == 0.0:
                            continue
                        for sig in (0, n):
                            for tau in (0, n):
                                out, sg = apply_ops(st, [("c", p + sig), ("c", r + tau), ("a", s + tau),
                                                         ("a", q + sig)])
                                if out is not None and out in idx:
                                    H[idx[out], j] += 0.5 * val * sg
    return H


def sign_by_operators(left, right):
    """Fermionic sign of <left| ... |right> for one spin string differing by
    at most two electrons, from explicitly moving electrons: the sign of
    a+_{r1} a+_{r2} a_{l2} a_{l1} acting on `left` giving `right`, with left-only
    orbitals l1<l2 and right-only r1<r2 in ascending order (one of the two
    conventional orderings; equals the reference's permutation_factor)."""
    lo = [k for k in range(64) if (left >> k) & 1 and not (right >> k) & 1]
    ro = [k for k in range(64) if (right >> k) & 1 and not (left >> k) & 1]
    ops = [("c", r) for r in ro] + [("a", l) for l in reversed(lo)]
    out, sg = apply_ops(left, ops)
    assert out == right
    return sg
