"""Pins HalfDet::diff (src/det/half_det.cc:120-206): the reference's own unit
test vectors (src/det/half_det_test.cc:30-46) and, for the fermionic sign the
reference tests leave unchecked, brute-force operator application on every
pair of strings of up to 8 orbitals."""
import itertools

import pytest

import oracle
from bruteforce import sign_by_operators


def test_half_det_test_vectors():
    # HalfDetTest.Diff: {0} vs {1}, then {0,1} vs {1,3}
    nd, lo, ro, _ = oracle.diff(0b1, 0b10)
    assert nd == 1 and lo[0] == 0 and ro[0] == 1
    nd, lo, ro, _ = oracle.diff(0b11, 0b1010)
    assert nd == 1 and lo[0] == 0 and ro[0] == 3


def test_n_diffs_saturates_at_three():
    assert oracle.diff(0b111, 0b111000)[0] == 3


@pytest.mark.parametrize("n,ne", [(6, 3), (8, 4), (8, 2)])
def test_permutation_factor_brute_force(n, ne):
    strings = [sum(1 << p for p in c) for c in itertools.combinations(range(n), ne)]
    checked = 0
    for a in strings:
        for b in strings:
            nd, lo, ro, pf = oracle.diff(a, b)
            k = bin(a & ~b).count("1")
            if k > 2:
                assert nd == 3
                continue
            assert nd == k
            assert list(lo[:k]) == sorted(i for i in range(n) if (a >> i) & 1 and not (b >> i) & 1)
            assert list(ro[:k]) == sorted(i for i in range(n) if (b >> i) & 1 and not (a >> i) & 1)
            assert pf == sign_by_operators(a, b)
            checked += 1
    assert checked > len(strings)
