"""Pins the oracle's restatement of the device-permutation algorithm:
Philox4x32-10 against the Random123 known-answer vectors, and the
cycle-walking Feistel network as a bijection that, for an even bit count,
equals the swapping form of the same network."""
import numpy as np
import pytest

from oracle import oracle

# Random123 kat_vectors, philox4x32 with 10 rounds
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2,
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize('ctr,key,expect', KAT)
def test_philox_kat(ctr, key, expect):
    assert oracle.philox4x32_10(ctr, key) == expect


@pytest.mark.parametrize('M', [1, 2, 3, 12, 480, 1000])
def test_feistel_is_permutation(M):
    for run in (0, 5):
        perm = oracle.feistel_permutation(1234, run, M)
        assert sorted(perm.tolist()) == list(range(M))


def test_feistel_runs_differ():
    a = oracle.feistel_permutation(42, 0, 480)
    b = oracle.feistel_permutation(42, 1, 480)
    c = oracle.feistel_permutation(43, 0, 480)
    assert (a != b).mean() > 0.9 and (a != c).mean() > 0.9


def _swapping_feistel(seed, run, M):
    """The same network in the swapping form (L, R) -> (R, L ^ T_j[R]) on
    two halves of b / 2 bits, for an even bit count b: after an even number
    of rounds it must equal the oracle's alternating form."""
    bits = max(1, (M - 1).bit_length())
    assert bits % 2 == 0
    h = bits // 2
    mask = (1 << h) - 1
    entries = oracle.PERM_ROUNDS << h
    table = []
    for g in range((entries + 7) // 8):
        words = oracle.philox4x32_10((g, run & 0xFFFFFFFF, run >> 32, oracle.PERM_TAG),
                                     (seed & 0xFFFFFFFF, seed >> 32))
        for j in range(8):
            table.append((words[j >> 1] >> (16 * (j & 1))) & mask)
    out = []
    for i in range(M):
        y = i
        while True:
            L, R = y >> h, y & mask
            for j in range(oracle.PERM_ROUNDS):
                L, R = R, L ^ table[(j << h) | R]
            y = (L << h) | R
            if y < M:
                break
        out.append(y)
    return out


@pytest.mark.parametrize('M', [3, 4, 12, 200, 256, 3000])
def test_feistel_even_bits_equals_swapping_form(M):
    for run in (0, 9):
        assert oracle.feistel_permutation(77, run, M).tolist() == \
            _swapping_feistel(77, run, M)


@pytest.mark.parametrize('M', [2, 5, 480, 32512])
def test_feistel_odd_bits_is_permutation(M):
    """Odd bit counts (halves of unequal size; the L = 128 lattice has
    b = 15) still give a permutation of 0..M-1."""
    bits = max(1, (M - 1).bit_length())
    assert bits % 2 == 1
    perm = oracle.feistel_permutation(5, 3, M)
    assert sorted(perm.tolist()) == list(range(M))
