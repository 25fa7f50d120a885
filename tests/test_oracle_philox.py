"""Pins the oracle's restatement of the device-permutation algorithm:
Philox4x32-10 against the Random123 known-answer vectors, and the
cycle-walking Feistel network as a bijection."""
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
