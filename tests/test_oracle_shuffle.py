"""Pins the oracle's restatement of the device permutations
(pc_permute_philox, a Philox-keyed Rao-Sandelius shuffle; no reference
counterpart: the reference draws numpy RandomState permutations,
percolate/hpc.py:195,206) and checks that its permutations are uniform.

* Philox4x32-10, in Python and in the C restatement, against the Random123
  known-answer vectors.
* The C restatement (oracle/nz_shuffle.c) against a second, pure-Python
  restatement of the same layout on top of the pinned Philox: one bucket,
  many buckets and the wide Fisher-Yates path of large buckets (small inputs
  reach the latter two through shrunk size parameters).
* Uniformity: the full distribution of all M! orderings for M <= 5 (chi
  square), in every form; the parity of the permutation (a keyed Feistel
  network on a power-of-two domain only makes even permutations); the
  position of single elements and the order of neighbours, with one bucket
  and with many (M = 40000: sixteen buckets).
"""
import math

import numpy as np
import pytest
import scipy.stats

from oracle import oracle

# Random123 kat_vectors, philox4x32 with 10 rounds
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2,
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]

DEVICE = (oracle.SHUFFLE_BUCKET, oracle.SHUFFLE_GROUP_CAP,
          oracle.SHUFFLE_WHOLE, oracle.SHUFFLE_SUB)
# shrunk size parameters: (bucket, group_cap, whole, sub)
MANY = (16, 64, 64, 4)      # many buckets, each through sub-buckets of ~4
WIDE = (2, 2, 2, 64)        # buckets above 2 through the wide Fisher-Yates path
MIXED = (4, 5, 5, 2)        # buckets of both kinds
SUBS = (65536, 65536, 65536, 1)      # one bucket, sub-buckets of ~1 element
ONE = (65536, 65536, 65536, 65536)   # one bucket, one sub-bucket: long runs
SPLIT = (4, 8, 100, 4)      # whole runs up to 100 (above the group cap), then
                            # buckets of ~4 with the wide path above 8


@pytest.mark.parametrize('ctr,key,expect', KAT)
def test_philox_kat(ctr, key, expect):
    assert oracle.philox4x32_10(ctr, key) == expect
    assert oracle.philox4x32_10_c(ctr, key) == expect


@pytest.mark.parametrize('params', [DEVICE, MANY, WIDE, MIXED, SUBS, ONE,
                                    SPLIT, (40, 20, 20, 8)])
@pytest.mark.parametrize('M', [1, 2, 3, 7, 33, 64, 65, 300, 1000, 5000])
def test_c_restatement_equals_python(M, params):
    got = oracle.bucket_permutations(0x123456789, 7, 3, M, *params)
    for r in range(3):
        want = oracle.bucket_permutation_py(0x123456789, 7 + r, M, *params)
        np.testing.assert_array_equal(got[r], want)
        assert sorted(got[r].tolist()) == list(range(M))


def test_restatements_reach_rejections_and_wide_buckets():
    """The cases above take the replacement stream of rejected 16-bit words
    and the wide Fisher-Yates path, so the C-Python equality covers both."""
    st = {}
    for r in range(3):
        oracle.bucket_permutation_py(0x123456789, 7 + r, 5000, *ONE, stats=st)
    assert st['rejected'] > 20 and st['wide'] == 0
    st = {}
    oracle.bucket_permutation_py(0x123456789, 7, 1000, *WIDE, stats=st)
    assert st['wide'] > 50


def test_sizes_outside_the_layout_are_refused():
    with pytest.raises(ValueError):
        oracle.bucket_permutations(1, 0, 1, 10, 16, 1 << 17, 64, 64)
    with pytest.raises(ValueError):
        oracle.bucket_permutations(1, 0, 1, 10, 16, 64, 1 << 17, 64)


def test_device_sizes_are_permutations_and_runs_differ():
    for M in (4096, 4097, 8192, 8193, 32512, 32768, 32769, 40000, 130560):
        p = oracle.bucket_permutations(1234, 0, 3, M)
        for r in range(3):
            assert np.array_equal(np.sort(p[r]), np.arange(M))
        assert (p[0] != p[1]).mean() > 0.99
    q = oracle.bucket_permutations(1235, 0, 1, 32512)
    assert (q[0] != oracle.bucket_permutations(1234, 0, 1, 32512)[0]).mean() > 0.99


def _lehmer_rank(perms):
    """Rank 0..M!-1 of every row of an (R, M) array of permutations."""
    R, M = perms.shape
    rank = np.zeros(R, dtype=np.int64)
    for i in range(M):
        smaller = (perms[:, i + 1:] < perms[:, i:i + 1]).sum(axis=1)
        rank = rank * (M - i) + smaller
    return rank


@pytest.mark.parametrize('params', [DEVICE, WIDE, MIXED, SUBS, (1, 8, 8, 2),
                                    (1, 2, 3, 1)])
@pytest.mark.parametrize('M', [2, 3, 4, 5])
def test_every_ordering_equally_likely(M, params):
    """All M! orderings of a small list, over 400 * M! runs: chi square of
    the counts against the uniform distribution (p > 1e-6)."""
    n_out = math.factorial(M)
    R = 400 * n_out
    perms = oracle.bucket_permutations(99, 0, R, M, *params)
    counts = np.bincount(_lehmer_rank(perms), minlength=n_out)
    stat, p = scipy.stats.chisquare(counts)
    assert p > 1e-6, (M, params, stat)


def _odd(p):
    """1 if the permutation p is odd (M - #cycles odd), else 0."""
    M = len(p)
    seen = np.zeros(M, dtype=bool)
    cycles = 0
    for s in range(M):
        if seen[s]:
            continue
        cycles += 1
        j = s
        while not seen[j]:
            seen[j] = True
            j = p[j]
    return (M - cycles) & 1


@pytest.mark.parametrize('M', [16, 64, 1024, 40000])
def test_parity_is_balanced(M):
    """Odd and even permutations equally likely (binomial test, 4000 runs;
    600 at M = 40000)."""
    R = 4000 if M < 40000 else 600
    perms = oracle.bucket_permutations(2024, 0, R, M)
    odd = sum(_odd(p) for p in perms)
    assert scipy.stats.binomtest(odd, R, 0.5).pvalue > 1e-6, odd


@pytest.mark.parametrize('M,params', [(1000, DEVICE), (40000, DEVICE),
                                      (3000, MANY), (3000, SUBS),
                                      (3000, SPLIT)])
def test_positions_uniform(M, params):
    """Where single elements land (20 equal bins, chi square) and whether a
    neighbour pair keeps its order (binomial), over many runs; elements at
    both ends and in the middle, so label boundaries of the split form and
    the sub-buckets would show."""
    R = 3000
    perms = oracle.bucket_permutations(5, 0, R, M, *params)
    pos = np.argsort(perms, axis=1)
    for e in (0, 1, M // 2, M - 1):
        counts = np.bincount(pos[:, e] * 20 // M, minlength=20)
        assert scipy.stats.chisquare(counts).pvalue > 1e-6, (M, e)
    for i in (0, M // 3, M - 2):
        before = int((perms[:, i] < perms[:, i + 1]).sum())
        assert scipy.stats.binomtest(before, R, 0.5).pvalue > 1e-6, (M, i)


def test_first_edges_pairs_uniform():
    """The unordered pair of the first two edges of M = 5 (10 pairs): the
    statistic a keyed Feistel network failed badly (chi square 486)."""
    R = 20000
    perms = oracle.bucket_permutations(3, 0, R, 5)
    a, b = np.minimum(perms[:, 0], perms[:, 1]), np.maximum(perms[:, 0], perms[:, 1])
    counts = np.bincount(a * 5 + b, minlength=25)[[i * 5 + j for i in range(5)
                                                    for j in range(i + 1, 5)]]
    assert scipy.stats.chisquare(counts).pvalue > 1e-6


def test_ordering_test_catches_a_biased_shuffle():
    """Planted mistake: the naive shuffle (swap each position with any
    position) is not uniform for M = 3; the ordering test above, at the
    same threshold, rejects it."""
    rng = np.random.RandomState(0)
    R = 20000
    perms = np.tile(np.arange(3), (R, 1))
    for j in range(3):
        k = rng.randint(0, 3, size=R)
        a = perms[np.arange(R), j].copy()
        perms[np.arange(R), j] = perms[np.arange(R), k]
        perms[np.arange(R), k] = a
    counts = np.bincount(_lehmer_rank(perms), minlength=6)
    assert scipy.stats.chisquare(counts).pvalue < 1e-6


def test_parity_test_catches_even_only_permutations():
    """Planted mistake: permutations forced even (as a Feistel network on a
    power-of-two domain makes them) fail the parity test's binomial bound."""
    R, M = 4000, 16
    rng = np.random.RandomState(1)
    odd = 0
    for _ in range(R):
        p = rng.permutation(M)
        if _odd(p):
            p[[0, 1]] = p[[1, 0]]
        odd += _odd(p)
    assert scipy.stats.binomtest(odd, R, 0.5).pvalue < 1e-6
