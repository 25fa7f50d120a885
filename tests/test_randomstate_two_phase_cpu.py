"""The two-phase form of numpy's legacy shuffle, restated on the CPU.

The device draws numpy ``RandomState(seed).permutation(n)`` (the reference's
permutation, percolate/hpc.py:195,206) for n <= 32768 in two kernels:
``randomstate_targets_kernel`` settles the MT19937 words 64 at a time into
the swap target k_i of every step i, and ``randomstate_apply_kernel``
applies all swaps at once through F (the first later step writing a
position) and the chain roots V.  This file restates both phases in plain
Python, step for step like the kernels, and checks the result against
numpy itself; the GPU tests compare the kernels with numpy directly.  One
test plants a mistake in the restatement (the latest instead of the first
later writer) and shows that the comparison catches it.
"""
import numpy as np
import pytest

MT_N, MT_M = 624, 397


def mt_words(seed):
    """Tempered MT19937 words after init_genrand(seed) (Matsumoto &
    Nishimura 1998), in blocks of at most 64 aligned to the 624-word state,
    the way the targets kernel consumes them."""
    mt = [seed & 0xFFFFFFFF]
    for i in range(1, MT_N):
        mt.append((1812433253 * (mt[-1] ^ (mt[-1] >> 30)) + i) & 0xFFFFFFFF)
    while True:
        for k in range(MT_N):
            y = (mt[k] & 0x80000000) | (mt[(k + 1) % MT_N] & 0x7FFFFFFF)
            mt[k] = mt[(k + MT_M) % MT_N] ^ (y >> 1) ^ (0x9908B0DF if y & 1 else 0)
        words = []
        for y in mt:
            y ^= y >> 11
            y ^= (y << 7) & 0x9D2C5680
            y ^= (y << 15) & 0xEFC60000
            y ^= y >> 18
            words.append(y)
        for b in range(0, MT_N, 64):
            yield words[b:b + 64]


def targets(seed, n):
    """Phase one: k[i] for steps i = n-1 .. 1.  Under one mask m the word of
    lane l serves step i - (accepted lanes below l): accepted for sure when
    x <= i - 63, rejected for sure when x > i, the lanes in between decided
    in lane order; the mask holds for c = i - m // 2 more steps, and lanes
    after the c-th accepted one are settled again under the next mask."""
    k = [0] * n
    i = n - 1 if n else 0
    blocks = mt_words(seed)
    while i >= 1:
        w = next(blocks)
        nw, l0 = len(w), 0
        while l0 < nw and i >= 1:
            m = (1 << i.bit_length()) - 1
            c = i - (m >> 1)
            x = [v & m for v in w]
            acc = [False] * nw
            for lane in range(l0, nw):
                if x[lane] + 63 <= i:
                    acc[lane] = True
            for lane in range(l0, nw):
                if not acc[lane] and x[lane] <= i:
                    acc[lane] = x[lane] + sum(acc[:lane]) <= i
            taken = [lane for lane in range(nw) if acc[lane]]
            if len(taken) >= c:
                taken = taken[:c]
                l0 = taken[-1] + 1
            else:
                l0 = nw
            for below, lane in enumerate(taken):
                k[i - below] = x[lane]
            i -= len(taken)
    return k


def apply_swaps(k, n, first=True):
    """Phase two: result[i] = V(s) for the smallest step s > i with
    k_s = k_i (k_i itself without one), result[0] = V(0), where
    V(p) = V(F(p)) or p, F(p) the smallest step s > p with k_s = p.  V by
    pointer jumping.  first=False takes the largest such step instead (a
    planted mistake)."""
    pick = min if first else max
    lists = [[] for _ in range(n)]
    for s in range(1, n):
        lists[k[s]].append(s)
    F = []
    for p in range(n):
        later = [s for s in lists[p] if s > p]
        F.append(pick(later) if later else p)
    # the kernel's way: the smallest step j of each list knows F of its
    # target p, j itself when j > p, else (j = p) its successor
    G = list(range(n))
    for j in range(1, n):
        p = k[j]
        if min(lists[p]) == j:
            later = [s for s in lists[p] if s > j]
            G[p] = j if j > p else (min(later) if later else p)
    if first:
        assert G == F
    V = F[:]
    while True:
        nxt = [V[V[p]] for p in range(n)]
        if nxt == V:
            break
        V = nxt
    out = [V[0]] if n else []
    for i in range(1, n):
        later = [s for s in lists[k[i]] if s > i]
        out.append(V[pick(later)] if later else k[i])
    return out


def sequential(k, n):
    """numpy's own order: for i = n-1 .. 1 swap a[i] and a[k[i]]."""
    a = list(range(n))
    for i in range(n - 1, 0, -1):
        a[i], a[k[i]] = a[k[i]], a[i]
    return a


def test_words_are_numpys_stream():
    """The restated MT19937 stream is numpy's: full-range 32-bit draws take
    one word each."""
    for seed in (0, 1, 4294967295):
        words = [v for _, blk in zip(range(40), mt_words(seed)) for v in blk][:2000]
        want = np.random.RandomState(seed).randint(0, 2 ** 32, size=2000, dtype=np.uint64)
        assert words == [int(v) for v in want]


@pytest.mark.parametrize('n', [1, 2, 3, 5, 63, 64, 65, 129, 1000, 4097])
def test_two_phases_equal_numpy(n):
    for seed in (0, 1, 12345, 4294967295):
        k = targets(seed, n)
        want = list(np.random.RandomState(seed).permutation(n))
        assert sequential(k, n) == want, (n, seed)
        assert apply_swaps(k, n) == want, (n, seed)


def test_planted_mistake_is_caught():
    """Taking the latest instead of the first later writer breaks the
    equality with numpy."""
    n, seed = 1000, 7
    k = targets(seed, n)
    assert apply_swaps(k, n) == list(np.random.RandomState(seed).permutation(n))
    assert apply_swaps(k, n, first=False) != list(np.random.RandomState(seed).permutation(n))
