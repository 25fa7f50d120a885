"""Pins the oracle's exact (unwrapped) moments and sums, and its
microcanonical confidence bounds where the standard deviation is positive.

The reference keeps the moments in numpy uint64 (percolate/hpc.py:286,297,
302), which wrap mod 2**64; the oracle carries them exactly in 128 bits and
its low 64 bits are the reference's values.  Nothing in the reference fixes
the exact values, so they are pinned here by brute force: connected
components of a chain counted with numpy and summed with Python ints, on
runs whose clusters pass 2**16 nodes (size**4 > 2**64)."""
import numpy as np
import pytest
import scipy.stats

from oracle import oracle
from tests.helpers import spanning_2d_grid

BIG = 1 << 17


def chain_graph(n):
    return {'num_nodes': n, 'num_edges': n - 1, 'spanning': False,
            'edges': np.stack([np.arange(n - 1), np.arange(1, n)], axis=1),
            'aux_side': np.zeros(0, np.int32),
            'aux_edges': np.zeros((0, 2), np.int64),
            'aux_edge_side': np.zeros(0, np.int32)}


def chain_permutation(n, sizes, rng):
    """Bonds of a chain of n nodes that first grow clusters of the given
    sizes on consecutive node ranges, one after the other, then every other
    bond in random order."""
    first, at = [], 0
    for s in sizes:
        first.append(np.arange(at, at + s - 1))
        at += s
    first = np.concatenate(first)
    rest = np.setdiff1d(np.arange(n - 1), first)
    return np.concatenate([first, rng.permutation(rest)])


def chain_cluster_sizes(n, perm, s):
    """Cluster sizes of a chain after the first s bonds of perm: the chain
    is cut at every unoccupied bond."""
    occupied = np.zeros(n - 1, dtype=bool)
    occupied[perm[:s]] = True
    cuts = np.flatnonzero(~occupied)
    return np.diff(np.concatenate([[-1], cuts, [n - 1]]))


def brute_observables(sizes):
    """Largest cluster and exact moments k = 0..4 of all other clusters
    (Python ints): the moments the reference's bookkeeping defines
    (percolate/hpc.py:282-305), without any wrap."""
    sizes = [int(x) for x in sizes]
    big = max(sizes)
    return big, [sum(x ** k for x in sizes) - big ** k for k in range(5)]


def _chain_runs(n, size_sets, seed):
    rng = np.random.RandomState(seed)
    return np.stack([chain_permutation(n, s, rng) for s in size_sets])


N_CHAIN = 2 * BIG + 70000 + 10
SIZE_SETS = [(BIG, BIG - 3), (BIG - 5, BIG, 70000), (3, 65537, BIG + 9)]


def _check_steps(n):
    """Steps around the phases of the runs above, plus a spread."""
    marks = [0, 1, 2, 65535, 65536, 70000, BIG - 1, BIG, BIG + 70000,
             2 * BIG - 10, 2 * BIG - 4, 2 * BIG, n - 2000, n - 1]
    return sorted(set(m for m in marks if 0 <= m <= n - 1) |
                  set(np.linspace(0, n - 1, 25).astype(int).tolist()))


def test_exact_moments_past_2_64_match_brute_force():
    """Runs whose non-largest clusters pass 2**16 nodes: the oracle's exact
    moments equal brute-force component sums at every checked step, its
    low words equal them mod 2**64, and some values do exceed 2**64."""
    n = N_CHAIN
    g = chain_graph(n)
    perms = _chain_runs(n, SIZE_SETS, 3)
    tr = oracle.bond_runs(g, perms, exact=True)
    exact = oracle.exact_moments(tr)
    wide = 0
    for r in range(perms.shape[0]):
        for s in _check_steps(n):
            big, mom = brute_observables(chain_cluster_sizes(n, perms[r], s))
            assert tr['max_cluster_size'][r, s] == big, (r, s)
            assert exact[r, s].tolist() == mom, (r, s)
            assert tr['moments'][r, s].tolist() == [m % (1 << 64) for m in mom]
            wide += any(m >= 1 << 64 for m in mom)
    assert wide > 0, 'no checked value exceeds 2**64'


def test_exact_sums_match_python_ints():
    """Exact per-step sums over runs (sum of x in 128 bits, of x**2 in 192)
    equal Python-int sums of the exact traces, both from the traces and
    straight from the sweep (nzo_bond_sums, here in two chunks)."""
    n = N_CHAIN
    g = chain_graph(n)
    perms = _chain_runs(n, SIZE_SETS, 4)
    tr = oracle.bond_runs(g, perms, exact=True)
    exact = oracle.exact_moments(tr)
    cols = [tr['max_cluster_size'].astype(object),
            tr['has_spanning_cluster'].astype(object)] + \
        [exact[:, :, k] for k in range(5)]
    s1, s2 = oracle.microcanonical_sums(tr)
    steps = _check_steps(n)
    for c, x in enumerate(cols):
        xs = x[:, steps]
        assert (s1[c, steps] == xs.sum(axis=0)).all(), c
        assert (s2[c, steps] == (xs * xs).sum(axis=0)).all(), c
    assert max(int(v) for v in s2[6, steps]) >= 1 << 128
    a1, a2 = oracle.bond_sums(g, perms[:1])
    oracle.bond_sums(g, perms[1:], sums=(a1, a2), threads=2)
    assert (oracle.limbs_to_int(a1) == s1).all()
    assert (oracle.limbs_to_int(a2) == s2).all()


def test_bond_sums_equal_trace_sums_on_lattice():
    """On spanning_2d_grid(12) with RandomState permutations the sweep's
    direct sums equal the sums of its traces (spanning column included)."""
    g = oracle.percolation_graph_arrays(spanning_2d_grid(12), True)
    perms = oracle.randomstate_permutations(range(30), g['num_edges'])
    s1, s2 = oracle.microcanonical_sums(oracle.bond_runs(g, perms))
    a1, a2 = oracle.bond_sums(g, perms[:7], threads=3)
    oracle.bond_sums(g, perms[7:], sums=(a1, a2), threads=5)
    assert (oracle.limbs_to_int(a1) == s1).all()
    assert (oracle.limbs_to_int(a2) == s2).all()


@pytest.mark.parametrize('runs,alpha', [(40, 2 * scipy.stats.norm.cdf(-1.0)),
                                        (7, 0.05), (2, 0.3)])
def test_microcanonical_step_stats_invert_through_t_and_beta_cdf(runs, alpha):
    """percolate.py:557-570, 613-635, 681-705 where the samples vary: the
    mean equals the exact rational mean, the standard deviation the exact
    sample variance (ddof = 1), and the confidence bounds invert through
    scipy's Student-t CDF (R - 1 degrees of freedom, scale std / sqrt(R))
    to alpha / 2 and 1 - alpha / 2; the spanning bounds invert through the
    beta CDF with parameters (k + 1, R - k + 1)."""
    from fractions import Fraction
    rng = np.random.RandomState(runs)
    mx = rng.randint(1, 5000, size=runs).astype(np.float64)
    mom = rng.randint(0, 10 ** 6, size=(runs, 5)).astype(np.float64)
    mom[:, 2] = 17.0  # one constant column: its interval collapses
    span = np.arange(runs) % 3 == 0
    st = oracle.microcanonical_step_stats(mx, mom, span, alpha)

    def exact_mean_std(x):
        q = [Fraction(int(v)) for v in x]
        m = sum(q) / len(q)
        var = sum((v - m) ** 2 for v in q) / (len(q) - 1)
        return float(m), float(var) ** 0.5

    def check(value, ci, x):
        m, sd = exact_mean_std(x)
        assert value == pytest.approx(m, rel=1e-15)
        if sd == 0:
            assert ci.tolist() == [m, m]
            return
        cdf = scipy.stats.t.cdf(ci, df=runs - 1, loc=m, scale=sd / np.sqrt(runs))
        np.testing.assert_allclose(cdf, [alpha / 2, 1 - alpha / 2], rtol=1e-9)

    check(st['max_cluster_size'], st['max_cluster_size_ci'], mx)
    for k in range(5):
        check(st['moments'][k], st['moments_ci'][k], mom[:, k])
    k = int(span.sum())
    assert st['spanning_cluster'] == pytest.approx((k + 1) / (runs + 2), rel=1e-15)
    cdf = scipy.stats.beta.cdf(st['spanning_cluster_ci'], k + 1, runs - k + 1)
    np.testing.assert_allclose(cdf, [alpha / 2, 1 - alpha / 2], rtol=1e-9)


def test_microcanonical_step_stats_catch_a_planted_error():
    """The inversion check above fails for a bound computed with the wrong
    degrees of freedom (R instead of R - 1), so it can tell them apart."""
    runs, alpha = 7, 0.05
    rng = np.random.RandomState(1)
    x = rng.randint(1, 100, size=runs).astype(np.float64)
    m, sd = x.mean(), x.std(ddof=1)
    wrong = scipy.stats.t.interval(1 - alpha, df=runs, loc=m, scale=sd / np.sqrt(runs))
    cdf = scipy.stats.t.cdf(wrong, df=runs - 1, loc=m, scale=sd / np.sqrt(runs))
    assert np.abs(cdf - [alpha / 2, 1 - alpha / 2]).max() > 1e-4


@pytest.mark.parametrize('length', [1, 2, 3, 7, 12])
def test_square_lattice_graph_equals_networkx(length):
    """The oracle's closed-form lattice gives the same traces as the
    networkx-built graph (same edge order, same sides)."""
    a = oracle.square_lattice_graph(length)
    b = oracle.percolation_graph_arrays(spanning_2d_grid(length), True)
    assert a['num_nodes'] == b['num_nodes'] and a['num_edges'] == b['num_edges']
    np.testing.assert_array_equal(a['edges'], b['edges'])
    perms = oracle.randomstate_permutations(range(6), a['num_edges'])
    ta, tb = oracle.bond_runs(a, perms), oracle.bond_runs(b, perms)
    for k in ta:
        np.testing.assert_array_equal(ta[k], tb[k])
