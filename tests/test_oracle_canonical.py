"""Pins the oracle's FP64 restatements: binomial weights, canonical
convolution, Chan merge, finalize and microcanonical statistics."""
import functools

import numpy as np
import pytest
import scipy.stats

from oracle import oracle
from tests.helpers import ratio_weights, spanning_2d_grid


@pytest.mark.parametrize('n,p', [(100, 0.1), (480, 0.5), (32512, 0.45), (1000, 0.01)])
def test_binomial_pmf_vs_scipy(golden, n, p):
    """percolate/test/test_percolate.py:283-295 (rtol 1e-7), at more sizes."""
    ref = golden('binomial.npz')['n{}_p{}'.format(n, p)]
    got = oracle.binomial_pmf(n, p)
    assert got.sum() == pytest.approx(1.0, abs=1e-12)
    big = ref > 1e-250
    np.testing.assert_allclose(got[big], ref[big], rtol=1e-7)


@pytest.mark.parametrize('n,p,k0,k1', [(32512, 0.47, 15000, 15600),
                                        (480, 0.5, 0, 481), (480, 0.3, 100, 190),
                                        (2095104, 0.3, 628400, 628700),
                                        (7, 0.0, 0, 8), (7, 1.0, 0, 8)])
def test_ratio_weights_exact(n, p, k0, k1):
    """The long-double reference weights of the fused-contraction tests
    against exact rational arithmetic (Fraction of the binary p), to 1e-15
    of the largest weight: C(n, k) p^k (1-p)^(n-k) itself for n < 1000, the
    exact ratio recursion for the large n."""
    import math
    from fractions import Fraction
    k = np.arange(k0, k1)
    got = ratio_weights(n, p, k).astype(np.float64)
    P = Fraction(p)
    Q = 1 - P
    if n < 1000:
        terms = [math.comb(n, int(j)) * P ** int(j) * Q ** (n - int(j)) for j in k]
    else:
        terms = [Fraction(1)]
        for j in range(k0 + 1, k1):
            terms.append(terms[-1] * Fraction(n - j + 1, j) * P / Q)
    tot = sum(terms)
    exact = np.array([float(t / tot) for t in terms])
    assert np.abs(got - exact).max() <= 1e-15 * exact.max()


def test_binomial_pmf_edges():
    np.testing.assert_array_equal(oracle.binomial_pmf(10, 0.0), np.eye(11)[0])
    np.testing.assert_array_equal(oracle.binomial_pmf(10, 1.0), np.eye(11)[10])
    assert oracle.binomial_pmf(0, 0.3).tolist() == [1.0]


def test_canonical_statistics_closed_form():
    """hpc.py:443-515 is np.sum(f * X) per field (test_hpc.py:337-378)."""
    g = oracle.percolation_graph_arrays(spanning_2d_grid(3), True)
    tr = oracle.bond_runs(g, oracle.randomstate_permutations([5], g['num_edges']))
    one = {k: v[0] for k, v in tr.items()}
    f = oracle.binomial_pmf(g['num_edges'], 0.4)
    st = oracle.canonical_statistics(one, f)
    assert st['percolation_probability'] == pytest.approx(
        sum(f[i] * one['has_spanning_cluster'][i] for i in range(f.size)))
    for k in range(5):
        assert st['moments'][k] == pytest.approx(
            sum(f[i] * float(one['moments'][i, k]) for i in range(f.size)))


def test_bond_reduce_matches_numpy():
    """Chan merge == numpy mean and sum of squared deviations
    (test_hpc.py:454-524, rtol 1e-7 / atol 1e-6)."""
    rng = np.random.RandomState(3)
    vals = rng.rand(17, 6) * 100
    inits = [{'number_of_runs': np.ones(6, dtype=np.uint32),
              'x_mean': v.copy(), 'x_m2': np.zeros(6)} for v in vals]
    res = functools.reduce(oracle.bond_reduce, inits)
    assert (res['number_of_runs'] == 17).all()
    np.testing.assert_allclose(res['x_mean'], vals.mean(axis=0), rtol=1e-12)
    np.testing.assert_allclose(res['x_m2'], ((vals - vals.mean(axis=0)) ** 2).sum(axis=0),
                               rtol=1e-10)
    # associativity: a different bracketing gives the same result
    left = functools.reduce(oracle.bond_reduce, inits[:8])
    right = functools.reduce(oracle.bond_reduce, inits[8:])
    both = oracle.bond_reduce(left, right)
    np.testing.assert_allclose(both['x_m2'], res['x_m2'], rtol=1e-12)


def test_finalize_ci_inverts_through_t_cdf():
    """hpc.py:752-834: CI bounds invert through t.cdf to alpha/2, 1 - alpha/2
    (test_hpc.py:527-694)."""
    rng = np.random.RandomState(1)
    runs = 12
    vals = rng.rand(runs, 4) * 10 + 1
    avg = {'number_of_runs': np.full(4, runs, dtype=np.uint32),
           'percolation_probability_mean': vals.mean(axis=0),
           'percolation_probability_m2': ((vals - vals.mean(0)) ** 2).sum(0),
           'max_cluster_size_mean': vals.mean(axis=0) * 3,
           'max_cluster_size_m2': ((vals * 3 - vals.mean(0) * 3) ** 2).sum(0),
           'moments_mean': np.tile(vals.mean(axis=0)[:, None], (1, 5)),
           'moments_m2': np.tile(((vals - vals.mean(0)) ** 2).sum(0)[:, None], (1, 5))}
    alpha = 0.1
    out = oracle.finalize_canonical_averages(7, np.linspace(0.1, 0.4, 4), avg, alpha)
    np.testing.assert_allclose(out['percolation_strength_mean'],
                               avg['max_cluster_size_mean'] / 7)
    for key in ('percolation_probability', 'percolation_strength'):
        for bound, a in ((0, alpha / 2), (1, 1 - alpha / 2)):
            cdf = scipy.stats.t.cdf(out[key + '_ci'][:, bound], loc=out[key + '_mean'],
                                    scale=out[key + '_std'] / np.sqrt(runs), df=runs - 1)
            np.testing.assert_allclose(cdf, a, rtol=1e-4)


def test_microcanonical_initial_step_kat(golden):
    """percolate/test/test_percolate.py:234-265: n = 0 with the default 40
    runs: spanning 1/42, beta CI inverting to Phi(-1), collapsed t CIs."""
    alpha = 2 * scipy.stats.norm.cdf(-1.0)
    runs = 40
    st = oracle.microcanonical_step_stats(np.ones(runs), np.ones((runs, 5)) * 8,
                                          np.zeros(runs, dtype=bool), alpha)
    assert st['max_cluster_size'] == 1.0
    np.testing.assert_allclose(st['max_cluster_size_ci'], np.ones(2))
    np.testing.assert_allclose(st['moments_ci'], np.ones((5, 2)) * 8)
    assert st['spanning_cluster'] == 1 / 42
    np.testing.assert_allclose(
        np.array([0, 1]) + np.array([1, -1]) *
        scipy.stats.beta.cdf(st['spanning_cluster_ci'], a=1, b=41),
        scipy.stats.norm.cdf(-1) * np.ones(2))


def test_canonical_averages_linearity():
    """percolate.py:1112-1223: a constant column convolves to itself."""
    arrays = {'N': 4, 'M': 10, 'max_cluster_size': np.full(11, 0.5),
              'max_cluster_size_ci': np.full((11, 2), 0.25),
              'moments': np.full((5, 11), 2.0), 'moments_ci': np.full((5, 11, 2), 3.0)}
    out = oracle.canonical_averages(np.array([0.2, 0.7]), arrays)
    np.testing.assert_allclose(out['max_cluster_size'], 0.5)
    np.testing.assert_allclose(out['moments_ci'], 3.0)
