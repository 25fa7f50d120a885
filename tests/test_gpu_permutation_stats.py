"""The device permutations (pc_permute_philox) give the same physics as
numpy's RandomState permutations the reference draws (percolate/hpc.py:
195,206): per-step means of all seven observables (largest cluster,
spanning, moments 0..4), from the exact sums over many runs of each kind,
agree within their standard errors (Welch z, Bonferroni bound over every
non-constant (column, step) pair)."""
import numpy as np
import pytest
import scipy.stats

from pypercolate_amd import _graph
from pypercolate_amd._engine import BondEngine, Pipeline, host_sums

pytestmark = pytest.mark.gpu


def _moments(s1, s2, R):
    """Means and variances of the mean (float) from exact Python-int sums."""
    mean = np.array([[int(v) / R for v in row] for row in s1])
    var = np.array([[(R * int(b) - int(a) * int(a)) / (R * R * (R - 1.0))
                     for a, b in zip(ra, rb)] for ra, rb in zip(s1, s2)])
    return mean, var


@pytest.mark.parametrize('length,runs', [(16, 200000), (128, 10000)])
def test_observables_match_randomstate(gpu, length, runs):
    torch = gpu
    eng = BondEngine(_graph.square_lattice(length))
    a = Pipeline(eng, runs, seed=20241203)
    sa = host_sums(*a.run())
    torch.cuda.synchronize()
    del a
    b = Pipeline(eng, runs, seeds=np.arange(runs) + 7)
    sb = host_sums(*b.run())
    torch.cuda.synchronize()
    del b
    ma, va = _moments(sa[0], sa[1], runs)
    mb, vb = _moments(sb[0], sb[1], runs)
    both0 = (va == 0) & (vb == 0)
    # constant in both samples: the same constant (e.g. step 0, step M)
    assert np.array_equal(ma[both0], mb[both0])
    live = ~both0
    z = (ma[live] - mb[live]) / np.sqrt(va[live] + vb[live])
    tests = int(live.sum())
    bound = scipy.stats.norm.isf(0.001 / 2 / tests)
    worst = np.abs(z).max()
    assert worst < bound, (worst, bound, tests)
    # and the z values look standard normal overall (they are correlated
    # along the steps, so only a loose check of their spread)
    assert 0.5 < np.std(z) < 2.0
