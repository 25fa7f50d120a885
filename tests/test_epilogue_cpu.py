"""Host epilogues of the drop-in (no GPU): bond_initialize and bond_reduce
(hpc.py:561-702) and the vectorised microcanonical confidence intervals
(percolate.py:557-705) against the oracle's per-step restatement.  Finalize
runs on the device (tests/test_gpu_finalize.py)."""
import functools

import numpy as np
import pytest
import scipy.stats

from oracle import oracle
from pypercolate_amd import hpc
from pypercolate_amd import percolate as P


def _canonical_rows(vals, spanning=True):
    st = np.empty(vals.shape[1], dtype=hpc.canonical_statistics_dtype(spanning))
    if spanning:
        st['percolation_probability'] = vals[0]
    st['max_cluster_size'] = vals[1]
    st['moments'] = vals[2:7].T
    return st


@pytest.mark.parametrize('spanning', [True, False])
def test_reduce_vs_oracle(spanning):
    rng = np.random.RandomState(0)
    runs = [rng.rand(7, 4) * 50 + 1 for _ in range(13)]
    inits = [hpc.bond_initialize_canonical_averages(_canonical_rows(v, spanning))
             for v in runs]
    for a in inits:
        assert (a['number_of_runs'] == 1).all()
        assert (a['max_cluster_size_m2'] == 0).all()
    res = functools.reduce(hpc.bond_reduce, inits)
    stack = np.stack(runs)
    np.testing.assert_allclose(res['max_cluster_size_mean'], stack[:, 1].mean(0),
                               rtol=1e-12)
    np.testing.assert_allclose(res['moments_m2'],
                               ((stack[:, 2:7] - stack[:, 2:7].mean(0)) ** 2)
                               .sum(0).T, rtol=1e-10)
    odicts = [{k: a[k] for k in a.dtype.names} for a in inits]
    ores = functools.reduce(oracle.bond_reduce, odicts)
    for k in res.dtype.names:
        np.testing.assert_allclose(res[k], ores[k], rtol=1e-12)


def test_microcanonical_ci_vectorised():
    rng = np.random.RandomState(2)
    runs, alpha = 25, 0.2
    mx = rng.randint(1, 50, size=(runs, 30)).astype(float)
    mx[:, 3] = 7.0  # std 0 -> collapsed interval
    mom = rng.rand(runs, 30, 5) * 100
    span = rng.rand(runs, 30) < 0.4
    ci = P._t_interval(mx.mean(0), mx.std(0, ddof=1), runs, alpha)
    mci = P._t_interval(mom.mean(0).T, mom.std(0, ddof=1).T, runs, alpha)
    smean, sci = P._spanning_stats(span.sum(0), runs, alpha)
    for s in range(30):
        o = oracle.microcanonical_step_stats(mx[:, s], mom[:, s], span[:, s], alpha)
        np.testing.assert_allclose(ci[s], o['max_cluster_size_ci'], rtol=1e-13)
        np.testing.assert_allclose(mci[:, s], o['moments_ci'], rtol=1e-13)
        assert smean[s] == o['spanning_cluster']
        np.testing.assert_allclose(sci[s], o['spanning_cluster_ci'], rtol=1e-13)
