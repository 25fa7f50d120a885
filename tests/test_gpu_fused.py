"""Fused canonical averages (weights generated inside the contraction)
against an independent numpy/scipy FP64 reference:
out[p][c] = sum_k w(k) x[k][c] over the window of p, with w the normalised
weights multiplied up from the exact ratio in long double
(tests/helpers.py:ratio_weights, pinned against exact rationals in
tests/test_oracle_canonical.py), not the library's weight code.

Tolerances, relative to sum_k w_k |x_k| (the scale of the sum), floored at
1e-6 of the column's largest such scale (values that only come from weights
near exp(-72) of the row's maximum are compared absolutely at that floor):
  fp64_mfma   1e-12 (ratio recursion + FP64 matrix cores)
  split_bf16  1e-9  (exact byte split; the bar the precision forms must meet)
"""
import numpy as np
import pytest
from pypercolate_amd import _engine as E

from tests.helpers import ratio_weights

pytestmark = pytest.mark.gpu

TOL = {'fp64_mfma': 1e-12, 'split_bf16': 1e-9}


def _reference(x, ps):
    n = x.shape[0] - 1
    start, length = E.binomial_windows(n, ps)
    out = np.zeros((len(ps), x.shape[1]))
    mag = np.zeros_like(out)
    for i, p in enumerate(ps):
        k = np.arange(start[i], start[i] + length[i])
        w = ratio_weights(n, p, k)[:, None]
        out[i] = (w * x[k].astype(np.longdouble)).sum(axis=0)
        mag[i] = (w * np.abs(x[k])).sum(axis=0)
    return out, mag


def _check(gpu, x, ps, precision):
    got = E.canonical_fused(gpu.from_numpy(x).cuda(), ps, precision).cpu().numpy()
    ref, mag = _reference(x, ps)
    floor = 1e-6 * mag.max(axis=0, keepdims=True)
    err = np.abs(got - ref) / np.maximum(mag, np.where(floor > 0, floor, 1.0))
    assert np.isfinite(got).all()
    assert err.max() < TOL[precision], (precision, err.max(), np.unravel_index(
        err.argmax(), err.shape))
    return got


@pytest.mark.parametrize('precision', sorted(TOL))
@pytest.mark.parametrize('S,P,C,lo,hi', [
    (481, 101, 7, 0.4, 0.6),        # L=16 lattice, 101 p
    (32513, 1000, 14, 0.45, 0.55),  # L=128 lattice, 1000 p, 14 columns
    (4001, 37, 21, 0.05, 0.95),     # 21 columns: two column tiles
    (2000, 16, 1, 0.2, 0.8),
])
def test_fused_against_scipy(gpu, precision, S, P, C, lo, hi):
    rng = np.random.RandomState(S + P + C)
    x = rng.rand(S, C) * (10.0 ** (np.arange(C) % 5))[None, :]
    ps = np.linspace(lo, hi, P)
    _check(gpu, x, ps, precision)


@pytest.mark.parametrize('precision', sorted(TOL))
def test_fused_unsorted_and_degenerate_p(gpu, precision):
    rng = np.random.RandomState(7)
    x = rng.rand(3001, 5) + 0.5
    ps = np.array([0.7, 0.0, 0.31, 1.0, 0.5, 1e-4, 0.9999, 0.31, 0.02])
    got = _check(gpu, x, ps, precision)
    # p = 0 and p = 1 put all mass on the first and last step
    tol = 1e-15 if precision == 'fp64_mfma' else TOL[precision]
    np.testing.assert_allclose(got[1], x[0], rtol=tol)
    np.testing.assert_allclose(got[3], x[-1], rtol=tol)


@pytest.mark.parametrize('precision', sorted(TOL))
def test_fused_wide_range_and_signs(gpu, precision):
    """Columns spanning 1e-300..2^154 and a mixed-sign column: the split's
    per-window column scales keep them exact where FP32 would overflow."""
    rng = np.random.RandomState(11)
    S = 20001
    t = np.arange(S) / (S - 1)
    x = np.stack([
        2.0 ** 154 * np.exp(-((t - 0.5) * 40) ** 2) + 1e20,
        1e-300 * (1 + rng.rand(S)),
        np.where(t < 0.5, 0.0, 1.0),
        rng.rand(S) - 0.2,
    ], axis=1)
    ps = np.linspace(0.3, 0.7, 64)
    _check(gpu, x, ps, precision)


@pytest.mark.parametrize('precision', sorted(TOL))
def test_fused_small_n(gpu, precision):
    for S in (1, 2, 3, 10):
        x = np.arange(1.0, S + 1.0)[:, None] * np.array([[1.0, 2.0, -1.0]])
        ps = np.linspace(0.0, 1.0, 11)
        _check(gpu, x, ps, precision)


def test_fused_long_walk(gpu):
    """Windows longer than 64 x 1024 steps at n = 4e7: every lane's walk
    passes the periodic direct re-evaluation of the weights."""
    torch = gpu
    n = 40_000_000
    g = torch.Generator(device='cuda').manual_seed(3)
    x = torch.rand((n + 1, 2), dtype=torch.float64, device='cuda', generator=g)
    ps = np.array([0.5, 0.3])
    got = E.canonical_fused(x, ps, 'fp64_mfma').cpu().numpy()
    start, length = E.binomial_windows(n, ps)
    assert length.min() > 64 * 1024
    for i, p in enumerate(ps):
        k = np.arange(start[i], start[i] + length[i])
        w = ratio_weights(n, p, k)
        xs = x[int(start[i]):int(start[i] + length[i])].cpu().numpy()
        ref = (w[:, None] * xs.astype(np.longdouble)).sum(axis=0)
        err = np.abs(got[i] - ref) / (w[:, None] * np.abs(xs)).sum(axis=0)
        assert float(err.max()) < 1e-12, (p, err)


def test_fused_deterministic_and_matches_unfused(gpu):
    rng = np.random.RandomState(5)
    x = rng.rand(8001, 14)
    ps = np.linspace(0.3, 0.7, 64)
    dx = gpu.from_numpy(x).cuda()
    for prec in sorted(TOL):
        plan = E.FusedPlan(dx, ps, prec)
        a = plan.run().cpu().numpy().copy()
        b = plan.run().cpu().numpy()
        assert np.array_equal(a, b)
    unfused = E.contract(dx, E.binomial_weights(8000, ps), 'fp64_mfma').cpu().numpy()
    fused = E.canonical_fused(dx, ps, 'fp64_mfma').cpu().numpy()
    np.testing.assert_allclose(fused, unfused, rtol=1e-12)
