"""GPU contraction at every precision against numpy FP64 (and so against
each other): Q[p][c] = sum_k w[p][k] X[start_p + k][c].  Tolerances:
FP64 forms 1e-13 relative; the FP32-input forms 2e-6 relative to the sum of
|w X| (FP32 inputs round at 2^-24)."""
import numpy as np
import pytest

from pypercolate_amd import _engine as E
from pypercolate_amd import _native as nat

pytestmark = pytest.mark.gpu

TOL = {'fp64': 1e-13, 'fp64_mfma': 1e-13, 'fp32_mfma': 2e-6, 'bf16x3_mfma': 2e-6}


def _numpy_contract(x, wt):
    w = wt.w.cpu().numpy()
    out = np.zeros((w.shape[0], x.shape[1]))
    mag = np.zeros_like(out)
    for i in range(w.shape[0]):
        s, L = int(wt.host_start[i]), int(wt.host_length[i])
        out[i] = w[i, :L] @ x[s:s + L]
        mag[i] = w[i, :L] @ np.abs(x[s:s + L])
    return out, mag


@pytest.mark.parametrize('precision', sorted(TOL))
@pytest.mark.parametrize('S,P,C', [(481, 101, 7), (4001, 37, 14), (32513, 40, 21),
                                   (2000, 16, 1)])
def test_contract_precisions(gpu, precision, S, P, C):
    rng = np.random.RandomState(S + P + C)
    # column c scaled by 10^(c mod 5): asymmetric in p and c, wide range
    x = rng.rand(S, C) * (10.0 ** (np.arange(C) % 5))[None, :]
    x[:, 0] -= 0.3  # one column with mixed signs
    ps = np.sort(rng.uniform(0.05, 0.95, size=P))
    wt = E.binomial_weights(S - 1, ps)
    dx = gpu.from_numpy(x).cuda()
    got = E.contract(dx, wt, nat.PRECISIONS[precision]).cpu().numpy()
    ref, mag = _numpy_contract(x, wt)
    err = np.abs(got - ref) / mag
    assert err.max() < TOL[precision], (precision, err.max())


def test_contract_deterministic(gpu):
    rng = np.random.RandomState(0)
    x = rng.rand(8001, 14)
    wt = E.binomial_weights(8000, np.linspace(0.3, 0.7, 64))
    dx = gpu.from_numpy(x).cuda()
    for prec in ('fp64_mfma', 'fp32_mfma', 'bf16x3_mfma'):
        plan = E.ContractPlan(dx, wt, nat.PRECISIONS[prec])
        a = plan.run().cpu().numpy().copy()
        b = plan.run().cpu().numpy()
        assert np.array_equal(a, b)


@pytest.mark.parametrize('S,P,C,lo,hi', [(481, 101, 70, 0.05, 0.95),
                                         (8001, 130, 1100, 0.2, 0.8),
                                         (4001, 1000, 4160, 0.45, 0.55)])
def test_contract_wide_fp64_mfma(gpu, S, P, C, lo, hi):
    """The wide FP64 matrix-core kernel (64 x 64 tiles through LDS, from 64
    columns on): with K-slices (few tiles) and without (>= 1024 tiles),
    partial tiles in p and in columns."""
    rng = np.random.RandomState(S + P)
    x = rng.rand(S, C) * (10.0 ** (np.arange(C) % 5))[None, :]
    x[:, 1] -= 0.5
    ps = np.sort(rng.uniform(lo, hi, size=P))
    wt = E.binomial_weights(S - 1, ps)
    dx = gpu.from_numpy(x).cuda()
    plan = E.ContractPlan(dx, wt, 'fp64_mfma')
    got = plan.run().cpu().numpy().copy()
    ref, mag = _numpy_contract(x, wt)
    err = np.abs(got - ref) / mag
    assert err.max() < 1e-13, err.max()
    assert np.array_equal(got, plan.run().cpu().numpy())


@pytest.mark.parametrize('R,cols', [(1, 7), (255, 6), (1000, 7), (4099, 8)])
def test_run_moments(gpu, R, cols):
    """pc_run_moments: mean and sum of squared deviations over runs of
    q[p][r * cols + c], against numpy (one workgroup per p, runs strided
    over 256 threads)."""
    rng = np.random.RandomState(R)
    P = 9
    q = rng.rand(P, R, cols) * 100.0 + rng.rand(P, 1, cols) * 1e4
    dq = gpu.from_numpy(q.reshape(P, R * cols)).cuda()
    mean = gpu.empty((P, cols), dtype=gpu.float64, device='cuda')
    m2 = gpu.empty_like(mean)
    lib = nat.load_library()
    nat.check(lib.pc_run_moments(nat.ptr(dq), P, R, cols, nat.ptr(mean),
                                 nat.ptr(m2), nat.stream_handle(gpu)),
              'pc_run_moments')
    mu = q.mean(axis=1)
    np.testing.assert_allclose(mean.cpu().numpy(), mu, rtol=1e-13)
    d2 = ((q - mu[:, None, :]) ** 2).sum(axis=1)
    np.testing.assert_allclose(m2.cpu().numpy(), d2, rtol=1e-10, atol=1e-12)
