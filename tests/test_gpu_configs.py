"""The finite-size batch and the large canonical contraction at the sizes
of the benchmark configurations.

* Finite-size batch: L in {16, 32, 64, 128, 256, 512}, 2^26 / L^2 runs per
  L (262144 ... 256), Philox permutations; the HIP-graph replays equal the
  eager per-L pipelines bit for bit, and every L's exact per-step sums over
  all its runs equal the oracle's over the same (device-drawn)
  permutations; the persistent form (one sweep launch over all sizes)
  gives the same results.  The reference runs this study one size at a time
  (percolate/share/jugfile.py:166-259).
* Canonical averages at L = 1024: M = 2,095,104, 4096 p in [0.3, 0.7], the
  14 columns of the benchmark (mean and mean square of the seven
  observables, the fourth moment's mean square past the FP32 range), fused FP64 and
  exact-split forms against long-double ratio weights on 64 sampled p
  (percolate/percolate.py:1196-1221)."""
import numpy as np
import pytest

from oracle import oracle
from pypercolate_amd import _graph
from pypercolate_amd import _engine as E
from pypercolate_amd.batch import FiniteSizeBatch, default_runs
from tests.helpers import lattice_oracle_graph, ratio_weights

pytestmark = pytest.mark.gpu

LENGTHS = [16, 32, 64, 128, 256, 512]


def test_finite_size_batch_at_size(gpu):
    torch = gpu
    runs = default_runs(LENGTHS)
    assert runs == [262144, 65536, 16384, 4096, 1024, 256]
    ps = np.linspace(0.4, 0.6, 21)
    batch = FiniteSizeBatch(LENGTHS, runs, seed=42, ps=ps, tune=False)
    batch.run_sequential()
    torch.cuda.synchronize()
    eager = batch.results()
    batch.capture()
    for k in range(2):
        batch.run()
        torch.cuda.synchronize()
        junk = torch.ones(1 << 20, device='cuda').sum().item()
        assert junk == float(1 << 20)
        got = batch.results()
        for L in LENGTHS:
            for key in ('sum1', 'sum2', 'mean', 'canonical'):
                assert np.array_equal(got[L][key], eager[L][key]), (k, L, key)
    # one persistent sweep launch over all six lattices: the same sums,
    # means and canonical averages
    pers = FiniteSizeBatch(LENGTHS, runs, seed=42, ps=ps, persistent=True)
    pers.run()
    torch.cuda.synchronize()
    got = pers.results()
    for L in LENGTHS:
        for key in ('sum1', 'sum2', 'mean', 'canonical'):
            assert np.array_equal(got[L][key], eager[L][key]), ('persistent', L, key)
    del pers, got
    for plan in batch.plans:
        L, R = plan.length, plan.runs
        g = lattice_oracle_graph(L)
        sums = None
        step = max(1, (1 << 24) // g['num_edges'])
        for r0 in range(0, R, step):
            perms = plan.perm[r0:r0 + step].cpu().numpy()
            sums = oracle.bond_sums(g, perms, sums=sums, threads=16)
        assert np.array_equal(eager[L]['sum1'], sums[0]), L
        assert np.array_equal(eager[L]['sum2'], sums[1]), L
        # the per-step means and their canonical averages at ps
        s1 = oracle.limbs_to_int(sums[0])
        mean = np.array([[float(v) / R for v in row] for row in s1])
        np.testing.assert_allclose(eager[L]['mean'], mean, rtol=1e-15, atol=0)
        M = g['num_edges']
        start, length = E.binomial_windows(M, ps)
        ref = np.zeros((ps.size, 7))
        for i, p in enumerate(ps):
            k = np.arange(start[i], start[i] + length[i])
            ref[i] = (ratio_weights(M, p, k)[:, None] *
                      mean[:, k].T.astype(np.longdouble)).sum(axis=0)
        np.testing.assert_allclose(eager[L]['canonical'], ref, rtol=1e-11,
                                   atol=1e-12 * np.abs(ref).max())


def _columns_L1024(torch, runs=64):
    """The 14 columns of the benchmark's canonical comparison from `runs`
    runs of the L = 1024 lattice: per step the mean and the mean square of
    each observable (mean square = std^2 (R-1)/R + mean^2)."""
    eng = E.BondEngine(_graph.square_lattice(1024))
    res = eng.sweep(eng.device_permutations(7, runs))
    s1, s2 = eng.microcanonical_sums(res)
    mean, std = E.micro_moments(s1, s2, eng.M, runs)
    second = std * std * ((runs - 1.0) / runs) + mean * mean
    return torch.cat([mean, second], dim=0).t().contiguous()


@pytest.mark.parametrize('precision,tol', [('fp64_mfma', 1e-12), ('split_bf16', 1e-9)])
def test_canonical_L1024_4096p_14_columns(gpu, precision, tol):
    torch = gpu
    x = _columns_L1024(torch)
    n = x.shape[0] - 1
    assert n == 2095104 and x.shape[1] == 14
    ps = np.linspace(0.3, 0.7, 4096)
    got = E.canonical_fused(x, ps, precision).cpu().numpy()
    assert np.isfinite(got).all()
    xh = x.cpu().numpy()
    assert xh[:, 13].max() > 2.0 ** 128   # past the FP32 range
    start, length = E.binomial_windows(n, ps)
    rows = np.linspace(0, ps.size - 1, 64).astype(int)
    for i in rows:
        k = np.arange(start[i], start[i] + length[i])
        w = ratio_weights(n, ps[i], k)[:, None]
        xs = xh[k].astype(np.longdouble)
        ref = (w * xs).sum(axis=0)
        mag = (w * np.abs(xs)).sum(axis=0)
        err = np.abs(got[i] - ref) / np.maximum(mag, 1e-300)
        assert float(err.max()) < tol, (precision, ps[i], err.max())
