"""GPU parity of the drop-in API (pypercolate_amd.percolate and .hpc)
against the oracle and the reference's own known answers.  Integer fields
bit-exact; FP64 statistics to the stated tolerances."""
import functools

import networkx as nx
import numpy as np
import pytest
import scipy.stats

import pypercolate_amd as percolate
from oracle import oracle
from pypercolate_amd import hpc
from tests.helpers import grid_3x3_graph

pytestmark = pytest.mark.gpu

MAX_SEED = 4294967296 - 1
SEEDS = [0, 1, 42, 12345, MAX_SEED]


# -- percolate.percolate ---------------------------------------------------
@pytest.mark.parametrize('spanning', [True, False])
def test_sample_states_kat(gpu, golden, spanning):
    """test_percolate.py:144-192 through sample_states (global seed 42)."""
    kat = golden('kat_3x3.json')
    g = grid_3x3_graph(spanning, kat)
    np.random.seed(kat['seed'])
    states = list(percolate.sample_states(g, spanning_cluster=spanning))
    assert [s['n'] for s in states] == list(range(13))
    assert [s['max_cluster_size'] for s in states] == kat['max_cluster_size']
    for s, m in zip(states, kat['moments']):
        assert np.array_equal(s['moments'], m)
    if spanning:
        assert [bool(s['has_spanning_cluster']) for s in states] == \
            kat['has_spanning_cluster']
    assert ('has_spanning_cluster' in states[0]) == spanning
    edges = list(nx.Graph(g).subgraph(range(9)).edges())
    assert states[1]['edge'] in (edges[10], edges[10][::-1])


def test_sample_states_copy_result_false(gpu, golden):
    kat = golden('kat_3x3.json')
    g = grid_3x3_graph(True, kat)
    np.random.seed(42)
    it = percolate.sample_states(g, copy_result=False)
    a = next(it)
    b = next(it)
    assert a is b and b['n'] == 1


def _oracle_micro(g_nx, spanning, seed, runs, alpha):
    og = oracle.percolation_graph_arrays(g_nx, spanning)
    np.random.seed(seed)
    perms = np.stack([np.random.permutation(og['num_edges']) for _ in range(runs)])
    tr = oracle.bond_runs(og, perms)
    if not spanning:
        tr = {k: v for k, v in tr.items() if k != 'has_spanning_cluster'}
    return oracle.microcanonical_arrays(tr, og['num_nodes'], alpha)


@pytest.mark.parametrize('spanning', [True, False])
def test_microcanonical_averages_vs_oracle(gpu, spanning):
    """The L = 16 configuration: 40 runs drawn from the global stream."""
    g = percolate.spanning_2d_grid(16) if spanning else nx.grid_2d_graph(16, 16)
    alpha = percolate.percolate.alpha_1sigma
    np.random.seed(7)
    mine = percolate.microcanonical_averages_arrays(
        percolate.microcanonical_averages(g, runs=40, spanning_cluster=spanning))
    ref = _oracle_micro(g, spanning, 7, 40, alpha)
    for key in ref:
        if key in ('N', 'M'):
            assert mine[key] == ref[key]
        else:
            np.testing.assert_allclose(mine[key], ref[key], rtol=1e-12, atol=0,
                                       err_msg=key)


def test_microcanonical_generator_matches_arrays(gpu, golden):
    g = grid_3x3_graph(True, golden('kat_3x3.json'))
    np.random.seed(3)
    steps = list(percolate.microcanonical_averages(g, runs=9))
    assert [s['n'] for s in steps] == list(range(13))
    np.random.seed(3)
    arr = percolate.microcanonical_averages_arrays(
        percolate.microcanonical_averages(g, runs=9))
    gen = percolate.microcanonical_averages_arrays(iter(steps))
    for key in arr:
        np.testing.assert_allclose(gen[key], arr[key], rtol=1e-15)


def test_microcanonical_initial_kat(gpu, golden):
    """test_percolate.py:234-265."""
    g = grid_3x3_graph(True, golden('kat_3x3.json'))
    ret = next(percolate.microcanonical_averages(g))
    assert ret['n'] == 0 and ret['max_cluster_size'] == 1.0
    np.testing.assert_allclose(ret['moments_ci'], np.ones((5, 2)) * 8)
    assert ret['spanning_cluster'] == 1 / 42
    np.testing.assert_allclose(
        np.array([0, 1]) + np.array([1, -1]) *
        scipy.stats.beta.cdf(ret['spanning_cluster_ci'], a=1, b=41),
        scipy.stats.norm.cdf(-1) * np.ones(2))


@pytest.mark.parametrize('n,p', [(100, 0.1), (480, 0.5), (32512, 0.45),
                                 (1000, 0.01), (10, 0.0), (10, 1.0)])
def test_binomial_pmf(gpu, golden, n, p):
    """test_percolate.py:283-295 and the oracle recursion."""
    got = percolate.percolate._binomial_pmf(n, p)
    assert got.shape == (n + 1,)
    np.testing.assert_allclose(got.sum(), 1.0, rtol=1e-13)
    ref = scipy.stats.binom.pmf(np.arange(n + 1), n, p)
    big = ref > 1e-250
    np.testing.assert_allclose(got[big], ref[big], rtol=1e-7)
    orc = oracle.binomial_pmf(n, p)
    bulk = orc > 1e-12
    np.testing.assert_allclose(got[bulk], orc[bulk], rtol=1e-10)
    assert np.abs(got - orc).max() < 1e-13


def test_canonical_averages_vs_oracle(gpu):
    g = percolate.spanning_2d_grid(16)
    alpha = percolate.percolate.alpha_1sigma
    ref_arrays = _oracle_micro(g, True, 11, 40, alpha)
    ps = np.linspace(0.4, 0.6, 101)
    mine = percolate.canonical_averages(ps, ref_arrays)
    ref = oracle.canonical_averages(ps, ref_arrays)
    for key in ref:
        if key in ('N', 'M', 'ps'):
            continue
        np.testing.assert_allclose(mine[key], ref[key], rtol=1e-11, atol=1e-14,
                                   err_msg=key)


def test_statistics_end_to_end(gpu):
    g = percolate.spanning_2d_grid(8)
    ps = np.linspace(0.3, 0.7, 21)
    np.random.seed(5)
    mine = percolate.statistics(g, ps, runs=12)
    ref = oracle.canonical_averages(
        ps, _oracle_micro(g, True, 5, 12, percolate.percolate.alpha_1sigma))
    for key in ref:
        if key in ('N', 'M', 'ps'):
            continue
        np.testing.assert_allclose(mine[key], ref[key], rtol=1e-11, atol=1e-14)


@pytest.mark.parametrize('spanning', [True, False])
def test_statistics_equals_three_calls(gpu, spanning):
    """statistics keeps the microcanonical statistics on the device; its
    results equal microcanonical_averages -> microcanonical_averages_arrays
    -> canonical_averages bit for bit (same global random stream)."""
    g = percolate.spanning_2d_grid(10)
    ps = np.linspace(0.2, 0.8, 33)
    np.random.seed(11)
    mine = percolate.statistics(g, ps, runs=25, spanning_cluster=spanning)
    np.random.seed(11)
    arrays = percolate.microcanonical_averages_arrays(
        percolate.microcanonical_averages(g, runs=25, spanning_cluster=spanning))
    ref = percolate.canonical_averages(ps, arrays)
    assert sorted(mine) == sorted(ref)
    for key in ref:
        np.testing.assert_array_equal(np.asarray(mine[key]), np.asarray(ref[key]))


def test_single_run_arrays(gpu, golden):
    kat = golden('kat_3x3.json')
    np.random.seed(42)
    ret = percolate.single_run_arrays(graph=grid_3x3_graph(True, kat))
    assert ret['N'] == 9 and ret['M'] == 12
    assert ret['max_cluster_size'].tolist() == kat['max_cluster_size']
    assert ret['moments'].T.tolist() == kat['moments']
    assert ret['has_spanning_cluster'].tolist() == kat['has_spanning_cluster']


# -- percolate.hpc ---------------------------------------------------------
@pytest.fixture(params=[True, False])
def percolation_graph(request):
    """test_hpc.py:144-154."""
    if request.param:
        graph = percolate.spanning_2d_grid(length=3)
    else:
        graph = nx.grid_2d_graph(3, 3)
    return percolate.percolation_graph(graph=graph, spanning_cluster=request.param)


def _oracle_trace(pg, seed):
    og = oracle.percolation_graph_arrays(pg['graph'], pg['spanning_cluster'])
    perm = np.random.RandomState(seed=seed).permutation(og['num_edges'])
    return oracle.bond_runs(og, perm[None, :]), perm


@pytest.mark.parametrize('seed', SEEDS)
def test_bond_sample_states(gpu, percolation_graph, seed):
    pg = percolation_graph
    tr, perm = _oracle_trace(pg, seed)
    seen = None
    for n, st in enumerate(hpc.bond_sample_states(seed=seed, **pg)):
        seen = st if seen is None else seen
        assert st is seen  # one array object, overwritten every step
        assert st['n'][0] == n
        if n:
            assert st['edge'][0] == perm[n - 1]
        assert st['max_cluster_size'][0] == tr['max_cluster_size'][0, n]
        assert st['moments'][0].tolist() == tr['moments'][0, n].tolist()
        if pg['spanning_cluster']:
            assert bool(st['has_spanning_cluster'][0]) == tr['has_spanning_cluster'][0, n]
    assert n == pg['num_edges']


@pytest.mark.parametrize('seed', SEEDS)
def test_bond_microcanonical_statistics(gpu, percolation_graph, seed):
    """test_hpc.py:306-334: rows equal the generator's states."""
    pg = percolation_graph
    res = hpc.bond_microcanonical_statistics(seed=seed, **pg)
    states = hpc.bond_sample_states(seed=seed, **pg)
    assert res.shape == (pg['num_edges'] + 1,)
    for n in range(pg['num_edges'] + 1):
        st = next(states)
        assert res[n]['n'] == n
        if n:
            assert res[n]['edge'] == st['edge'][0]
        assert res[n]['max_cluster_size'] == st['max_cluster_size'][0]
        assert (res[n]['moments'] == st['moments'][0]).all()


def test_bond_microcanonical_batch(gpu):
    pg = percolate.percolation_graph(percolate.spanning_2d_grid(16))
    seeds = list(range(40))
    res = hpc.bond_microcanonical_statistics_batch(seeds=seeds, **pg)
    og = oracle.percolation_graph_arrays(pg['graph'], True)
    tr = oracle.bond_runs(og, oracle.randomstate_permutations(seeds, og['num_edges']))
    np.testing.assert_array_equal(res['max_cluster_size'], tr['max_cluster_size'])
    np.testing.assert_array_equal(res['moments'], tr['moments'])
    np.testing.assert_array_equal(res['has_spanning_cluster'], tr['has_spanning_cluster'])
    np.testing.assert_array_equal(res['edge'][:, 1:], tr['edge'][:, 1:])


@pytest.mark.parametrize('p', [0.0, 0.1, 0.45, 0.5, 0.9, 1.0])
def test_bond_canonical_statistics(gpu, percolation_graph, p):
    """test_hpc.py:337-378 (7 decimals), and the oracle to 1e-12."""
    pg = percolation_graph
    ms = hpc.bond_microcanonical_statistics(seed=99, **pg)
    f = oracle.binomial_pmf(pg['num_edges'], p)
    res = hpc.bond_canonical_statistics(microcanonical_statistics=ms,
                                        convolution_factors=f)
    assert ('percolation_probability' in res.dtype.names) == pg['spanning_cluster']
    ref = oracle.canonical_statistics({k: ms[k] for k in ms.dtype.names}, f)
    if pg['spanning_cluster']:
        np.testing.assert_almost_equal(ref['percolation_probability'],
                                       res['percolation_probability'])
    np.testing.assert_almost_equal(ref['max_cluster_size'], res['max_cluster_size'])
    np.testing.assert_allclose(res['moments'][0], ref['moments'], rtol=1e-12)


def test_bond_pipeline_reduce_finalize(gpu, percolation_graph):
    """test_hpc.py:454-694 through the GPU pieces."""
    pg = percolation_graph
    ps = [0.2, 0.5, 0.8]
    alpha = 0.1
    factors = [oracle.binomial_pmf(pg['num_edges'], p) for p in ps]
    seeds = [3, 17, 256, 1000, 77, 4242, 9, 31, 64, 5]
    inits = []
    vals = []
    for seed in seeds:
        ms = hpc.bond_microcanonical_statistics(seed=seed, **pg)
        cs = np.concatenate([hpc.bond_canonical_statistics(ms, f) for f in factors])
        inits.append(hpc.bond_initialize_canonical_averages(cs))
        vals.append(cs)
    red = functools.reduce(hpc.bond_reduce, inits)
    assert (red['number_of_runs'] == len(seeds)).all()
    stack = np.stack([v['max_cluster_size'] for v in vals])
    np.testing.assert_allclose(red['max_cluster_size_mean'], stack.mean(0), rtol=1e-12)
    np.testing.assert_allclose(red['max_cluster_size_m2'],
                               ((stack - stack.mean(0)) ** 2).sum(0), atol=1e-6)
    fin = hpc.finalize_canonical_averages(pg['num_nodes'], ps, red, alpha)
    np.testing.assert_allclose(fin['percolation_strength_mean'],
                               red['max_cluster_size_mean'] / pg['num_nodes'])
    for bound, a in ((0, alpha / 2), (1, 1 - alpha / 2)):
        ok = fin['percolation_strength_std'] / fin['percolation_strength_mean'] >= 1e-10
        cdf = scipy.stats.t.cdf(fin['percolation_strength_ci'][ok, bound],
                                loc=fin['percolation_strength_mean'][ok],
                                scale=fin['percolation_strength_std'][ok] / np.sqrt(len(seeds)),
                                df=len(seeds) - 1)
        np.testing.assert_allclose(cdf, a, rtol=1e-4)


@pytest.mark.parametrize('L,spanning', [(3, True), (16, True), (16, False)])
def test_bond_canonical_averages_batch(gpu, L, spanning):
    """The jug study's map-reduce on the device equals the oracle's per-run
    canonical statistics reduced over runs (means rtol 1e-11, M2 rtol 1e-9
    of the largest M2 of the field)."""
    graph = percolate.spanning_2d_grid(L) if spanning else nx.grid_2d_graph(L, L)
    pg = percolate.percolation_graph(graph, spanning_cluster=spanning)
    ps = np.linspace(0.3, 0.7, 9)
    seeds = list(range(100, 140))
    res = hpc.bond_canonical_averages_batch(ps=ps, seeds=seeds, run_chunk=16, **pg)
    og = oracle.percolation_graph_arrays(graph, spanning)
    tr = oracle.bond_runs(og, oracle.randomstate_permutations(seeds, og['num_edges']))
    factors = [oracle.binomial_pmf(og['num_edges'], p) for p in ps]
    stats = []
    for r in range(len(seeds)):
        one = {k: v[r] for k, v in tr.items()}
        if not spanning:
            one.pop('has_spanning_cluster')
        stats.append(oracle.initialize_canonical_averages(
            [oracle.canonical_statistics(one, f) for f in factors]))
    ref = functools.reduce(oracle.bond_reduce, stats)
    assert (res['number_of_runs'] == len(seeds)).all()
    for key in ref:
        if key == 'number_of_runs':
            continue
        if key.endswith('_m2'):
            scale = np.abs(ref[key]).max(axis=0, keepdims=True) + 1e-300
            np.testing.assert_allclose(res[key] / scale, ref[key] / scale,
                                       rtol=0, atol=1e-9, err_msg=key)
        else:
            np.testing.assert_allclose(res[key], ref[key], rtol=1e-11,
                                       atol=1e-13, err_msg=key)


@pytest.mark.parametrize('length,runs,spanning', [(16, 40, True), (8, 1, True),
                                                  (12, 7, False), (32, 300, True)])
def test_device_ci_equals_host_epilogue(gpu, length, runs, spanning):
    """pc_microcanonical_ci (the confidence bounds on the device) equals the
    host epilogue (scipy t interval where std > 0, beta quantiles of the
    spanning counts) applied to the same exact means and stds, bit for bit."""
    from pypercolate_amd import _engine as E, _graph
    from pypercolate_amd.percolate import _spanning_stats, _t_interval
    eng = E.BondEngine(_graph.square_lattice(length))
    res = eng.sweep(eng.device_permutations(5, runs))
    s1, s2 = eng.microcanonical_sums(res)
    alpha = 2 * 0.15865525393145707
    out = E.micro_ci(s1, s2, eng.M, runs, alpha).cpu().numpy()
    mean, std = [t.cpu().numpy() for t in E.micro_moments(s1, s2, eng.M, runs)]
    ci_max = _t_interval(mean[0], std[0], runs, alpha)
    assert np.array_equal(out[0], mean[0])
    assert np.array_equal(out[1], ci_max[:, 0]) and np.array_equal(out[2], ci_max[:, 1])
    ci_mom = _t_interval(mean[2:7], std[2:7], runs, alpha)
    for j in range(5):
        assert np.array_equal(out[6 + 3 * j], mean[2 + j])
        assert np.array_equal(out[7 + 3 * j], ci_mom[j, :, 0])
        assert np.array_equal(out[8 + 3 * j], ci_mom[j, :, 1])
    k = s1[1, :, 0].cpu().numpy().view(np.uint64).astype(np.float64)
    span_mean, span_ci = _spanning_stats(k, runs, alpha)
    assert np.array_equal(out[3], span_mean)
    assert np.array_equal(out[4], span_ci[:, 0]) and np.array_equal(out[5], span_ci[:, 1])


@pytest.mark.parametrize('spanning', [True, False])
def test_batch_graph_arrays_prepared(gpu, spanning):
    """Prepared graph_arrays give the same averages as reading the networkx
    graph in the call: the closed form _graph.square_lattice(L) for the
    spanning lattice (same node order, edge order and sides), the arrays
    read once from the graph without spanning (the whole grid, auxiliary
    rows included)."""
    from pypercolate_amd import _graph
    pg = percolate.percolation_graph(percolate.spanning_2d_grid(12),
                                     spanning_cluster=spanning)
    ga = (_graph.square_lattice(12) if spanning
          else _graph.from_percolation_graph(**pg))
    ps = np.linspace(0.35, 0.65, 7)
    seeds = list(range(20))
    a = hpc.bond_canonical_averages_batch(ps=ps, seeds=seeds, **pg)
    b = hpc.bond_canonical_averages_batch(ps=ps, seeds=seeds,
                                          graph_arrays=ga, **pg)
    for key in a.dtype.names:
        assert np.array_equal(a[key], b[key]), key
