"""Host-side behaviour of the drop-in API that needs no GPU: names and
signatures (test_percolate.py:31-39,195-203), structured dtypes
(test_hpc.py:25-141) and the ValueError paths raised at the first next()
(test_percolate.py:76-89,206-231, test_hpc.py:295-303)."""
import inspect

import networkx as nx
import numpy as np
import pytest

import pypercolate_amd as percolate
from pypercolate_amd import hpc
from tests.helpers import grid_3x3_graph


@pytest.mark.parametrize('name', [
    'sample_states', 'single_run_arrays', 'microcanonical_averages',
    'microcanonical_averages_arrays', 'canonical_averages',
    'spanning_1d_chain', 'spanning_2d_grid', 'statistics'])
def test_exports(name):
    assert hasattr(percolate, name)


@pytest.mark.parametrize('func,params', [
    (percolate.sample_states, ['graph', 'spanning_cluster', 'model']),
    (percolate.microcanonical_averages,
     ['graph', 'runs', 'spanning_cluster', 'model', 'alpha']),
    (hpc.bond_sample_states, ['perc_graph', 'num_nodes', 'num_edges', 'seed',
                              'spanning_cluster', 'auxiliary_node_attributes',
                              'auxiliary_edge_attributes', 'spanning_sides']),
    (hpc.bond_microcanonical_statistics, ['perc_graph', 'num_nodes',
                                          'num_edges', 'seed']),
    (hpc.bond_canonical_statistics, ['microcanonical_statistics',
                                     'convolution_factors']),
    (hpc.bond_initialize_canonical_averages, ['canonical_statistics']),
    (hpc.bond_reduce, ['row_a', 'row_b']),
    (hpc.finalize_canonical_averages, ['number_of_nodes', 'ps',
                                       'canonical_averages', 'alpha']),
])
def test_signatures(func, params):
    args = inspect.signature(func).parameters
    for p in params:
        assert p in args


@pytest.mark.parametrize('spanning', [True, False])
@pytest.mark.parametrize('fn,field', [
    (hpc.microcanonical_statistics_dtype, 'has_spanning_cluster'),
    (hpc.canonical_statistics_dtype, 'percolation_probability'),
    (hpc.canonical_averages_dtype, 'percolation_probability_mean'),
    (hpc.finalized_canonical_averages_dtype, 'percolation_probability_ci'),
])
def test_dtypes(spanning, fn, field):
    a = np.empty(1, dtype=fn(spanning_cluster=spanning))
    assert (field in a.dtype.names) == spanning


def test_microcanonical_dtype_layout():
    dt = np.dtype(hpc.microcanonical_statistics_dtype(True))
    assert dt.names == ('n', 'edge', 'has_spanning_cluster', 'max_cluster_size',
                        'moments')
    assert dt['moments'].shape == (5,) and dt['moments'].base == np.uint64
    fdt = np.dtype(hpc.finalized_canonical_averages_dtype(True))
    assert fdt['moments_ci'].shape == (5, 2)


def test_sample_states_bad_model():
    with pytest.raises(ValueError):
        next(percolate.sample_states(nx.Graph(), model='site'))


def test_sample_states_no_aux():
    with pytest.raises(ValueError):
        next(percolate.sample_states(nx.Graph(), spanning_cluster=True))


def test_sample_states_one_sided():
    g = nx.Graph()
    g.add_node(1, span=0)
    with pytest.raises(ValueError):
        next(percolate.sample_states(g, spanning_cluster=True))


@pytest.mark.parametrize('kwargs', [{'runs': 'many'}, {'runs': 0},
                                    {'alpha': 'huge'}, {'alpha': 0.0},
                                    {'alpha': 1.0}])
def test_microcanonical_bad_args(golden, kwargs):
    g = grid_3x3_graph(True, golden('kat_3x3.json'))
    it = percolate.microcanonical_averages(g, **kwargs)
    with pytest.raises(ValueError):
        next(it)


def test_bond_sample_states_three_sides():
    pg = percolate.percolation_graph(percolate.spanning_2d_grid(3))
    pg['spanning_sides'] = [0, 1, 2]
    with pytest.raises(ValueError):
        next(hpc.bond_sample_states(seed=0, **pg))


def test_bond_sample_states_first_state_without_gpu():
    """n = 0 comes before any device work (hpc.py:208-221)."""
    pg = percolate.percolation_graph(percolate.spanning_2d_grid(3))
    st = next(hpc.bond_sample_states(seed=3, **pg))
    assert st['n'][0] == 0 and st['max_cluster_size'][0] == 1
    assert st['moments'][0].tolist() == [8] * 5
    assert not st['has_spanning_cluster'][0]


def test_percolation_graph_keys():
    pg = percolate.percolation_graph(percolate.spanning_2d_grid(4))
    assert pg['num_nodes'] == 16 and pg['num_edges'] == 24
    assert sorted(pg['spanning_sides']) == [0, 1]
    assert len(pg['auxiliary_edge_attributes']) == 8
    pg2 = percolate.percolation_graph(nx.grid_2d_graph(3, 3), spanning_cluster=False)
    assert pg2['num_edges'] == 12 and 'spanning_sides' not in pg2


def test_batch_graph_arrays_must_match_spanning():
    """graph_arrays without sides for a spanning-cluster call (or a graph of
    another size) is a ValueError, raised before any GPU work."""
    from pypercolate_amd import _graph
    pg = percolate.percolation_graph(percolate.spanning_2d_grid(4),
                                     spanning_cluster=True)
    with pytest.raises(ValueError):
        hpc.bond_canonical_averages_batch(
            ps=[0.5], num_runs=1,
            graph_arrays=_graph.square_lattice(4, spanning_cluster=False), **pg)
    with pytest.raises(ValueError):
        hpc.bond_canonical_averages_batch(
            ps=[0.5], num_runs=1, graph_arrays=_graph.square_lattice(5), **pg)


def test_arrays_from_a_plain_iterator_of_steps():
    """microcanonical_averages_arrays over any iterator of per-step dicts
    (percolate/percolate.py:968-1064): each entry stacked over the steps,
    moments with their 5 rows first, all but spanning divided by N."""
    from pypercolate_amd import percolate
    rng = np.random.RandomState(0)
    N, M = 9, 12
    steps = []
    for n in range(M + 1):
        steps.append({'n': n, 'N': N, 'M': M,
                      'max_cluster_size': rng.rand(),
                      'max_cluster_size_ci': rng.rand(2),
                      'spanning_cluster': rng.rand(),
                      'spanning_cluster_ci': rng.rand(2),
                      'moments': rng.rand(5), 'moments_ci': rng.rand(5, 2)})
    out = percolate.microcanonical_averages_arrays(iter(steps))
    assert out['N'] == N and out['M'] == M
    for n, st in enumerate(steps):
        assert out['max_cluster_size'][n] == st['max_cluster_size'] / N
        assert (out['max_cluster_size_ci'][n] == st['max_cluster_size_ci'] / N).all()
        assert out['spanning_cluster'][n] == st['spanning_cluster']
        assert (out['spanning_cluster_ci'][n] == st['spanning_cluster_ci']).all()
        assert (out['moments'][:, n] == st['moments'] / N).all()
        assert (out['moments_ci'][:, n] == st['moments_ci'] / N).all()
    assert out['moments'].shape == (5, M + 1)
    assert out['moments_ci'].shape == (5, M + 1, 2)
    with pytest.raises(AssertionError):
        percolate.microcanonical_averages_arrays(iter(steps[1:]))
