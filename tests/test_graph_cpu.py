"""Host graph ingestion: closed-form lattices equal the networkx ingestion,
and the per-node side flags reproduce the oracle's auxiliary union-find
(percolate/hpc.py:223-243,269-274) on ordinary and unusual graphs."""
import networkx as nx
import numpy as np
import pytest

from oracle import oracle
from pypercolate_amd import _graph, percolate
from tests.helpers import spanning_1d_chain, spanning_2d_grid


@pytest.mark.parametrize('L', [1, 2, 3, 5, 16])
def test_square_lattice_closed_form(L):
    pg = percolate.percolation_graph(percolate.spanning_2d_grid(L))
    a = _graph.from_percolation_graph(**pg)
    b = _graph.square_lattice(L)
    assert a.num_nodes == b.num_nodes == L * L
    np.testing.assert_array_equal(a.edges, b.edges)
    np.testing.assert_array_equal(a.node_sides, b.node_sides)
    assert a.span_on_first_merge == b.span_on_first_merge
    assert a.nodes == b.nodes


@pytest.mark.parametrize('n', [1, 2, 10])
def test_linear_chain_closed_form(n):
    pg = percolate.percolation_graph(percolate.spanning_1d_chain(n))
    a = _graph.from_percolation_graph(**pg)
    b = _graph.linear_chain(n)
    np.testing.assert_array_equal(a.edges, b.edges)
    np.testing.assert_array_equal(a.node_sides, b.node_sides)
    assert a.span_on_first_merge == b.span_on_first_merge


def _spanning_from_flags(ga, perm):
    """Spanning per step from the side flags alone (CPU union-find of the
    perc nodes, sides OR-ed at merges, evaluated at merge steps only)."""
    parent = list(range(ga.num_nodes))
    side = [int(x) for x in ga.node_sides]

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x
    out = [False]
    span = False
    for e in perm:
        u, v = (int(t) for t in ga.edges[e])
        ru, rv = find(u), find(v)
        if ru != rv:
            parent[ru] = rv
            side[rv] |= side[ru]
            span = span or ga.span_on_first_merge or side[rv] == 3
        out.append(span)
    return np.array(out)


def _odd_graphs():
    # (1) perc node 4 touches both sides
    g1 = nx.path_graph(8)
    for a, s in ((100, 0), (101, 1)):
        g1.add_node(a, span=s)
    g1.add_edge(100, 4, span=0)
    g1.add_edge(101, 4, span=1)
    g1.add_edge(101, 7, span=1)
    # (2) an auxiliary edge joins the two sides directly
    g2 = nx.cycle_graph(6)
    g2.add_node('a', span=0)
    g2.add_node('b', span=1)
    g2.add_edge('a', 'b', span=0)
    g2.add_edge('a', 0, span=0)
    # (3) several aux nodes per side, sides on random nodes
    g3 = nx.gnm_random_graph(40, 90, seed=5)
    for j in range(4):
        g3.add_node(('L', j), span='left')
        g3.add_node(('R', j), span='right')
        g3.add_edge(('L', j), j, span='left')
        g3.add_edge(('R', j), 39 - j, span='right')
    return [g1, g2, g3, spanning_2d_grid(4), spanning_1d_chain(6)]


@pytest.mark.parametrize('which', range(5))
def test_side_flags_match_oracle(which):
    g = _odd_graphs()[which]
    pg = percolate.percolation_graph(g)
    ga = _graph.from_percolation_graph(**pg)
    og = oracle.percolation_graph_arrays(g, True)
    np.testing.assert_array_equal(ga.edges, og['edges'])
    perms = oracle.randomstate_permutations(range(6), ga.num_edges)
    tr = oracle.bond_runs(og, perms)
    for r in range(perms.shape[0]):
        np.testing.assert_array_equal(_spanning_from_flags(ga, perms[r]),
                                      tr['has_spanning_cluster'][r])


def test_graph_arrays_validation():
    with pytest.raises(ValueError):
        _graph.GraphArrays(3, [[0, 3]])
    with pytest.raises(ValueError):
        _graph.GraphArrays(0, np.zeros((0, 2)))
    with pytest.raises(ValueError):
        _graph.GraphArrays(2, [[0, 1]], node_sides=[0, 3])


def test_locality_order_tiles_grid_nodes():
    """locality_order: a permutation of the nodes that puts 4 x 8 blocks of
    grid nodes at consecutive indices; None for labels that are not pairs
    of integers."""
    from pypercolate_amd import _graph
    g = _graph.square_lattice(20)
    order = _graph.locality_order(g)
    assert sorted(order.tolist()) == list(range(400))
    coords = np.array(g.nodes)
    # every aligned 4 x 8 block of the grid is one run of 32 indices
    for i0 in range(0, 20, 4):
        for j0 in (0, 8):
            sel = ((coords[:, 0] - 1) // 4 == i0 // 4) & (coords[:, 1] // 8 == j0 // 8)
            idx = np.sort(order[sel])
            assert (np.diff(idx) == 1).all()
    assert _graph.locality_order(_graph.linear_chain(10)) is None
