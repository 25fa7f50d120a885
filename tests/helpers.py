"""Graph builders and reference weights shared by the tests."""
import networkx as nx
import numpy as np


def spanning_2d_grid(length):
    """percolate/percolate.py:931-965 with networkx 3 attribute access."""
    g = nx.grid_2d_graph(length + 2, length)
    for i in range(length):
        g.nodes[(0, i)]['span'] = 0
        g[(0, i)][(1, i)]['span'] = 0
        g.nodes[(length + 1, i)]['span'] = 1
        g[(length + 1, i)][(length, i)]['span'] = 1
    return g


def spanning_1d_chain(length):
    """percolate/percolate.py:899-928 with networkx 3 attribute access."""
    g = nx.grid_graph(dim=[int(length + 2)])
    g.nodes[0]['span'] = 0
    g[0][1]['span'] = 0
    g.nodes[length + 1]['span'] = 1
    g[length][length + 1]['span'] = 1
    return g


def grid_3x3_graph(spanning, kat):
    """The fixture of percolate/test/test_percolate.py:49-73."""
    g = nx.Graph()
    g.add_nodes_from(range(9))
    g.add_edges_from(kat['edges_added_by_fixture'])
    if spanning:
        g.add_nodes_from(range(9, 12), span=0)
        g.add_nodes_from(range(12, 15), span=1)
        g.add_edges_from(kat['aux_side0_edges'], span=0)
        g.add_edges_from(kat['aux_side1_edges'], span=1)
    return g


def ratio_weights(n, p, k):
    """Normalised binomial weights over the consecutive steps k (an integer
    array) from the exact ratio B(k) / B(k - 1) = (n - k + 1) / k * p / (1 - p),
    multiplied up in 80-bit long double: relative error below len(k) * 1e-19,
    independent of any pmf routine (scipy.stats.binom.pmf is off by ~1e-12
    at n ~ 3e4).  p = 0 and p = 1 put all mass on k = 0 and k = n."""
    k = np.asarray(k)
    if p <= 0.0 or p >= 1.0:
        w = (k == (0 if p <= 0.0 else n)).astype(np.longdouble)
        return w / w.sum()
    ld = np.longdouble
    kk = k[1:].astype(ld)
    r = (ld(n) - kk + 1) / kk * (ld(p) / (1 - ld(p)))
    w = np.concatenate([np.ones(1, dtype=ld), np.cumprod(r)])
    return w / w.sum()


def lattice_oracle_graph(length):
    """The oracle's graph dict of spanning_2d_grid(length), built from the
    closed-form lattice (whose edge order tests/test_graph_cpu.py pins to
    networkx's) instead of a networkx graph, which takes minutes and GBs at
    L = 1024.  Auxiliary nodes: one per column on each side, attached to the
    first and last row, as percolate/percolate.py:931-965 builds them."""
    from pypercolate_amd import _graph
    ga = _graph.square_lattice(length)
    L, N = int(length), ga.num_nodes
    return {'num_nodes': N, 'num_edges': ga.num_edges, 'spanning': True,
            'edges': ga.edges.astype(np.int64),
            'aux_side': np.array([0] * L + [1] * L, dtype=np.int32),
            'aux_edges': np.array([(i, N + i) for i in range(L)] +
                                  [(N - L + i, N + L + i) for i in range(L)],
                                  dtype=np.int64).reshape(-1, 2),
            'aux_edge_side': np.array([0] * L + [1] * L, dtype=np.int32)}
