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
    """The oracle's graph dict of spanning_2d_grid(length) in closed form
    (oracle.square_lattice_graph, pinned to networkx on small lattices by
    tests/test_oracle_exact.py)."""
    from oracle import oracle
    return oracle.square_lattice_graph(length)
