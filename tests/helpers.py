"""Graph builders shared by the tests (networkx 3 API)."""
import networkx as nx


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
