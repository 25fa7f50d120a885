"""Host-side graph ingestion: a percolation graph as flat integer arrays.

The device kernels see a graph as
  * ``edges``: int32 array (M, 2) of node indices, in ``perc_graph.edges()``
    order (the order the reference indexes with its permutation,
    percolate/hpc.py:205,255);
  * ``node_sides``: uint8 array (N,), which spanning side (1 or 2, 0 none)
    each node is attached to through the auxiliary nodes and edges;
  * ``span_on_first_merge``: True when the auxiliary structure alone joins
    both sides.

``node_sides`` restates the reference's second union-find over the
auxiliary nodes (percolate/hpc.py:223-243): all auxiliary nodes of a side
are merged, every auxiliary edge is merged with its side's root, and a
percolation cluster spans once it contains nodes attached to both sides
(hpc.py:269-274).  The reference evaluates spanning only on merge steps, so
a pre-joined structure reports spanning from the first merge on.
"""

import numpy as np


class GraphArrays(object):
    """Integer form of a percolation graph (see module docstring)."""

    def __init__(self, num_nodes, edges, node_sides=None,
                 span_on_first_merge=False, nodes=None, edge_list=None):
        self.num_nodes = int(num_nodes)
        self.edges = np.ascontiguousarray(edges, dtype=np.int32).reshape(-1, 2)
        self.num_edges = int(self.edges.shape[0])
        self.node_sides = (
            None if node_sides is None
            else np.ascontiguousarray(node_sides, dtype=np.uint8)
        )
        self.span_on_first_merge = bool(span_on_first_merge)
        if self.num_nodes < 1:
            raise ValueError('a percolation graph needs at least one node')
        if self.edges.size and (self.edges.min() < 0 or
                                self.edges.max() >= self.num_nodes):
            raise ValueError('edge endpoints must be node indices < num_nodes')
        if self.node_sides is not None and (
                self.node_sides.shape != (self.num_nodes,) or
                (self.node_sides.size and self.node_sides.max() > 2)):
            raise ValueError('node_sides must hold one value 0, 1 or 2 per node')
        self.nodes = nodes
        self._edge_list = edge_list
        # side L when this is square_lattice(L)'s closed form (its edges then
        # follow from their index, pc_sweep_lattice), else 0
        self.lattice_length = 0

    @property
    def spanning(self):
        return self.node_sides is not None

    def edge_labels(self, index):
        """Node-label pair of edge ``index`` (as perc_graph.edges() lists it)."""
        if self._edge_list is not None:
            return self._edge_list[index]
        u, v = self.edges[index]
        if self.nodes is None:
            return (int(u), int(v))
        return (self.nodes[u], self.nodes[v])


class _DisjointSets(object):
    """Plain dict union-find for the (small) auxiliary structure."""

    def __init__(self):
        self.parent = {}

    def find(self, x):
        parent = self.parent
        parent.setdefault(x, x)
        root = x
        while parent[root] != root:
            root = parent[root]
        while parent[x] != root:
            parent[x], x = root, parent[x]
        return root

    def union(self, *objects):
        roots = [self.find(x) for x in objects]
        for r in roots[1:]:
            if r != roots[0]:
                self.parent[r] = roots[0]
        return roots[0]


def spanning_side_flags(node_index, auxiliary_node_attributes,
                        auxiliary_edge_attributes, spanning_sides):
    """Per-node side flags and the pre-joined flag (module docstring).

    ``node_index`` maps perc-graph node labels to indices 0..N-1.
    Raises ValueError unless there are exactly two sides with nodes.
    """
    sides = list(spanning_sides)
    if len(sides) != 2:
        raise ValueError(
            'Spanning cluster is to be detected, but auxiliary nodes '
            'of less or more than 2 types (sides) given.'
        )
    ds = _DisjointSets()
    side_roots = {}
    for side in sides:
        nodes = [
            node for (node, node_side) in auxiliary_node_attributes.items()
            if node_side == side
        ]
        if not nodes:
            raise ValueError(
                'Spanning cluster is to be detected, but side {!r} has no '
                'auxiliary nodes.'.format(side))
        ds.union(*nodes)
        side_roots[side] = ds.find(nodes[0])
    for (edge, edge_side) in auxiliary_edge_attributes.items():
        ds.union(side_roots[edge_side], *edge)
    r0 = ds.find(side_roots[sides[0]])
    r1 = ds.find(side_roots[sides[1]])

    flags = np.zeros(len(node_index), dtype=np.uint8)
    if r0 == r1:
        return flags, True
    for node in list(ds.parent):
        idx = node_index.get(node)
        if idx is None:
            continue
        root = ds.find(node)
        if root == r0:
            flags[idx] = 1
        elif root == r1:
            flags[idx] = 2
    return flags, False


def from_percolation_graph(perc_graph, spanning_cluster=True,
                           auxiliary_node_attributes=None,
                           auxiliary_edge_attributes=None,
                           spanning_sides=None, **kwargs):
    """GraphArrays of the keyword dict percolate.percolation_graph() returns
    (percolate/percolate.py:31-100), or of the same keywords passed to the
    percolate.hpc functions."""
    # comprehensions, not list(): list() asks the networkx views for their
    # length first, which sums the degrees of every node
    nodes = [node for node in perc_graph.nodes()]
    index = {node: i for i, node in enumerate(nodes)}
    edge_list = [edge for edge in perc_graph.edges()]
    edges = np.fromiter((index[node] for edge in edge_list for node in edge),
                        dtype=np.int32, count=2 * len(edge_list)).reshape(-1, 2)
    node_sides = None
    span_first = False
    if spanning_cluster:
        node_sides, span_first = spanning_side_flags(
            index, auxiliary_node_attributes or {},
            auxiliary_edge_attributes or {}, spanning_sides or [],
        )
    return GraphArrays(
        len(nodes), edges, node_sides, span_first, nodes=nodes,
        edge_list=edge_list,
    )


def locality_order(graph, tile_rows=4, tile_cols=8):
    """Node relabelling old -> new (int64 array) that stores the union-find
    codes of 2-D grid nodes in tiles of tile_rows x tile_cols (32 codes =
    one 128-byte line), so the two endpoints of most bonds share a line; or
    None when the node labels are not pairs of integers.  The sweep's
    results do not depend on node labels (only on cluster sizes and the
    spanning sides), so any relabelling gives the same records."""
    nodes = graph.nodes
    if nodes is None or len(nodes) != graph.num_nodes or not len(nodes):
        return None
    try:
        coords = np.array(nodes, dtype=np.int64)
    except (TypeError, ValueError):
        return None
    if coords.ndim != 2 or coords.shape[1] != 2:
        return None
    a = coords[:, 0] - coords[:, 0].min()
    b = coords[:, 1] - coords[:, 1].min()
    key = np.lexsort((b % tile_cols, a % tile_rows, b // tile_cols, a // tile_rows))
    new = np.empty(graph.num_nodes, dtype=np.int64)
    new[key] = np.arange(graph.num_nodes, dtype=np.int64)
    return new


def square_lattice(length, spanning_cluster=True):
    """Closed form of percolation_graph(spanning_2d_grid(length)) with
    spanning_cluster=True (with spanning_cluster=False: the L x L grid
    alone, without spanning sides; percolation_graph then keeps the whole
    grid, auxiliary rows included, which this is not).

    Nodes (i, j), i = 1..L, j = 0..L-1 of the grid get index (i-1)*L + j; the
    edges come in networkx's order for grid_2d_graph subgraphs: row by row,
    for each node its edge to the next row, then to the next column.  Rows 1
    and L attach to spanning sides 0 and 1 (percolate/percolate.py:931-965).
    """
    L = int(length)
    if L < 1:
        raise ValueError('length must be a positive integer')
    N = L * L
    idx = np.arange(N, dtype=np.int64).reshape(L, L)
    blocks = []
    for r in range(L):
        if r < L - 1:
            row = np.empty((2 * L - 1, 2), dtype=np.int64)
            row[0::2, 0] = idx[r]
            row[0::2, 1] = idx[r + 1]
            row[1::2, 0] = idx[r, :-1]
            row[1::2, 1] = idx[r, 1:]
        else:
            row = np.stack([idx[r, :-1], idx[r, 1:]], axis=1)
        blocks.append(row)
    edges = np.concatenate(blocks).astype(np.int32) if blocks else np.zeros((0, 2), np.int32)
    node_sides = None
    span_first = False
    if spanning_cluster:
        node_sides = np.zeros(N, dtype=np.uint8)
        if L == 1:
            span_first = True
        else:
            node_sides[idx[0]] = 1
            node_sides[idx[L - 1]] = 2
    nodes = [(i, j) for i in range(1, L + 1) for j in range(L)]
    ga = GraphArrays(N, edges, node_sides, span_first, nodes=nodes)
    ga.lattice_length = L
    return ga


def linear_chain(length, spanning_cluster=True):
    """Closed form of percolation_graph(spanning_1d_chain(length)).

    Perc nodes 1..length get indices 0..length-1, edges (i, i+1) in order;
    the end nodes attach to sides 0 and 1 (percolate/percolate.py:899-928).
    """
    n = int(length)
    if n < 1:
        raise ValueError('length must be a positive integer')
    edges = np.stack([np.arange(n - 1), np.arange(1, n)], axis=1)
    node_sides = None
    span_first = False
    if spanning_cluster:
        node_sides = np.zeros(n, dtype=np.uint8)
        if n == 1:
            span_first = True
        else:
            node_sides[0] = 1
            node_sides[n - 1] = 2
    nodes = list(range(1, n + 1))
    return GraphArrays(n, edges, node_sides, span_first, nodes=nodes)
