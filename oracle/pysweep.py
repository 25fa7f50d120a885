"""TEST INFRASTRUCTURE -- never imported by the product path.

A pure-Python restatement of the reference's per-edge loop
(percolate/hpc.py:224-307), at the reference's own cost class: a
dictionary-based weighted union-find (the structure of networkx's
UnionFind: full path compression, union by weight, called by hpc.py:224-277)
and numpy uint64 moment vectors updated with `w ** arange(5)` on every
merge.  bench.py times it on one core and with one process per core as the
reference-class CPU baseline; tests/test_oracle_sweep.py checks its traces
against the C oracle.
"""
import numpy as np

_POWERS = np.arange(5, dtype='uint64')


class UnionFind(object):
    """Weighted union-find over hashable items (networkx's UnionFind
    semantics: items appear on first lookup with weight 1)."""

    def __init__(self):
        self.parents = {}
        self.weights = {}

    def __getitem__(self, item):
        parents = self.parents
        if item not in parents:
            parents[item] = item
            self.weights[item] = 1
            return item
        path = [item]
        root = parents[item]
        while root != path[-1]:
            path.append(root)
            root = parents[root]
        for node in path:
            parents[node] = root
        return root

    def union(self, *items):
        roots = [self[x] for x in items]
        heaviest = max(roots, key=lambda r: self.weights[r])
        for r in roots:
            if r != heaviest:
                self.weights[heaviest] += self.weights[r]
                self.parents[r] = heaviest


def bond_run(num_nodes, edges, perm, side_nodes=None, side_edges=None):
    """One run (percolate/hpc.py:194-307): edges is a list of node pairs,
    perm the edge order.  side_nodes: two lists of auxiliary nodes (the
    spanning sides), side_edges: two lists of auxiliary edges.  Returns
    (max_cluster_size uint32 (M+1,), moments uint64 (M+1, 5), spanning bool
    (M+1,) or None), the per-step arrays of bond_microcanonical_statistics."""
    M = len(perm)
    max_size = np.empty(M + 1, dtype='uint32')
    moments = np.empty((M + 1, 5), dtype='uint64')
    spanning = np.zeros(M + 1, dtype=bool) if side_nodes is not None else None
    cur_moments = np.ones(5, dtype='uint64') * np.uint64(num_nodes - 1)
    cur_max = 1
    cur_span = False
    max_size[0] = cur_max
    moments[0] = cur_moments
    ds = UnionFind()
    if side_nodes is not None:
        dss = UnionFind()
        side_roots = []
        for nodes, aux in zip(side_nodes, side_edges):
            dss.union(*nodes)
            for u, v in aux:
                dss.union(nodes[0], u, v)
            side_roots.append(nodes[0])
    max_root = edges[0][0] if edges else 0
    for n in range(M):
        u, v = edges[perm[n]]
        ru, rv = ds[u], ds[v]
        wu, wv = ds.weights[ru], ds.weights[rv]
        if ru != rv:
            ds.union(ru, rv)
            if spanning is not None:
                dss.union(ru, rv)
                cur_span = dss[side_roots[0]] == dss[side_roots[1]]
            root = ds[u]
            weight = ds.weights[root]
            if ru != max_root:
                cur_moments -= np.uint64(wu) ** _POWERS
            if rv != max_root:
                cur_moments -= np.uint64(wv) ** _POWERS
            if max_root in (ru, rv):
                max_root = root
                cur_max = weight
            elif cur_max >= weight:
                cur_moments += np.uint64(weight) ** _POWERS
            else:
                max_root = root
                cur_moments += np.uint64(cur_max) ** _POWERS
                cur_max = weight
        max_size[n + 1] = cur_max
        moments[n + 1] = cur_moments
        if spanning is not None:
            spanning[n + 1] = cur_span
    return max_size, moments, spanning


def lattice_run(length, seed):
    """One run on spanning_2d_grid(length) in the closed-form networkx-3
    edge order, with numpy's RandomState(seed) permutation (hpc.py:195,206):
    the reference-class timing workload."""
    L = length
    N = L * L
    edges = []
    for r in range(L):
        for c in range(L):
            i = r * L + c
            if r + 1 < L:
                edges.append((i, i + L))
            if c + 1 < L:
                edges.append((i, i + 1))
    top = [('top', c) for c in range(L)]
    bottom = [('bottom', c) for c in range(L)]
    side_edges = [[(('top', c), c) for c in range(L)],
                  [(('bottom', c), N - L + c) for c in range(L)]]
    perm = np.random.RandomState(seed).permutation(len(edges))
    return bond_run(N, edges, perm, [top, bottom], side_edges)
