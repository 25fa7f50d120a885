"""Writes the golden fixtures under tests/golden/ (run in the build container,
where the reference checkout is readable as text at /root/reference).

* kat_3x3.json: the known-answer test of percolate/test/test_percolate.py:
  the 3x3 grid fixture (lines 49-73), its edge diagram and the permutation
  numpy's legacy RandomState(42) draws (lines 132-142), and the expected
  max_cluster_size (lines 148-150), moments (lines 160-174) and spanning
  flags (line 187) per step.
* tutorial_traces.json: the two printed sample-state traces of
  docs/tutorial-bond-square-lattice.ipynb (linear chain of length 10, and
  the 3x3 spanning grid), parsed from the notebook's stored outputs.
* brute_force.npz: per-step observables of small lattices computed by brute
  force with networkx connected components (the strategy of
  percolate/test/test_hpc.py:157-213,250-292), independent of the oracle.
* binomial.npz: scipy.stats.binom.pmf at the (n, p) pairs the tests use.
"""
import ast
import json
import os
import re

import networkx as nx
import numpy as np
import scipy.stats

HERE = os.path.dirname(os.path.abspath(__file__))
REF = '/root/reference'


def kat_3x3():
    # numbers as written in percolate/test/test_percolate.py:141-187
    return {
        'source': 'percolate/test/test_percolate.py:49-73,132-192',
        'edges_added_by_fixture': [[i, i + j] for i in [1, 4, 7] for j in [-1, 1]]
        + [[i, i + j] for i in [3, 4, 5] for j in [-3, 3]],
        'aux_side0_edges': [[0, 9], [3, 10], [6, 11]],
        'aux_side1_edges': [[2, 12], [5, 13], [8, 14]],
        'seed': 42,
        'permutation': [10, 9, 0, 8, 5, 2, 1, 11, 4, 7, 3, 6],
        'max_cluster_size': [1, 2, 2, 2, 3, 4, 4, 7, 9, 9, 9, 9, 9],
        'moments': [
            [8, 8, 8, 8, 8], [7, 7, 7, 7, 7], [6, 7, 9, 13, 21],
            [5, 7, 11, 19, 35], [4, 6, 10, 18, 34], [3, 5, 9, 17, 33],
            [2, 5, 13, 35, 97], [1, 2, 4, 8, 16], [0, 0, 0, 0, 0],
            [0, 0, 0, 0, 0], [0, 0, 0, 0, 0], [0, 0, 0, 0, 0],
            [0, 0, 0, 0, 0],
        ],
        'has_spanning_cluster': 6 * [False] + 7 * [True],
    }


def tutorial_traces():
    nb = json.load(open(os.path.join(REF, 'docs',
                                     'tutorial-bond-square-lattice.ipynb')))
    traces = []
    for cell in nb['cells']:
        if cell['cell_type'] != 'code':
            continue
        src = ''.join(cell['source'])
        if 'percolate.sample_states(' not in src or 'pprint' not in src:
            continue
        text = ''.join(''.join(o.get('text', '')) for o in cell['outputs'])
        states = []
        for block in re.findall(r'\{.*?\}', text, flags=re.S):
            block = block.replace('array(', '(').replace('\n', ' ')
            block = re.sub(r'(\d)\.(?=[,\s\]\)])', r'\1', block)
            d = ast.literal_eval(block)
            states.append({
                'n': d['n'], 'N': d['N'], 'M': d['M'],
                'edge': list(map(list, d['edge'])) if isinstance(d.get('edge'), tuple)
                and isinstance(d['edge'][0], tuple) else
                (list(d['edge']) if 'edge' in d else None),
                'max_cluster_size': d['max_cluster_size'],
                'moments': [int(x) for x in d['moments']],
                'has_spanning_cluster': bool(d['has_spanning_cluster']),
            })
        kind = 'chain' if 'chain' in src else 'grid'
        traces.append({'graph': kind, 'length': 10 if kind == 'chain' else 3,
                       'states': states})
    return {'source': 'docs/tutorial-bond-square-lattice.ipynb', 'traces': traces}


def spanning_2d_grid(length):
    g = nx.grid_2d_graph(length + 2, length)
    for i in range(length):
        g.nodes[(0, i)]['span'] = 0
        g[(0, i)][(1, i)]['span'] = 0
        g.nodes[(length + 1, i)]['span'] = 1
        g[(length + 1, i)][(length, i)]['span'] = 1
    return g


def brute_force(length, seeds):
    """Observables per step by brute force on spanning_2d_grid(length)."""
    g = spanning_2d_grid(length)
    perc = g.subgraph([n for n in g.nodes() if 'span' not in g.nodes[n]])
    nodes = list(perc.nodes())
    edges = list(perc.edges())
    M = len(edges)
    aux_edges = [e for e, d in ((e, g.edges[e]) for e in g.edges()) if 'span' in d]
    side = {n: d['span'] for n, d in g.nodes(data=True) if 'span' in d}
    out_max, out_mom, out_span, out_perm = [], [], [], []
    for seed in seeds:
        perm = np.random.RandomState(seed).permutation(M)
        sub = nx.Graph()
        sub.add_nodes_from(nodes)
        aux = nx.Graph(aux_edges)
        mx, mom, sp = [], [], []
        for n in range(M + 1):
            if n:
                sub.add_edge(*edges[perm[n - 1]])
                aux.add_edge(*edges[perm[n - 1]])
            sizes = sorted((len(c) for c in nx.connected_components(sub)),
                           reverse=True)
            mx.append(sizes[0])
            mom.append([sum(s ** k for s in sizes[1:]) for k in range(5)])
            comps = list(nx.connected_components(aux))
            spans = any(
                any(side.get(x) == 0 for x in c) and any(side.get(x) == 1 for x in c)
                for c in comps)
            # the reference re-evaluates spanning on merge steps only
            sp.append(bool(spans) if n else False)
        out_max.append(mx)
        out_mom.append(mom)
        out_span.append(sp)
        out_perm.append(perm)
    return (np.array(out_perm, dtype=np.int32), np.array(out_max, dtype=np.uint32),
            np.array(out_mom, dtype=np.uint64), np.array(out_span, dtype=bool))


def main():
    with open(os.path.join(HERE, 'kat_3x3.json'), 'w') as f:
        json.dump(kat_3x3(), f, indent=1)
    with open(os.path.join(HERE, 'tutorial_traces.json'), 'w') as f:
        json.dump(tutorial_traces(), f, indent=1)
    arrays = {}
    for length, seeds in ((2, range(16)), (3, range(32)), (4, range(16)),
                          (6, range(8))):
        perm, mx, mom, sp = brute_force(length, list(seeds))
        arrays['L{}_perm'.format(length)] = perm
        arrays['L{}_max'.format(length)] = mx
        arrays['L{}_moments'.format(length)] = mom
        arrays['L{}_spanning'.format(length)] = sp
    np.savez_compressed(os.path.join(HERE, 'brute_force.npz'), **arrays)
    binom = {}
    for n, p in ((100, 0.1), (480, 0.5), (32512, 0.45), (1000, 0.01)):
        binom['n{}_p{}'.format(n, p)] = scipy.stats.binom.pmf(np.arange(n + 1), n, p)
    np.savez_compressed(os.path.join(HERE, 'binomial.npz'), **binom)


if __name__ == '__main__':
    main()
