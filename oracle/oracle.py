"""TEST INFRASTRUCTURE ONLY: CPU oracle of the Newman-Ziff hot path.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg may
import this module, and only as the checker; the product package
(pypercolate_amd) never imports, calls or links it.

Each function restates the reference (andsor/pypercolate 0.4.6) step for
step, citing the lines it follows:

* ``percolation_graph_arrays``  percolate/percolate.py:31-100 (networkx 3 API)
* ``square_lattice_graph``      the same for spanning_2d_grid(L), closed form
* ``bond_runs``                 percolate/hpc.py:194-307 via nz_oracle.c
* ``microcanonical_sums``       exact sums over runs (nz_oracle.c)
* ``bond_sums``                 sweep + exact sums without traces (nz_oracle.c)
* ``binomial_pmf``              percolate/percolate.py:1067-1109
* ``canonical_statistics``      percolate/hpc.py:443-515
* ``initialize_canonical_averages``  percolate/hpc.py:561-635
* ``online_variance``           simoa.stats.online_variance (Chan et al.), the
                                third-party merge behind hpc.py:677-692
* ``bond_reduce``               percolate/hpc.py:638-702
* ``finalize_canonical_averages``  percolate/hpc.py:752-834
* ``microcanonical_step_stats`` percolate/percolate.py:450-705, 862-891
* ``microcanonical_arrays``     percolate/percolate.py:968-1064
* ``canonical_averages``        percolate/percolate.py:1112-1223
* ``philox4x32_10``, ``bucket_permutations`` (C, nz_shuffle.c) and
  ``bucket_permutation_py``: the device-permutation algorithm of this package
  (no reference counterpart), a Philox-keyed Rao-Sandelius shuffle.

Pinning (tests/test_oracle_*.py): the sweep against the reference's own
known-answer tests, the tutorial traces, brute-force connected components
and closed-form end states; the binomial weights against scipy.stats.binom;
the Chan merge against numpy mean/variance; finalize against scipy's t CDF;
Philox against the Random123 known-answer vectors.  ``online_variance`` is
pinned by the closed form (numpy's two-pass mean and sum of squared
deviations of the concatenated samples) and by associativity under
re-bracketing (mean to 1e-12, sum of squares to 1e-10): FP64 merges
agree to rounding.  What no check here
fixes is simoa's own operation order, hence its last bits (simoa is absent
here, and unpinned in requirements.txt).
"""

import ctypes
import os
import subprocess

import numpy as np
import scipy.stats

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libnzoracle.so')

_lib = None


def build():
    """Compile nz_oracle.c into oracle/libnzoracle.so."""
    subprocess.check_call(['make', '-s', '-C', HERE])


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        i64 = ctypes.c_int64
        lib.nzo_bond_runs.restype = ctypes.c_int
        lib.nzo_bond_runs.argtypes = [i64, i64, i64, vp, ctypes.c_int, i64,
                                      vp, i64, vp, vp, vp, vp, vp, vp, vp,
                                      vp, ctypes.c_int]
        lib.nzo_microcanonical_sums.restype = ctypes.c_int
        lib.nzo_microcanonical_sums.argtypes = [i64, i64, vp, vp, vp, vp, vp,
                                                vp]
        lib.nzo_philox4x32_10.restype = None
        lib.nzo_philox4x32_10.argtypes = [vp, vp, vp]
        lib.nzo_bucket_permutations.restype = ctypes.c_int
        lib.nzo_bucket_permutations.argtypes = [ctypes.c_uint64, ctypes.c_uint32,
                                                i64, i64, i64, i64, i64, i64, vp,
                                                ctypes.c_int]
        lib.nzo_bond_sums.restype = ctypes.c_int
        lib.nzo_bond_sums.argtypes = [i64, i64, i64, vp, ctypes.c_int, i64,
                                      vp, i64, vp, vp, vp, vp, vp,
                                      ctypes.c_int, ctypes.c_int]
        _lib = lib
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------------------
# Graph preparation: percolate/percolate.py:31-100 with networkx 3 calls
# ---------------------------------------------------------------------------
def percolation_graph_arrays(graph, spanning_cluster=True):
    """Integer arrays of percolation_graph(graph, spanning_cluster).

    Perc nodes are numbered in perc_graph.nodes() order, edges in
    perc_graph.edges() order; auxiliary nodes follow the perc nodes.
    """
    ret = {'spanning': bool(spanning_cluster)}
    if spanning_cluster:
        aux_attr = {n: d['span'] for n, d in graph.nodes(data=True) if 'span' in d}
        if not aux_attr:
            raise ValueError('no auxiliary nodes')
        sides = list(set(aux_attr.values()))
        if len(sides) != 2:
            raise ValueError('auxiliary nodes of less or more than 2 sides')
        perc_graph = graph.subgraph(
            [n for n in graph.nodes() if 'span' not in graph.nodes[n]])
    else:
        perc_graph = graph
        aux_attr = {}
        sides = []
    nodes = list(perc_graph.nodes())
    index = {n: i for i, n in enumerate(nodes)}
    aux_nodes = list(aux_attr)
    for j, n in enumerate(aux_nodes):
        index[n] = len(nodes) + j
    edge_list = list(perc_graph.edges())
    ret['num_nodes'] = len(nodes)
    ret['num_edges'] = len(edge_list)
    ret['nodes'] = nodes
    ret['edge_list'] = edge_list
    ret['edges'] = np.array([(index[u], index[v]) for u, v in edge_list],
                            dtype=np.int64).reshape(-1, 2)
    ret['aux_side'] = np.array([sides.index(aux_attr[n]) for n in aux_nodes],
                               dtype=np.int32)
    aux_edges = []
    aux_edge_side = []
    if spanning_cluster:
        for (u, v), side in _edge_span_attributes(graph).items():
            aux_edges.append((index[u], index[v]))
            aux_edge_side.append(sides.index(side))
    ret['aux_edges'] = np.array(aux_edges, dtype=np.int64).reshape(-1, 2)
    ret['aux_edge_side'] = np.array(aux_edge_side, dtype=np.int32)
    return ret


def square_lattice_graph(length):
    """percolation_graph_arrays(spanning_2d_grid(length)) without networkx
    (which takes minutes and GBs at L = 1024): perc node (i, j), i = 1..L,
    j = 0..L-1, gets index (i-1) L + j; the edges in networkx's order for
    the grid's perc subgraph (row by row, each node's edge down, then
    right); one auxiliary node per column on each side, joined to the first
    and last row (percolate/percolate.py:931-965).  tests/test_oracle_*
    pin it to the networkx-built arrays on small lattices."""
    L = int(length)
    N = L * L
    edges = []
    for i in range(L):
        for j in range(L):
            u = i * L + j
            if i + 1 < L:
                edges.append((u, u + L))
            if j + 1 < L:
                edges.append((u, u + 1))
    return {'num_nodes': N, 'num_edges': len(edges), 'spanning': True,
            'edges': np.array(edges, dtype=np.int64).reshape(-1, 2),
            'aux_side': np.array([0] * L + [1] * L, dtype=np.int32),
            'aux_edges': np.array([(j, N + j) for j in range(L)] +
                                  [((L - 1) * L + j, N + L + j) for j in range(L)],
                                  dtype=np.int64).reshape(-1, 2),
            'aux_edge_side': np.array([0] * L + [1] * L, dtype=np.int32)}


def _edge_span_attributes(graph):
    return {(u, v): d['span'] for u, v, d in graph.edges(data=True) if 'span' in d}


# ---------------------------------------------------------------------------
# Sweep (nz_oracle.c)
# ---------------------------------------------------------------------------
def bond_runs(g, perms, threads=None, exact=False):
    """Traces of runs with the given permutations (R, M).

    Returns dict edge (R,S) u32, has_spanning_cluster (R,S) bool,
    max_cluster_size (R,S) u32, moments (R,S,5) u64 (the reference's values,
    mod 2**64); S = M + 1.  With exact=True also moments_hi (R,S,5) u64, the
    high words of the exact 128-bit moments.
    """
    lib = _load()
    M = g['num_edges']
    perms = np.ascontiguousarray(perms, dtype=np.int64).reshape(-1, M) if M else \
        np.zeros((np.asarray(perms).shape[0], 0), dtype=np.int64)
    R = perms.shape[0]
    S = M + 1
    edge = np.zeros((R, S), dtype=np.uint32)
    span = np.zeros((R, S), dtype=np.uint8)
    mx = np.zeros((R, S), dtype=np.uint32)
    mom = np.zeros((R, S, 5), dtype=np.uint64)
    mom_hi = np.zeros((R, S, 5), dtype=np.uint64) if exact else None
    edges = np.ascontiguousarray(g['edges'], dtype=np.int64)
    aux_side = np.ascontiguousarray(g['aux_side'], dtype=np.int32)
    aux_edges = np.ascontiguousarray(g['aux_edges'], dtype=np.int64)
    aux_edge_side = np.ascontiguousarray(g['aux_edge_side'], dtype=np.int32)
    rc = lib.nzo_bond_runs(
        R, g['num_nodes'], M, _p(edges), int(g['spanning']), aux_side.size,
        _p(aux_side), aux_edge_side.size, _p(aux_edges), _p(aux_edge_side),
        _p(perms), _p(edge), _p(span), _p(mx), _p(mom), _p(mom_hi),
        int(threads or os.cpu_count() or 1))
    if rc:
        raise ValueError('oracle: an auxiliary side has no nodes')
    out = {'edge': edge, 'has_spanning_cluster': span.astype(bool),
           'max_cluster_size': mx, 'moments': mom}
    if exact:
        out['moments_hi'] = mom_hi
    return out


def exact_moments(traces):
    """(R, S, 5) object array of the exact moments of bond_runs(exact=True)."""
    return (traces['moments_hi'].astype(object) * (1 << 64) +
            traces['moments'].astype(object))


def microcanonical_sums(traces, limbs=False):
    """Exact sums over runs per step: (sum1, sum2) object arrays (7, S) of
    Python ints, columns max, spanning, moments 0..4 (with limbs=True the
    uint64 limb arrays (7, S, 2) and (7, S, 3) instead).  The moments are the
    traces' exact values when they carry moments_hi (bond_runs(exact=True)),
    else their 64-bit values."""
    lib = _load()
    span = np.ascontiguousarray(traces['has_spanning_cluster'], dtype=np.uint8)
    mx = np.ascontiguousarray(traces['max_cluster_size'], dtype=np.uint32)
    mom = np.ascontiguousarray(traces['moments'], dtype=np.uint64)
    hi = traces.get('moments_hi')
    hi = None if hi is None else np.ascontiguousarray(hi, dtype=np.uint64)
    R, S = mx.shape
    s1 = np.zeros((7, S, 2), dtype=np.uint64)
    s2 = np.zeros((7, S, 3), dtype=np.uint64)
    if lib.nzo_microcanonical_sums(R, S, _p(span), _p(mx), _p(mom), _p(hi),
                                   _p(s1), _p(s2)):
        raise OverflowError('oracle: a sum exceeds its 128/192 bits')
    if limbs:
        return s1, s2
    return limbs_to_int(s1), limbs_to_int(s2)


def bond_sums(g, perms, sums=None, threads=None):
    """Exact per-step sums over the runs of perms (R, M), straight from the
    sweep with exact moments (no traces kept): (sum1, sum2) uint64 limb
    arrays (7, S, 2) and (7, S, 3).  With sums=(sum1, sum2) the new runs'
    sums are added to those arrays in place (chunks of a large sample)."""
    lib = _load()
    M = g['num_edges']
    S = M + 1
    perms = np.ascontiguousarray(perms, dtype=np.int32).reshape(-1, M)
    if sums is None:
        s1 = np.zeros((7, S, 2), dtype=np.uint64)
        s2 = np.zeros((7, S, 3), dtype=np.uint64)
        acc = 0
    else:
        s1, s2 = sums
        acc = 1
    edges = np.ascontiguousarray(g['edges'], dtype=np.int64)
    aux_side = np.ascontiguousarray(g['aux_side'], dtype=np.int32)
    aux_edges = np.ascontiguousarray(g['aux_edges'], dtype=np.int64)
    aux_edge_side = np.ascontiguousarray(g['aux_edge_side'], dtype=np.int32)
    rc = lib.nzo_bond_sums(
        perms.shape[0], g['num_nodes'], M, _p(edges), int(g['spanning']),
        aux_side.size, _p(aux_side), aux_edge_side.size, _p(aux_edges),
        _p(aux_edge_side), _p(perms), _p(s1), _p(s2), acc,
        int(threads or os.cpu_count() or 1))
    if rc == -1:
        raise ValueError('oracle: an auxiliary side has no nodes')
    if rc == -2:
        raise OverflowError('oracle: a sum exceeds its 128/192 bits')
    if rc:
        raise MemoryError('oracle: out of memory')
    return s1, s2


def limbs_to_int(limbs):
    limbs = np.asarray(limbs).view(np.uint64)
    out = np.zeros(limbs.shape[:-1], dtype=object)
    for j in range(limbs.shape[-1] - 1, -1, -1):
        out = out * (1 << 64) + limbs[..., j].astype(object)
    return out


def randomstate_permutations(seeds, num_edges):
    """numpy RandomState(seed).permutation(M) per seed (hpc.py:195,206)."""
    return np.stack([np.random.RandomState(seed=s).permutation(num_edges)
                     for s in seeds]).astype(np.int64) if len(seeds) else \
        np.zeros((0, num_edges), dtype=np.int64)


# ---------------------------------------------------------------------------
# Canonical ensemble: percolate/percolate.py:1067-1109 and hpc.py:443-834
# ---------------------------------------------------------------------------
def binomial_pmf(n, p):
    """percolate/percolate.py:1067-1109: ratio recursion outward from
    round(p n), then normalised by the sum."""
    n = int(n)
    ret = np.empty(n + 1)
    nmax = int(np.round(p * n))
    ret[nmax] = 1.0
    old = np.seterr(under='ignore', divide='ignore', invalid='ignore')
    for i in range(nmax + 1, n + 1):
        ret[i] = ret[i - 1] * (n - i + 1.0) / i * p / (1.0 - p)
    for i in range(nmax - 1, -1, -1):
        ret[i] = ret[i + 1] * (i + 1.0) / (n - i) * (1.0 - p) / p
    np.seterr(**old)
    return ret / ret.sum()


def canonical_statistics(trace, convolution_factors):
    """percolate/hpc.py:443-515 for one run and one p: dict of float64."""
    f = np.asarray(convolution_factors, dtype=np.float64)
    ret = {}
    if 'has_spanning_cluster' in trace:
        ret['percolation_probability'] = np.sum(f * trace['has_spanning_cluster'])
    ret['max_cluster_size'] = np.sum(f * trace['max_cluster_size'])
    ret['moments'] = np.sum(f[:, np.newaxis] * trace['moments'], axis=0)
    return ret


def initialize_canonical_averages(stats_list):
    """percolate/hpc.py:561-635 over a list of canonical_statistics dicts
    (one per p): n = 1, mean = value, m2 = 0."""
    out = {'number_of_runs': np.ones(len(stats_list), dtype=np.uint32)}
    for key in stats_list[0]:
        vals = np.array([s[key] for s in stats_list], dtype=np.float64)
        out[key + '_mean'] = vals
        out[key + '_m2'] = np.zeros_like(vals)
    return out


def online_variance(a, b):
    """Pairwise (n, mean, M2) merge of Chan, Golub & LeVeque (the algorithm
    simoa.stats.online_variance implements)."""
    n_a, mean_a, m2_a = a
    n_b, mean_b, m2_b = b
    n_a = np.asarray(n_a, dtype=np.float64)
    n_b = np.asarray(n_b, dtype=np.float64)
    n = n_a + n_b
    delta = np.asarray(mean_b) - np.asarray(mean_a)
    mean = np.asarray(mean_a) + delta * n_b / n
    m2 = np.asarray(m2_a) + np.asarray(m2_b) + delta ** 2 * n_a * n_b / n
    return n, mean, m2


def bond_reduce(a, b):
    """percolate/hpc.py:638-702 on dicts of arrays."""
    out = {'number_of_runs': a['number_of_runs'] + b['number_of_runs']}
    for key in a:
        if not key.endswith('_mean'):
            continue
        base = key[:-5]
        na = a['number_of_runs'].astype(np.float64)
        nb = b['number_of_runs'].astype(np.float64)
        if a[key].ndim == 2:
            na, nb = na[:, None], nb[:, None]
        _, mean, m2 = online_variance((na, a[key], a[base + '_m2']),
                                      (nb, b[key], b[base + '_m2']))
        out[key] = mean
        out[base + '_m2'] = m2
    return out


def finalize_canonical_averages(number_of_nodes, ps, avg, alpha):
    """percolate/hpc.py:752-834 on dicts of arrays."""
    n = avg['number_of_runs'].astype(np.float64)
    sqrt_n = np.sqrt(n)
    out = {'number_of_runs': avg['number_of_runs'], 'p': np.asarray(ps),
           'alpha': alpha}

    def transform(key, final, normalize):
        mean = np.array(avg[key + '_mean'], dtype=np.float64)
        m2 = avg[key + '_m2']
        nn = n[:, None] if m2.ndim == 2 else n
        std = np.sqrt(m2 / (nn - 1))
        if normalize:
            mean = mean / number_of_nodes
            std = std / number_of_nodes
        sq = sqrt_n[:, None] if m2.ndim == 2 else sqrt_n
        lo, hi = scipy.stats.t.interval(1 - alpha, df=nn - 1, loc=mean,
                                        scale=std / sq)
        out[final + '_mean'] = mean
        out[final + '_std'] = std
        out[final + '_ci'] = np.stack([lo, hi], axis=-1)

    if 'percolation_probability_mean' in avg:
        transform('percolation_probability', 'percolation_probability', False)
    transform('max_cluster_size', 'percolation_strength', True)
    transform('moments', 'moments', True)
    return out


# ---------------------------------------------------------------------------
# Microcanonical averages: percolate/percolate.py:450-705, 862-1064
# ---------------------------------------------------------------------------
def microcanonical_step_stats(max_sizes, moments, spanning, alpha):
    """One step over R runs: the dict the reference's per-step helpers build
    (percolate.py:557-570, 613-635, 681-705).  Inputs are float arrays."""
    ret = {}
    runs = max_sizes.size
    sqrt_n = np.sqrt(runs)
    m = max_sizes.mean()
    ret['max_cluster_size'] = m
    sd = max_sizes.std(ddof=1)
    if sd:
        ret['max_cluster_size_ci'] = np.array(scipy.stats.t.interval(
            1 - alpha, df=runs - 1, loc=m, scale=sd / sqrt_n))
    else:
        ret['max_cluster_size_ci'] = m * np.ones(2)
    mm = moments.mean(axis=0)
    ret['moments'] = mm
    msd = moments.std(axis=0, ddof=1)
    ci = np.empty((5, 2))
    for k in range(5):
        if msd[k]:
            ci[k] = scipy.stats.t.interval(1 - alpha, df=runs - 1, loc=mm[k],
                                           scale=msd[k] / sqrt_n)
        else:
            ci[k] = mm[k] * np.ones(2)
    ret['moments_ci'] = ci
    if spanning is not None:
        k = spanning.sum(dtype=np.float64)
        ret['spanning_cluster'] = (k + 1) / (runs + 2)
        ret['spanning_cluster_ci'] = scipy.stats.beta.ppf(
            [alpha / 2, 1 - alpha / 2], k + 1, runs - k + 1)
    return ret


def microcanonical_arrays(traces, num_nodes, alpha):
    """percolate.py:708-896 + 968-1064 from per-run traces: stacked arrays,
    all but spanning divided by N."""
    mx = traces['max_cluster_size'].astype(np.float64)
    mom = traces['moments'].astype(np.float64)
    span = traces.get('has_spanning_cluster')
    R, S = mx.shape
    out = {'max_cluster_size': np.empty(S), 'max_cluster_size_ci': np.empty((S, 2)),
           'moments': np.empty((5, S)), 'moments_ci': np.empty((5, S, 2))}
    if span is not None:
        out['spanning_cluster'] = np.empty(S)
        out['spanning_cluster_ci'] = np.empty((S, 2))
    for s in range(S):
        st = microcanonical_step_stats(
            mx[:, s], mom[:, s, :], None if span is None else span[:, s], alpha)
        out['max_cluster_size'][s] = st['max_cluster_size']
        out['max_cluster_size_ci'][s] = st['max_cluster_size_ci']
        out['moments'][:, s] = st['moments']
        out['moments_ci'][:, s] = st['moments_ci']
        if span is not None:
            out['spanning_cluster'][s] = st['spanning_cluster']
            out['spanning_cluster_ci'][s] = st['spanning_cluster_ci']
    for key in list(out):
        if 'spanning_cluster' in key:
            continue
        out[key] = out[key] / num_nodes
    out['N'] = num_nodes
    out['M'] = S - 1
    return out


def canonical_averages(ps, arrays):
    """percolate/percolate.py:1112-1223."""
    ps = np.asarray(ps)
    M = arrays['M']
    ret = {'ps': ps, 'N': arrays['N'], 'M': M}
    ret['max_cluster_size'] = np.empty(ps.size)
    ret['max_cluster_size_ci'] = np.empty((ps.size, 2))
    if 'spanning_cluster' in arrays:
        ret['spanning_cluster'] = np.empty(ps.size)
        ret['spanning_cluster_ci'] = np.empty((ps.size, 2))
    ret['moments'] = np.empty((5, ps.size))
    ret['moments_ci'] = np.empty((5, ps.size, 2))
    for i, p in enumerate(ps):
        b = binomial_pmf(M, p)
        for key, value in arrays.items():
            if len(key) <= 1:
                continue
            if key in ('max_cluster_size', 'spanning_cluster'):
                ret[key][i] = np.sum(b * value)
            elif key in ('max_cluster_size_ci', 'spanning_cluster_ci'):
                ret[key][i] = np.sum(np.tile(b, (2, 1)).T * value, axis=0)
            elif key == 'moments':
                ret[key][:, i] = np.sum(np.tile(b, (5, 1)) * value, axis=1)
            elif key == 'moments_ci':
                ret[key][:, i] = np.sum(
                    np.rollaxis(np.tile(b, (5, 2, 1)), 2, 1) * value, axis=1)
    return ret


# ---------------------------------------------------------------------------
# Device-permutation algorithm (this package's own; no reference counterpart)
# ---------------------------------------------------------------------------
PHILOX_M0, PHILOX_M1 = 0xD2511F53, 0xCD9E8D57
PHILOX_W0, PHILOX_W1 = 0x9E3779B9, 0xBB67AE85
MASK32 = 0xFFFFFFFF


def philox4x32_10(ctr, key):
    """Philox4x32 with 10 rounds (Salmon et al., SC'11) on Python ints."""
    c0, c1, c2, c3 = [int(x) & MASK32 for x in ctr]
    k0, k1 = [int(x) & MASK32 for x in key]
    for i in range(10):
        if i:
            k0 = (k0 + PHILOX_W0) & MASK32
            k1 = (k1 + PHILOX_W1) & MASK32
        p0 = PHILOX_M0 * c0
        p1 = PHILOX_M1 * c2
        c0, c1, c2, c3 = ((p1 >> 32) ^ c1 ^ k0, p1 & MASK32,
                          (p0 >> 32) ^ c3 ^ k1, p0 & MASK32)
    return c0, c1, c2, c3


# The size parameters of the device's layout (pc_permute_philox).
SHUFFLE_BUCKET = 4096
SHUFFLE_GROUP_CAP = 8192
SHUFFLE_WHOLE = 32768
SHUFFLE_SUB = 64


def philox4x32_10_c(ctr, key):
    """The C restatement's Philox (nz_shuffle.c), for pinning it against the
    known-answer vectors like philox4x32_10."""
    lib = _load()
    c = np.array(ctr, dtype=np.uint32)
    k = np.array(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib.nzo_philox4x32_10(_p(c), _p(k), _p(out))
    return tuple(int(v) for v in out)


def bucket_permutations(seed, first_run, runs, num_edges,
                        bucket=SHUFFLE_BUCKET, group_cap=SHUFFLE_GROUP_CAP,
                        whole=SHUFFLE_WHOLE, sub=SHUFFLE_SUB, threads=None):
    """The permutations pc_permute_philox draws for runs first_run ..
    first_run + runs - 1 (int32 array (runs, M)), restated sequentially in
    C (nz_shuffle.c): a Philox-keyed Rao-Sandelius shuffle.  The size
    parameters default to the device's; smaller ones put small inputs
    through many buckets and the wide Fisher-Yates path."""
    lib = _load()
    M = int(num_edges)
    out = np.zeros((int(runs), M), dtype=np.int32)
    rc = lib.nzo_bucket_permutations(
        int(seed) & 0xFFFFFFFFFFFFFFFF, int(first_run), int(runs), M,
        int(bucket), int(group_cap), int(whole), int(sub), _p(out),
        int(threads or os.cpu_count() or 1))
    if rc:
        raise ValueError('oracle: size outside the shuffle layout')
    return out


def _half16(words, h):
    return (words[h >> 1] >> (16 * (h & 1))) & 0xFFFF


def bucket_permutation_py(seed, run, num_edges, bucket=SHUFFLE_BUCKET,
                          group_cap=SHUFFLE_GROUP_CAP, whole=SHUFFLE_WHOLE,
                          sub=SHUFFLE_SUB, stats=None):
    """The same layout as bucket_permutations, written again in pure Python
    on top of philox4x32_10 (pinned by the Random123 vectors), for small
    inputs: it pins the C restatement's counters, labels and draws.
    ``stats`` (a dict), if given, counts the rejected 16-bit words
    (``'rejected'``) and the wide Fisher-Yates buckets (``'wide'``)."""
    stats = {} if stats is None else stats
    stats.setdefault('rejected', 0)
    stats.setdefault('wide', 0)
    key = (seed & MASK32, (seed >> 32) & MASK32)

    def key16(tag, b, j):
        w = philox4x32_10((((j >> 9) << 6) | (j & 63), run, b, tag << 28), key)
        return _half16(w, (j >> 6) & 7)

    def stable_group(items, labels, nlab):
        groups = [[] for _ in range(nlab)]
        for it, lab in zip(items, labels):
            groups[lab].append(it)
        return groups

    def shuffle_bucket(a, b, cap):
        n = len(a)
        if n <= 1:
            return list(a)
        if n > cap:
            stats['wide'] += 1
            a = list(a)
            words, u = [], 0
            for j in range(n - 1, 0, -1):
                r = j + 1
                while True:
                    if not words:
                        words = list(philox4x32_10((u, run, b, 4 << 28), key))
                        u += 1
                    m = words.pop(0) * r
                    lo = m & MASK32
                    if lo < r and lo < (2 ** 32 - r) % r:
                        continue
                    break
                k = m >> 32
                a[j], a[k] = a[k], a[j]
            return a
        lb = 0
        while (sub << lb) < n:
            lb += 1
        labels = [key16(2, b, j) >> (16 - lb) if lb else 0 for j in range(n)]
        groups = stable_group(a, labels, 1 << lb)
        out, st = [], 0
        for g in groups:
            c = len(g)
            main, spare = [], []       # the sub-bucket's two word streams
            u = v = 0
            for j in range(c - 1, 0, -1):
                r = j + 1
                if not main:
                    w = philox4x32_10((st, run, b, (3 << 28) | u), key)
                    main = [_half16(w, h) for h in range(8)]
                    u += 1
                m = main.pop(0) * r
                while (m & 0xFFFF) < r and (m & 0xFFFF) < (65536 - r) % r:
                    stats['rejected'] += 1
                    if not spare:
                        w = philox4x32_10((st, run, b, (5 << 28) | v), key)
                        spare = [_half16(w, h) for h in range(8)]
                        v += 1
                    m = spare.pop(0) * r
                k = m >> 16
                g[j], g[k] = g[k], g[j]
            out.extend(g)
            st += c
        return out

    M = int(num_edges)
    lb1 = 0
    while M > whole and (bucket << lb1) < M:
        lb1 += 1
    labels = [key16(1, 0, i) >> (16 - lb1) if lb1 else 0 for i in range(M)]
    buckets = stable_group(list(range(M)), labels, 1 << lb1)
    out = []
    for b, items in enumerate(buckets):
        out.extend(shuffle_bucket(items, b, group_cap if lb1 else whole))
    return np.array(out, dtype=np.int64)
