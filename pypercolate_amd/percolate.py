"""Drop-in for ``percolate.percolate`` (andsor/pypercolate 0.4.6).

Same functions, signatures, return values and errors as
percolate/percolate.py, with the sweep, the reduction over runs and the
canonical convolution running on the GPU (libpercolate_hip).  Differences
that follow from running on networkx 3 / numpy 2 and from exact integer
arithmetic are listed where they occur:

* graphs use networkx 3 (``graph.nodes[n]`` for ``graph.node[n]``), and edge
  order is ``perc_graph.edges()`` of the installed networkx;
* moments are exact integers (128-bit) rounded to float64 at the end; the
  reference accumulates them in float64 with int64 powers
  (percolate.py:259,334-350), which agrees exactly while values stay below
  2**53 (every lattice up to L = 128) and wraps int64 for clusters above
  55108 nodes.
"""

import copy

import networkx as nx
import numpy as np
import scipy.stats

from . import _graph

alpha_1sigma = 2 * scipy.stats.norm.cdf(-1.0)
"""
The alpha for the 1 sigma confidence level
"""


_NO_AUX = ('Spanning cluster is to be detected, but no auxiliary nodes '
           'given.')
_NOT_TWO_SIDES = ('Spanning cluster is to be detected, but auxiliary nodes '
                  'of less or more than 2 types (sides) given.')


def _auxiliary_sides(graph):
    """The auxiliary nodes of a spanning-cluster graph (those with a 'span'
    attribute) as {node: side}, and the list of distinct sides; ValueError
    (the reference's messages) unless there are nodes of exactly two
    sides."""
    node_side = {node: data['span'] for node, data in graph.nodes(data=True)
                 if 'span' in data}
    if not node_side:
        raise ValueError(_NO_AUX)
    sides = list(set(node_side.values()))
    if len(sides) != 2:
        raise ValueError(_NOT_TWO_SIDES)
    return node_side, sides


def percolation_graph(graph, spanning_cluster=True):
    """The percolation graph of a graph and the keyword dict the
    ``percolate.hpc`` functions take (percolate/percolate.py:31-100): with
    spanning detection the auxiliary nodes (attribute 'span') are left out
    of ``perc_graph`` and described by ``auxiliary_node_attributes``,
    ``auxiliary_edge_attributes`` and the two ``spanning_sides``."""
    detect = bool(spanning_cluster)
    keywords = {'graph': graph, 'spanning_cluster': detect}
    perc_graph = graph
    if detect:
        node_side, sides = _auxiliary_sides(graph)
        keywords['auxiliary_node_attributes'] = node_side
        keywords['spanning_sides'] = sides
        keywords['auxiliary_edge_attributes'] = nx.get_edge_attributes(graph, 'span')
        perc_graph = graph.subgraph([n for n in graph.nodes() if n not in node_side])
    keywords['perc_graph'] = perc_graph
    keywords['num_nodes'] = perc_graph.number_of_nodes()
    keywords['num_edges'] = perc_graph.number_of_edges()
    return keywords


def spanning_1d_chain(length):
    """Linear chain with auxiliary nodes for spanning cluster detection
    (percolate/percolate.py:899-928)."""
    ret = nx.grid_graph(dim=[int(length + 2)])
    ret.nodes[0]['span'] = 0
    ret[0][1]['span'] = 0
    ret.nodes[length + 1]['span'] = 1
    ret[length][length + 1]['span'] = 1
    return ret


def spanning_2d_grid(length):
    """Square lattice with auxiliary nodes for spanning detection
    (percolate/percolate.py:931-965)."""
    ret = nx.grid_2d_graph(length + 2, length)
    for i in range(length):
        ret.nodes[(0, i)]['span'] = 0
        ret[(0, i)][(1, i)]['span'] = 0
        ret.nodes[(length + 1, i)]['span'] = 1
        ret[(length + 1, i)][(length, i)]['span'] = 1
    return ret


def _graph_arrays(graph, spanning_cluster):
    pg = percolation_graph(graph, spanning_cluster)
    return _graph.from_percolation_graph(**pg)


def _check_model_and_sides(graph, spanning_cluster, model):
    """The eager checks of percolate/percolate.py:210-227."""
    if model != 'bond':
        raise ValueError('Only bond percolation supported.')
    if spanning_cluster:
        _auxiliary_sides(graph)


def _engine(graph_arrays):
    from ._engine import BondEngine
    return BondEngine(graph_arrays)


def _run_batch(ga, perms):
    """Sweep the given (R, M) host permutations on the GPU; returns the
    engine and the SweepResult."""
    eng = _engine(ga)
    res = eng.sweep(eng.upload_permutations(perms))
    return eng, res


# largest graph whose exact per-step sums pc_microcanonical_sums takes
_SUMS_MAX_NODES = 1 << 21


def _single_run(ga, perm):
    """One run with the host permutation perm: largest cluster (int64),
    spanning flag (bool) and the five moments as float64, (S,) and (S, 5).
    The moments are the run's exact integers rounded to float64 (within one
    unit in the last place; the reference sums them in float64 with int64
    powers, percolate/percolate.py:259,334-350, which agrees while they stay
    below 2**53 and wraps for clusters above 55108 nodes): for one run the
    exact 128-bit per-step sums are the run's own values.  Graphs above
    2**21 nodes take the per-run traces, whose moments are kept mod 2**64."""
    eng, res = _run_batch(ga, perm[np.newaxis, :])
    if ga.num_nodes <= _SUMS_MAX_NODES:
        s1, _ = eng.microcanonical_sums(res)
        limbs = s1.cpu().numpy().view(np.uint64)          # (7, S, 2)
        exact = limbs[:, :, 0].astype(np.float64) + \
            limbs[:, :, 1].astype(np.float64) * 18446744073709551616.0
        return (limbs[0, :, 0].astype(np.int64), limbs[1, :, 0] != 0,
                exact[2:].T.copy())
    mx, sp, mom = eng.host_traces(res)
    return mx[0].astype(np.int64), sp[0], mom[0].astype(np.float64)


def sample_states(
    graph, spanning_cluster=True, model='bond', copy_result=True
):
    '''Generate successive sample states of the percolation model
    (percolate/percolate.py:103-356).

    Yields a dict per occupation number n = 0..M with keys n, N, M,
    max_cluster_size, moments (float64, 5), has_spanning_cluster (if
    detected) and, from n = 1 on, edge (the node pair added).  As in the
    reference the permutation is drawn from numpy's global random stream at
    the second ``next()``; the whole run is then swept on the GPU.
    '''
    _check_model_and_sides(graph, spanning_cluster, model)
    ga = _graph_arrays(graph, spanning_cluster)
    N, M = ga.num_nodes, ga.num_edges

    ret = dict()
    ret['n'] = 0
    ret['N'] = N
    ret['M'] = M
    ret['max_cluster_size'] = 1
    ret['moments'] = np.ones(5) * (N - 1)
    if spanning_cluster:
        ret['has_spanning_cluster'] = False

    if copy_result:
        yield copy.deepcopy(ret)
    else:
        yield ret

    perm_edges = np.random.permutation(M)
    if M == 0:
        return
    mx, sp, moments = _single_run(ga, perm_edges)
    for n in range(1, M + 1):
        ret['n'] = n
        ret['edge'] = ga.edge_labels(perm_edges[n - 1])
        ret['max_cluster_size'] = int(mx[n])
        if copy_result:
            ret['moments'] = moments[n].copy()
        else:
            ret['moments'][:] = moments[n]
        if spanning_cluster:
            ret['has_spanning_cluster'] = bool(sp[n])
        if copy_result:
            yield copy.deepcopy(ret)
        else:
            yield ret


def single_run_arrays(spanning_cluster=True, **kwargs):
    r'''Statistics of a single run as arrays (percolate/percolate.py:359-447):
    N, M, max_cluster_size (M+1,), moments (5, M+1) and, with spanning
    detection, has_spanning_cluster (M+1,).  Draws the permutation from
    numpy's global random stream like ``sample_states``.'''
    graph = kwargs.pop('graph')
    model = kwargs.pop('model', 'bond')
    kwargs.pop('copy_result', None)
    if kwargs:
        raise TypeError('unexpected arguments {}'.format(sorted(kwargs)))
    _check_model_and_sides(graph, spanning_cluster, model)
    ga = _graph_arrays(graph, spanning_cluster)
    N, M = ga.num_nodes, ga.num_edges
    perm = np.random.permutation(M)
    ret = {'N': N, 'M': M}
    if M:
        mx, sp, mom = _single_run(ga, perm)
    else:
        mx = np.ones(1, dtype=np.int64)
        sp = np.zeros(1, dtype=bool)
        mom = np.full((1, 5), N - 1.0)
    ret['max_cluster_size'] = mx.astype(np.float64)
    ret['moments'] = mom.T.copy()
    if spanning_cluster:
        ret['has_spanning_cluster'] = sp.astype(bool)
    return ret


def _t_interval(mean, std, runs, alpha):
    """percolate.py:621-633 / 690-703 over arrays: the t interval where
    std > 0, the mean twice where std == 0."""
    mean = np.asarray(mean, dtype=np.float64)
    std = np.asarray(std, dtype=np.float64)
    ci = np.empty(mean.shape + (2,))
    ci[..., 0] = mean
    ci[..., 1] = mean
    nz = std != 0
    if runs > 1 and np.any(nz):
        lo, hi = scipy.stats.t.interval(
            1 - alpha, df=runs - 1, loc=mean[nz], scale=std[nz] / np.sqrt(runs))
        ci[nz, 0] = lo
        ci[nz, 1] = hi
    return ci


def _spanning_stats(k, runs, alpha):
    """percolate.py:557-570: posterior mean (k+1)/(R+2) and the beta
    credible interval, with one quantile call per distinct count k."""
    k = np.asarray(k, dtype=np.float64)
    mean = (k + 1) / (runs + 2)
    ci = np.empty(k.shape + (2,))
    uniq, inv = np.unique(k, return_inverse=True)
    q = scipy.stats.beta.ppf(np.array([alpha / 2, 1 - alpha / 2])[:, None],
                             uniq + 1, runs - uniq + 1)
    ci[..., 0] = q[0][inv].reshape(k.shape)
    ci[..., 1] = q[1][inv].reshape(k.shape)
    return mean, ci


class _MicrocanonicalStats(object):
    """Per-step microcanonical statistics over R runs, computed on the GPU
    from the exact sums, confidence bounds included (pc_microcanonical_ci;
    columns: max, spanning, moments 0..4)."""

    def __init__(self, ga, runs, alpha, spanning_cluster, perms):
        from . import _engine as E
        self.N, self.M = ga.num_nodes, ga.num_edges
        self.runs, self.alpha = runs, alpha
        self.spanning = spanning_cluster
        S = self.M + 1
        self.device_out = None
        if self.M:
            eng, res = _run_batch(ga, perms)
            s1, s2 = eng.microcanonical_sums(res)
            self.device_out = E.micro_ci(s1, s2, self.M, runs, alpha)
            out = self.device_out.cpu().numpy()
            mom = [6, 9, 12, 15, 18]
            self.max_mean = out[0]
            self.max_ci = np.stack([out[1], out[2]], axis=-1)
            self.mom_mean = out[mom]
            self.mom_ci = np.stack([out[[m + 1 for m in mom]],
                                    out[[m + 2 for m in mom]]], axis=-1)
            if spanning_cluster:
                self.span_mean = out[3]
                self.span_ci = np.stack([out[4], out[5]], axis=-1)
            return
        # a graph without edges: the n = 0 state only
        init = np.array([1.0, 0.0] + [N - 1.0 for N in [self.N] * 5])
        mean = init[:, None] * np.ones((1, S))
        std = np.zeros((7, S))
        self.max_mean = mean[0]
        self.max_ci = _t_interval(mean[0], std[0], runs, alpha)
        self.mom_mean = mean[2:7]
        self.mom_ci = _t_interval(mean[2:7], std[2:7], runs, alpha)
        if spanning_cluster:
            self.span_mean, self.span_ci = _spanning_stats(np.zeros(S), runs, alpha)

    def step(self, n, ret):
        ret['n'] = n
        ret['N'] = self.N
        ret['M'] = self.M
        ret['max_cluster_size'] = self.max_mean[n]
        ret['max_cluster_size_ci'] = self.max_ci[n].copy()
        ret['moments'] = self.mom_mean[:, n].copy()
        ret['moments_ci'] = self.mom_ci[:, n].copy()
        if self.spanning:
            ret['spanning_cluster'] = self.span_mean[n]
            ret['spanning_cluster_ci'] = self.span_ci[n].copy()
        return ret

    # rows of pc_microcanonical_ci's output in the column order of
    # canonical_averages over arrays(): max, its bounds, moments 0..4, their
    # (lower, upper) bounds, then spanning and its bounds
    _DEVICE_COLUMNS = [0, 1, 2, 6, 9, 12, 15, 18, 7, 8, 10, 11, 13, 14, 16, 17,
                       19, 20]
    _DEVICE_LAYOUT = [('max_cluster_size', 0, 1), ('max_cluster_size_ci', 1, 2),
                      ('moments', 3, 5), ('moments_ci', 8, 10)]

    def device_columns(self):
        """The (S, C) device matrix canonical_averages builds from arrays(),
        made on the device (the same values: the divisions by N are the same
        IEEE operations), and its column layout; None without edges."""
        if self.device_out is None:
            return None
        rows = list(self._DEVICE_COLUMNS)
        layout = list(self._DEVICE_LAYOUT)
        picked = self.device_out[rows]
        # element-wise true division by a tensor of N (torch turns division
        # by a Python scalar into a multiplication by its reciprocal, which
        # is not the host's numpy division)
        scaled = picked / picked.new_full(picked.shape, float(self.N))
        if self.spanning:
            x = self.device_out.new_empty((len(rows) + 3, picked.shape[1]))
            x[:len(rows)] = scaled
            x[len(rows):] = self.device_out[[3, 4, 5]]
            layout += [('spanning_cluster', 18, 1), ('spanning_cluster_ci', 19, 2)]
        else:
            x = scaled
        return x.t().contiguous(), layout

    def arrays(self):
        """microcanonical_averages_arrays of all steps (percolate.py:968-1064)."""
        ret = {
            'max_cluster_size': self.max_mean / self.N,
            'max_cluster_size_ci': self.max_ci / self.N,
            'moments': self.mom_mean / self.N,
            'moments_ci': self.mom_ci / self.N,
        }
        if self.spanning:
            ret['spanning_cluster'] = self.span_mean.copy()
            ret['spanning_cluster_ci'] = self.span_ci.copy()
        ret['M'] = self.M
        ret['N'] = self.N
        return ret


def _initial_step_stats(N, M, runs, alpha, spanning_cluster):
    """n = 0: every run has max 1, moments N - 1 and no spanning cluster."""
    ret = {'n': 0, 'N': N, 'M': M, 'max_cluster_size': 1.0,
           'max_cluster_size_ci': np.ones(2),
           'moments': np.ones(5) * (N - 1),
           'moments_ci': np.ones((5, 2)) * (N - 1)}
    if spanning_cluster:
        mean, ci = _spanning_stats(np.zeros(1), runs, alpha)
        ret['spanning_cluster'] = mean[0]
        ret['spanning_cluster_ci'] = ci[0]
    return ret


class _MicrocanonicalAverages(object):
    """Iterator of microcanonical_averages (percolate.py:708-896).

    Validation happens at the first ``next()``.  The n = 0 ensemble is
    yielded before any permutation is drawn; at the second ``next()`` the
    permutations of runs 0..R-1 are drawn in order from numpy's global
    random stream (as the reference's lock-step zip over R generators does)
    and all runs are swept and reduced on the GPU at once.
    """

    def __init__(self, graph, runs, spanning_cluster, model, alpha,
                 copy_result):
        self._args = (graph, runs, spanning_cluster, model, alpha)
        self._copy = copy_result
        self._n = None
        self._stats = None
        self._ret = dict()
        self._ga = None

    def __iter__(self):
        return self

    def _start(self):
        graph, runs, spanning_cluster, model, alpha = self._args
        try:
            runs = int(runs)
        except Exception:
            raise ValueError("runs needs to be a positive integer")
        if runs <= 0:
            raise ValueError("runs needs to be a positive integer")
        try:
            alpha = float(alpha)
        except Exception:
            raise ValueError("alpha needs to be a float in the interval (0, 1)")
        if alpha <= 0.0 or alpha >= 1.0:
            raise ValueError("alpha needs to be a float in the interval (0, 1)")
        _check_model_and_sides(graph, spanning_cluster, model)
        self._args = (graph, runs, spanning_cluster, model, alpha)
        self._ga = _graph_arrays(graph, spanning_cluster)

    def _compute(self):
        graph, runs, spanning_cluster, model, alpha = self._args
        ga = self._ga
        perms = np.stack([np.random.permutation(ga.num_edges)
                          for _ in range(runs)])
        self._stats = _MicrocanonicalStats(ga, runs, alpha, spanning_cluster,
                                           perms)

    def __next__(self):
        if self._n is None:
            self._start()
            graph, runs, spanning_cluster, model, alpha = self._args
            ret = _initial_step_stats(self._ga.num_nodes, self._ga.num_edges,
                                      runs, alpha, spanning_cluster)
            self._n = 0
        else:
            if self._n >= self._ga.num_edges:
                raise StopIteration
            if self._stats is None:
                self._compute()
            self._n += 1
            ret = self._stats.step(self._n, dict())
        if self._copy:
            return copy.deepcopy(ret)
        self._ret.clear()
        self._ret.update(ret)
        return self._ret

    next = __next__

    def _all_stats(self):
        """All steps' statistics of a fresh iterator (or None)."""
        if self._n is not None:
            return None
        self._start()
        self._n = 0
        if self._ga.num_edges == 0:
            return None
        self._compute()
        self._n = self._ga.num_edges
        return self._stats

    def _all_arrays(self):
        """Fast path for microcanonical_averages_arrays on a fresh iterator."""
        if self._n is not None:
            return None
        self._start()
        self._n = 0
        if self._ga.num_edges == 0:
            return None
        self._compute()
        self._n = self._ga.num_edges
        return self._stats.arrays()


def microcanonical_averages(
    graph, runs=40, spanning_cluster=True, model='bond', alpha=alpha_1sigma,
    copy_result=True
):
    r'''Generate successive microcanonical percolation ensemble averages
    (percolate/percolate.py:708-896): for n = 0..M the mean and confidence
    interval over ``runs`` runs of max_cluster_size, moments (and the
    Bayesian spanning_cluster estimate).  Raises ValueError for non-positive
    runs or alpha outside (0, 1) at the first ``next()``.'''
    return _MicrocanonicalAverages(graph, runs, spanning_cluster, model,
                                   alpha, copy_result)


# the per-step entries microcanonical_averages_arrays stacks, with the axis
# the steps take in the result (the moments keep their 5 rows first)
_STEP_ARRAYS = (('max_cluster_size', 0), ('max_cluster_size_ci', 0),
                ('spanning_cluster', 0), ('spanning_cluster_ci', 0),
                ('moments', 1), ('moments_ci', 1))


def microcanonical_averages_arrays(microcanonical_averages):
    """Compile microcanonical averages over all steps into arrays, all but
    the spanning statistics divided by N (percolate/percolate.py:968-1064).

    A fresh iterator of this package's microcanonical_averages is compiled
    on the device without visiting the steps; any other iterator of
    per-step dicts is stacked entry by entry."""
    fast = getattr(microcanonical_averages, '_all_arrays', None)
    if fast is not None:
        arrays = fast()
        if arrays is not None:
            return arrays
    steps = list(microcanonical_averages)
    if [st['n'] for st in steps] != list(range(len(steps))):
        raise AssertionError('microcanonical averages out of step order')
    num_nodes, num_edges = steps[0]['N'], steps[0]['M']
    ret = {}
    for key, axis in _STEP_ARRAYS:
        if key not in steps[0]:
            continue
        stacked = np.moveaxis(
            np.stack([np.asarray(st[key], dtype=np.float64) for st in steps]),
            0, axis)
        ret[key] = stacked if key.startswith('spanning') else stacked / num_nodes
    ret['M'] = num_edges
    ret['N'] = num_nodes
    return ret


def _binomial_pmf(n, p):
    """Binomial PMF B(n, k; p) for k = 0..n (percolate/percolate.py:1067-1109),
    computed on the GPU with Loader's saddle-point form for every k and
    normalised to sum 1.  (canonical_averages uses the same weights over the
    window |k - np| <= 12 sigma + 48 only.)"""
    from . import _engine as E
    n = int(n)
    wt = E.binomial_weights(n, [p], window=False)
    ret = np.zeros(n + 1)
    start, length = int(wt.host_start[0]), int(wt.host_length[0])
    ret[start:start + length] = wt.w[0, :length].cpu().numpy()
    return ret


_CANONICAL_KEYS = ('max_cluster_size', 'max_cluster_size_ci',
                   'spanning_cluster', 'spanning_cluster_ci',
                   'moments', 'moments_ci')


def _canonical_columns(arrays):
    """The (S, C) matrix of every microcanonical array, one column per
    convolved series, and its layout [(key, first column, width)]."""
    cols, layout, at = [], [], 0
    for key, value in arrays.items():
        if len(key) <= 1:
            continue
        if key not in _CANONICAL_KEYS:
            raise NotImplementedError('{}-dimensional array'.format(
                np.ndim(value)))
        v = np.asarray(value, dtype=np.float64)
        if key.endswith('_ci'):
            # (S, 2), or moments (5, S, 2) -> (S, 10) as (lower, upper) pairs
            block = v.reshape(-1, 2) if v.ndim == 2 else \
                v.transpose(1, 0, 2).reshape(v.shape[1], -1)
        else:
            block = v.reshape(-1, 1) if v.ndim == 1 else v.T
        layout.append((key, at, block.shape[1]))
        at += block.shape[1]
        cols.append(block)
    return cols, layout


def _canonical_unpack(q, layout, ps, num_sites, num_edges):
    """canonical_averages' result dict from the (P, C) convolved columns."""
    ret = {'ps': ps, 'N': num_sites, 'M': num_edges}
    for key, off, width in layout:
        block = q[:, off:off + width]
        if key == 'moments':
            ret[key] = block.T.copy()
        elif key == 'moments_ci':
            ret[key] = block.reshape(ps.size, 5, 2).transpose(1, 0, 2).copy()
        elif width == 1:
            ret[key] = block[:, 0].copy()
        else:
            ret[key] = block.copy()
    return ret


def _convolve(x, ps):
    """(P, C) host array: the fused FP64 canonical convolution of the
    columns of the (S, C) device matrix x at every p (pc_canonical_fused)."""
    from . import _engine as E
    if ps.size and x.shape[1]:
        return E.canonical_fused(x, ps.reshape(-1), 'fp64_mfma').cpu().numpy()
    return np.zeros((ps.size, x.shape[1]))


def canonical_averages(ps, microcanonical_averages_arrays):
    """Canonical cluster statistics from microcanonical ones
    (percolate/percolate.py:1112-1223): every array (means and confidence
    bounds) convolved with the binomial weights of each p, on the GPU in
    FP64 with the weights generated inside the contraction
    (pc_canonical_fused)."""
    from . import _engine as E
    arrays = microcanonical_averages_arrays
    ps = np.asarray(ps)
    cols, layout = _canonical_columns(arrays)
    S = arrays['M'] + 1
    x = np.ascontiguousarray(np.concatenate(cols, axis=1)) if cols else \
        np.zeros((S, 0))
    torch = E.nat.torch_cuda()
    q = _convolve(torch.from_numpy(x).to(torch.device('cuda')), ps)
    return _canonical_unpack(q, layout, ps, arrays['N'], arrays['M'])


def statistics(
    graph, ps, spanning_cluster=True, model='bond', alpha=alpha_1sigma, runs=40
):
    """Canonical cluster statistics of a graph at every p
    (percolate/percolate.py:1226-1252): the microcanonical averages over
    ``runs`` runs convolved with the binomial weights.  The microcanonical
    statistics stay on the device between the two stages (the per-step
    arrays are never built on the host); the results equal those of
    microcanonical_averages -> microcanonical_averages_arrays ->
    canonical_averages bit for bit."""
    averages = microcanonical_averages(
        graph=graph, runs=runs, spanning_cluster=spanning_cluster, model=model,
        alpha=alpha)
    ps = np.asarray(ps)
    stats = averages._all_stats()
    cols = None if stats is None else stats.device_columns()
    if cols is None:
        return canonical_averages(ps, microcanonical_averages_arrays(averages))
    x, layout = cols
    return _canonical_unpack(_convolve(x, ps), layout, ps, stats.N, stats.M)
