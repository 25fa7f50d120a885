# encoding: utf-8
"""Drop-in for ``percolate.hpc`` (andsor/pypercolate 0.4.6,
percolate/hpc.py): the Newman-Ziff routines for HPC use, with the sweep and
the convolution on the GPU.

The structured-array layouts are the reference's (hpc.py:22-70, 407-440,
518-558, 705-749) so results interoperate with code written against it.
Integer outputs (max_cluster_size, has_spanning_cluster, moments modulo
2**64) are bit-identical to the reference given the same edge order and
numpy RandomState permutation.  Extensions for batches of runs carry a
``_batch`` suffix.
"""

import numpy as np

from . import _graph


def _ndarray_dtype(fields):
    """NumPy structured dtype from (name, type) pairs (hpc.py:22-31).  The
    names are plain str: they compare equal to the reference's np.str_
    names, and numpy 2 would otherwise keep np.str_ objects in the dtype,
    whose repr makes .npy files unreadable without pickle."""
    return [(str(key), values) for key, values in fields]


def microcanonical_statistics_dtype(spanning_cluster=True):
    """Dtype of sample states (hpc.py:34-70)."""
    fields = [('n', 'uint32'), ('edge', 'uint32')]
    if spanning_cluster:
        fields.append(('has_spanning_cluster', 'bool'))
    fields.extend([('max_cluster_size', 'uint32'), ('moments', '(5,)uint64')])
    return _ndarray_dtype(fields)


def canonical_statistics_dtype(spanning_cluster=True):
    """Dtype of the canonical statistics of one run (hpc.py:407-440)."""
    fields = []
    if spanning_cluster:
        fields.append(('percolation_probability', 'float64'))
    fields.extend([('max_cluster_size', 'float64'), ('moments', '(5,)float64')])
    return _ndarray_dtype(fields)


def canonical_averages_dtype(spanning_cluster=True):
    """Dtype of canonical averages over runs (hpc.py:518-558)."""
    fields = [('number_of_runs', 'uint32')]
    if spanning_cluster:
        fields.extend([('percolation_probability_mean', 'float64'),
                       ('percolation_probability_m2', 'float64')])
    fields.extend([('max_cluster_size_mean', 'float64'),
                   ('max_cluster_size_m2', 'float64'),
                   ('moments_mean', '(5,)float64'),
                   ('moments_m2', '(5,)float64')])
    return _ndarray_dtype(fields)


def finalized_canonical_averages_dtype(spanning_cluster=True):
    """Dtype of finalized canonical averages (hpc.py:705-749)."""
    fields = [('number_of_runs', 'uint32'), ('p', 'float64'), ('alpha', 'float64')]
    if spanning_cluster:
        fields.extend([('percolation_probability_mean', 'float64'),
                       ('percolation_probability_std', 'float64'),
                       ('percolation_probability_ci', '(2,)float64')])
    fields.extend([('percolation_strength_mean', 'float64'),
                   ('percolation_strength_std', 'float64'),
                   ('percolation_strength_ci', '(2,)float64'),
                   ('moments_mean', '(5,)float64'),
                   ('moments_std', '(5,)float64'),
                   ('moments_ci', '(5,2)float64')])
    return _ndarray_dtype(fields)


def _graph_arrays(perc_graph, spanning_cluster, auxiliary_node_attributes,
                  auxiliary_edge_attributes, spanning_sides):
    return _graph.from_percolation_graph(
        perc_graph, spanning_cluster=spanning_cluster,
        auxiliary_node_attributes=auxiliary_node_attributes,
        auxiliary_edge_attributes=auxiliary_edge_attributes,
        spanning_sides=spanning_sides)


def _check_sides(spanning_cluster, spanning_sides):
    if spanning_cluster and len(spanning_sides) != 2:
        raise ValueError(
            'Spanning cluster is to be detected, but auxiliary nodes '
            'of less or more than 2 types (sides) given.'
        )


def _traces_to_structured(ga, perms, mx, sp, mom, spanning_cluster):
    """(R, S) device-trace copies -> (R, S) structured arrays."""
    R, S = mx.shape
    out = np.empty((R, S), dtype=microcanonical_statistics_dtype(spanning_cluster))
    out['n'] = np.arange(S, dtype=np.uint32)[np.newaxis, :]
    out['edge'][:, 0] = 0
    if S > 1:
        out['edge'][:, 1:] = perms
    if spanning_cluster:
        out['has_spanning_cluster'] = sp
    out['max_cluster_size'] = mx
    out['moments'] = mom
    return out


def _int_seeds(seeds):
    """The seeds as uint32 values when every one is an integer in
    [0, 2**32) (RandomState's integer seeding), else None."""
    out = []
    for s in seeds:
        if isinstance(s, (bool, np.bool_)) or not isinstance(s, (int, np.integer)):
            return None
        if s < 0 or s > 0xFFFFFFFF:
            return None
        out.append(int(s))
    return out


def _permutations(engine, seeds):
    """Device tensor of RandomState(seed).permutation(M) for every seed:
    drawn on the GPU for integer seeds, by numpy on the host otherwise."""
    ints = _int_seeds(seeds)
    if ints is not None:
        return engine.randomstate_permutations(ints)
    perms = np.stack([np.random.RandomState(seed=s).permutation(engine.M)
                      for s in seeds])
    return engine.upload_permutations(perms)


def _sweep_host_perms(ga, perms, spanning_cluster):
    """Run the GPU sweep for (R, M) permutations; structured (R, S) array."""
    from ._engine import BondEngine
    R = perms.shape[0]
    S = ga.num_edges + 1
    if ga.num_edges == 0:
        mx = np.ones((R, 1), dtype=np.uint32)
        sp = np.zeros((R, 1), dtype=bool)
        mom = np.full((R, 1, 5), ga.num_nodes - 1, dtype=np.uint64)
        return _traces_to_structured(ga, perms, mx, sp, mom, spanning_cluster)
    eng = BondEngine(ga)
    res = eng.sweep(eng.upload_permutations(perms))
    mx, sp, mom = eng.host_traces(res)
    assert mx.shape == (R, S)
    return _traces_to_structured(ga, perms, mx, sp, mom, spanning_cluster)


def bond_sample_states(
    perc_graph, num_nodes, num_edges, seed, spanning_cluster=True,
    auxiliary_node_attributes=None, auxiliary_edge_attributes=None,
    spanning_sides=None,
    **kwargs
):
    '''Generate successive sample states of the bond percolation model
    (hpc.py:73-307).

    Yields one structured row (dtype microcanonical_statistics_dtype) per
    occupation number n = 0..M; as in the reference the same array object
    is returned every time and overwritten by the next step.  The run's
    permutation is numpy's RandomState(seed).permutation(num_edges); the run
    itself is swept on the GPU at the second ``next()``.
    '''
    rng = np.random.RandomState(seed=seed)
    if spanning_cluster:
        _check_sides(spanning_cluster, spanning_sides)
    perm_edges = rng.permutation(num_edges)

    ret = np.empty(1, dtype=microcanonical_statistics_dtype(spanning_cluster))
    ret['n'] = 0
    ret['max_cluster_size'] = 1
    ret['moments'] = np.ones(5, dtype='uint64') * (num_nodes - 1)
    if spanning_cluster:
        ret['has_spanning_cluster'] = False
    yield ret

    ga = _graph_arrays(perc_graph, spanning_cluster, auxiliary_node_attributes,
                       auxiliary_edge_attributes, spanning_sides)
    if ga.num_nodes != num_nodes or ga.num_edges != num_edges:
        raise ValueError('num_nodes / num_edges do not match perc_graph')
    trace = _sweep_host_perms(ga, perm_edges[np.newaxis, :], spanning_cluster)[0]
    for n in range(1, num_edges + 1):
        ret[0] = trace[n]
        yield ret


def bond_microcanonical_statistics(
    perc_graph, num_nodes, num_edges, seed,
    spanning_cluster=True,
    auxiliary_node_attributes=None, auxiliary_edge_attributes=None,
    spanning_sides=None,
    **kwargs
):
    """Evolve a single run over all microstates (hpc.py:310-404): structured
    array of num_edges + 1 rows.  ``ret['edge'][0]`` is 0 (undefined in the
    reference)."""
    return bond_microcanonical_statistics_batch(
        perc_graph, num_nodes, num_edges, [seed], spanning_cluster,
        auxiliary_node_attributes, auxiliary_edge_attributes, spanning_sides,
    )[0]


def bond_microcanonical_statistics_batch(
    perc_graph, num_nodes, num_edges, seeds,
    spanning_cluster=True,
    auxiliary_node_attributes=None, auxiliary_edge_attributes=None,
    spanning_sides=None,
    **kwargs
):
    """``bond_microcanonical_statistics`` for many seeds in one GPU sweep:
    structured array (len(seeds), num_edges + 1)."""
    _check_sides(spanning_cluster, spanning_sides)
    seeds = list(seeds)
    ga = _graph_arrays(perc_graph, spanning_cluster, auxiliary_node_attributes,
                       auxiliary_edge_attributes, spanning_sides)
    if ga.num_nodes != num_nodes or ga.num_edges != num_edges:
        raise ValueError('num_nodes / num_edges do not match perc_graph')
    if not seeds or num_edges == 0:
        perms = np.zeros((len(seeds), num_edges), dtype=np.int64)
        return _sweep_host_perms(ga, perms, spanning_cluster)
    from ._engine import BondEngine
    eng = BondEngine(ga)
    d_perm = _permutations(eng, seeds)
    res = eng.sweep(d_perm)
    mx, sp, mom = eng.host_traces(res)
    return _traces_to_structured(ga, d_perm.cpu().numpy(), mx, sp, mom,
                                 spanning_cluster)


def _trace_columns(microcanonical_statistics):
    """(S, C) float64 columns of a run: [spanning,] max, moments 0..4."""
    ms = microcanonical_statistics
    cols = []
    if 'has_spanning_cluster' in ms.dtype.names:
        cols.append(ms['has_spanning_cluster'].astype(np.float64)[:, None])
    cols.append(ms['max_cluster_size'].astype(np.float64)[:, None])
    cols.append(ms['moments'].astype(np.float64))
    return np.ascontiguousarray(np.concatenate(cols, axis=1))


def bond_canonical_statistics(
    microcanonical_statistics,
    convolution_factors,
    **kwargs
):
    """Canonical cluster statistics for a single run and a single p
    (hpc.py:443-515): the convolution of the run's observables with the
    factors, in FP64 on the GPU."""
    import torch
    from . import _engine as E
    spanning_cluster = (
        'has_spanning_cluster' in microcanonical_statistics.dtype.names
    )
    ret = np.empty(1, dtype=canonical_statistics_dtype(spanning_cluster))
    x = _trace_columns(microcanonical_statistics)
    f = np.asarray(convolution_factors, dtype=np.float64).reshape(-1)
    if f.size != x.shape[0]:
        raise ValueError('convolution_factors must have one entry per step')
    wt = E.dense_weights(f)
    q = E.contract(torch.from_numpy(x).to(wt.w.device), wt).cpu().numpy()[0]
    off = 0
    if spanning_cluster:
        ret['percolation_probability'] = q[0]
        off = 1
    ret['max_cluster_size'] = q[off]
    ret['moments'] = q[off + 1:off + 6]
    return ret


# The averaged statistics: (field of the canonical statistics, name in the
# finalized averages, divided by the number of nodes when finalized).
_STATISTICS = (('percolation_probability', 'percolation_probability', False),
               ('max_cluster_size', 'percolation_strength', True),
               ('moments', 'moments', True))


def _statistics_in(names, suffixes=('',)):
    """The entries of _STATISTICS whose fields (with every suffix) are all
    among the dtype field names."""
    return [st for st in _STATISTICS
            if all(st[0] + sfx in names for sfx in suffixes)]


def _per_row(values, like):
    """A per-row (per-p) array as float64, shaped to broadcast against a
    field with more axes (the moments' 5)."""
    v = np.asarray(values, dtype=np.float64)
    return v.reshape(v.shape + (1,) * (np.ndim(like) - v.ndim))


def bond_initialize_canonical_averages(
    canonical_statistics, **kwargs
):
    """Canonical averages of a single run (hpc.py:561-635): number_of_runs 1,
    every statistic its own mean, with zero sum of squared deviations.
    Spanning is detected from the field names, as in the reference."""
    names = canonical_statistics.dtype.names
    ret = np.empty_like(
        canonical_statistics,
        dtype=canonical_averages_dtype(
            spanning_cluster='percolation_probability' in names),
    )
    ret['number_of_runs'] = 1
    for field, _, _ in _statistics_in(names):
        ret[field + '_mean'] = canonical_statistics[field]
        ret[field + '_m2'] = 0.0
    return ret


def _online_variance(a, b):
    """Pairwise merge of (n, mean, M2) (Chan, Golub & LeVeque), the
    algorithm behind the reference's simoa.stats.online_variance."""
    n_a, mean_a, m2_a = a
    n_b, mean_b, m2_b = b
    n_a = np.asarray(n_a, dtype=np.float64)
    n_b = np.asarray(n_b, dtype=np.float64)
    n = n_a + n_b
    delta = mean_b - mean_a
    mean = mean_a + delta * (n_b / n)
    m2 = m2_a + m2_b + delta * delta * (n_a * n_b / n)
    return n, mean, m2


def bond_reduce(row_a, row_b):
    """Reduce the canonical averages over several runs (hpc.py:638-702):
    associative and commutative merge of (n, mean, M2) for every statistic
    whose mean and M2 both rows carry."""
    both = [st for st in _statistics_in(row_a.dtype.names, ('_mean', '_m2'))
            if st in _statistics_in(row_b.dtype.names, ('_mean', '_m2'))]
    ret = np.empty_like(row_a)
    for field, _, _ in both:
        mean, m2 = field + '_mean', field + '_m2'
        _, ret[mean], ret[m2] = _online_variance(
            (_per_row(row_a['number_of_runs'], row_a[mean]), row_a[mean], row_a[m2]),
            (_per_row(row_b['number_of_runs'], row_b[mean]), row_b[mean], row_b[m2]))
    ret['number_of_runs'] = row_a['number_of_runs'] + row_b['number_of_runs']
    return ret


def finalize_canonical_averages(
    number_of_nodes, ps, canonical_averages, alpha,
):
    """Finalize canonical averages (hpc.py:752-834): per statistic the
    mean, the sample standard deviation sqrt(M2 / (n - 1)) and the Student-t
    confidence interval of the mean; max_cluster_size becomes the
    percolation strength, and it and the moments are divided by the number
    of nodes.  Computed on the device (pc_finalize_canonical); scipy supplies
    the t quantiles of the distinct run counts."""
    from . import _engine as E
    torch = E.nat.torch_cuda()
    names = canonical_averages.dtype.names
    present = _statistics_in(names, ('_mean', '_m2'))
    spanning = any(st[0] == 'percolation_probability' for st in present)
    ret = np.empty_like(
        canonical_averages,
        dtype=finalized_canonical_averages_dtype(spanning_cluster=spanning),
    )
    runs = canonical_averages['number_of_runs']
    ret['number_of_runs'] = runs
    ret['p'] = ps
    ret['alpha'] = alpha
    rows = int(np.prod(ret.shape))
    flat = canonical_averages.reshape(rows)
    widths = [int(np.prod(flat.dtype[field + '_mean'].shape) or 1)
              for field, _, _ in present]
    mean = np.concatenate([np.asarray(flat[f + '_mean'], np.float64).reshape(rows, -1)
                           for f, _, _ in present], axis=1)
    m2 = np.concatenate([np.asarray(flat[f + '_m2'], np.float64).reshape(rows, -1)
                         for f, _, _ in present], axis=1)
    div = np.concatenate([np.full(w, float(number_of_nodes) if per_node else 1.0)
                          for (_, _, per_node), w in zip(present, widths)])
    dev = torch.device('cuda', torch.cuda.current_device())
    out = E.finalize_columns(np.asarray(runs).reshape(rows),
                             torch.from_numpy(mean).to(dev),
                             torch.from_numpy(m2).to(dev), div,
                             alpha).cpu().numpy()
    off = 0
    for (_, final, _), w in zip(present, widths):
        shape = ret.shape + ret.dtype[final + '_mean'].shape
        sl = slice(off, off + w)
        ret[final + '_mean'] = out[0][:, sl].reshape(shape)
        ret[final + '_std'] = out[1][:, sl].reshape(shape)
        ret[final + '_ci'] = np.stack([out[2][:, sl], out[3][:, sl]],
                                      axis=-1).reshape(shape + (2,))
        off += w
    return ret


def bond_canonical_averages_batch(
    perc_graph, num_nodes, num_edges, ps, seeds=None, num_runs=None,
    philox_seed=0, spanning_cluster=True, auxiliary_node_attributes=None,
    auxiliary_edge_attributes=None, spanning_sides=None, run_chunk=None,
    graph_arrays=None, **kwargs
):
    """Canonical averages over many runs in batched GPU launches: the result
    of the reference's jug study for one system size,
    ``reduce(bond_reduce, [bond_initialize_canonical_averages(
    [bond_canonical_statistics(bond_microcanonical_statistics(seed), f_p)
    for p in ps]) for seed in seeds])`` (percolate/share/jugfile.py:57-135,
    percolate/hpc.py:443-702).

    With ``seeds`` the runs use numpy RandomState(seed).permutation(M)
    (bit-exact traces; drawn on the GPU for integer seeds); otherwise ``num_runs`` runs draw device permutations
    keyed by ``philox_seed``.  Per chunk of runs: sweep, per-run observables
    over the union of the convolution windows, one FP64 matrix-core
    contraction of all runs' columns with the binomial weights, and the
    per-(p, field) mean and M2 over the chunk's runs; chunks merge with the
    Chan formula.  ``run_chunk`` (default: three sweep residencies of the
    GPU, 3072 runs at L = 128) is capped so that one chunk's columns take
    at most 1 GiB.  ``graph_arrays`` (a ``_graph.GraphArrays`` of the same
    graph, e.g. ``_graph.square_lattice(L)``) skips reading the networkx
    graph, which is most of the host time of a call at L = 128.  Returns an array of canonical_averages_dtype (one row per
    p), ready for finalize_canonical_averages.
    """
    import torch
    from . import _engine as E
    from ._engine import BondEngine, ContractPlan, Weights

    _check_sides(spanning_cluster, spanning_sides)
    if graph_arrays is not None:
        ga = graph_arrays
        if ga.spanning != bool(spanning_cluster):
            raise ValueError('graph_arrays and spanning_cluster disagree on '
                             'spanning sides')
    else:
        ga = _graph_arrays(perc_graph, spanning_cluster,
                           auxiliary_node_attributes,
                           auxiliary_edge_attributes, spanning_sides)
    if ga.num_nodes != num_nodes or ga.num_edges != num_edges:
        raise ValueError('num_nodes / num_edges do not match perc_graph')
    if ga.num_edges == 0:
        raise ValueError('the percolation graph has no edges')
    ps = np.atleast_1d(np.asarray(ps, dtype=np.float64))
    if seeds is not None:
        seeds = list(seeds)
        total = len(seeds)
    else:
        total = int(num_runs)
    if total < 1:
        raise ValueError('at least one run is needed')
    cols = 7 if spanning_cluster else 6
    eng = BondEngine(ga)
    M = eng.M
    wt = E.binomial_weights(M, ps)
    g_lo = int(wt.host_start.min())
    g_hi = int((wt.host_start + wt.host_length).max())
    g_len = g_hi - g_lo
    shifted = Weights(M, ps, wt.start - g_lo, wt.length, wt.kmax, wt.w)
    shifted.host_start = wt.host_start - g_lo
    shifted.host_length = wt.host_length
    if run_chunk is None:
        cus = torch.cuda.get_device_properties(eng.device).multi_processor_count
        run_chunk = 3 * max(eng.occupancy(), 1) * cus
    chunk = max(1, min(int(run_chunk), (1 << 30) // (g_len * cols * 8)))

    n_acc = None
    for first in range(0, total, chunk):
        rc = min(chunk, total - first)
        if seeds is not None:
            d_perm = _permutations(eng, seeds[first:first + rc])
        else:
            d_perm = eng.device_permutations(philox_seed, rc, first_run=first)
        res = eng.sweep(d_perm)
        x = torch.empty((g_len, rc * cols), dtype=torch.float64,
                        device=eng.device)
        lib = eng.lib
        from . import _native as nat
        nat.check(lib.pc_window_traces(
            eng.N, M, rc, nat.ptr(res.records), max(M, 1),
            nat.ptr(res.span_step), nat.ptr(res.chunk_state), g_lo, g_len,
            1 if spanning_cluster else 0, nat.ptr(x),
            nat.stream_handle(torch)), 'pc_window_traces')
        q = ContractPlan(x, shifted, nat.PC_PREC_FP64_MFMA).run()
        mean = torch.empty((ps.size, cols), dtype=torch.float64, device=eng.device)
        m2 = torch.empty_like(mean)
        nat.check(lib.pc_run_moments(nat.ptr(q), ps.size, rc, cols,
                                     nat.ptr(mean), nat.ptr(m2),
                                     nat.stream_handle(torch)), 'pc_run_moments')
        if n_acc is None:
            n_acc, mean_acc, m2_acc = float(rc), mean, m2
        else:
            # Chan merge on the device, into the accumulators
            n_a = torch.full((ps.size,), n_acc, dtype=torch.float64,
                             device=eng.device)
            n_b = torch.full((ps.size,), float(rc), dtype=torch.float64,
                             device=eng.device)
            E.merge_moments(n_a, mean_acc, m2_acc, n_b, mean, m2,
                            out=(mean_acc, m2_acc))
            n_acc += float(rc)
        del x, q, res, d_perm
    mean_acc = mean_acc.cpu().numpy()
    m2_acc = m2_acc.cpu().numpy()

    ret = np.empty(ps.size, dtype=canonical_averages_dtype(spanning_cluster))
    ret['number_of_runs'] = total
    off = 0
    if spanning_cluster:
        ret['percolation_probability_mean'] = mean_acc[:, 0]
        ret['percolation_probability_m2'] = m2_acc[:, 0]
        off = 1
    ret['max_cluster_size_mean'] = mean_acc[:, off]
    ret['max_cluster_size_m2'] = m2_acc[:, off]
    ret['moments_mean'] = mean_acc[:, off + 1:off + 6]
    ret['moments_m2'] = m2_acc[:, off + 1:off + 6]
    return ret
