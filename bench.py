"""Throughput of the Newman-Ziff bond-percolation hot path on one MI355X.

Workload (default): bond percolation on the L = 128 open square lattice
(spanning_2d_grid(128): N = 16384 nodes, M = 32512 bonds, top/bottom
spanning sides), R = 10^4 runs per step.  One step is one pass of the hot
path over the batch: device permutations (Philox seed 1234) -> union-find
sweep -> per-run chunk states -> exact microcanonical sums over runs for
every step n = 0..M.  Inputs (the lattice) are resident in HBM before the
timed region.  The metric is bond additions per second = R * M / time.

Output: one JSON line.  ``roofline`` describes the sweep kernel (the
dominant kernel), timed with HIP events on the stream it runs on;
``cpu_baseline`` times the CPU oracle (the C restatement of
percolate/hpc.py:194-307 plus the exact per-step sums) on a bounded sample
of the same workload on this host's cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = ('bond additions/sec at L=128 + canonical-avg speedup FP32-MFMA vs '
          'FP64 @1e-9')
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per bond addition in the LDS-state sweep: the 4-byte
# permutation read from HBM plus six 2-byte LDS accesses (two parent reads
# for the finds, two root-size reads, one parent and one size write)
ALG_BYTES_PER_ADD_LDS = 16.0
# the same six accesses with the state in HBM (4-byte codes)
ALG_BYTES_PER_ADD_HBM = 28.0


def parse_args(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split('\n\n')[0])
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--length', type=int, default=128)
    ap.add_argument('--runs', type=int, default=10000)
    ap.add_argument('--seed', type=int, default=1234)
    ap.add_argument('--cpu-sample-runs', type=int, default=49152,
                    help='runs in the CPU-oracle sample (0 skips it); the '
                         'default is about 10-20 s of work on 16 host cores')
    ap.add_argument('--copy-calibration', action='store_true',
                    help='also run a 1 GiB device copy (for PMC scaling)')
    ap.add_argument('--canonical-length', type=int, default=1024,
                    help='lattice of the large-lattice run whose exact sums '
                         'feed the canonical-average comparison (0 skips both)')
    ap.add_argument('--canonical-runs', type=int, default=1000)
    ap.add_argument('--canonical-ps', type=int, default=4096)
    ap.add_argument('--tune', action='store_true',
                    help='pick the sweep knobs with the autotuner first (the '
                         'default knobs are within 1%% of the tuned ones at '
                         'L = 128; the finite-size batch always tunes)')
    ap.add_argument('--no-bitexact', action='store_true',
                    help='skip the timing of the pipeline with bit-exact '
                         'device RandomState permutations')
    ap.add_argument('--per-run', action='store_true',
                    help='also time the per-run (hpc) flavour of the canonical '
                         'averages at L = 128: 10^4 runs, 1000 p')
    ap.add_argument('--batch', action='store_true',
                    help='time the finite-size batch (L = 16..512, 2^26/L^2 '
                         'runs each, one HIP graph) instead of L = 128')
    return ap.parse_args(argv)


def profiled_traffic_per_add(length, runs):
    """HBM bytes per bond addition of the sweep from the committed PMC
    summary of this workload (profiles/sweep_L<L>_R<R>.md), or None."""
    path = os.path.join(ROOT, 'profiles', 'sweep_L{}_R{}.md'.format(length, runs))
    try:
        for line in open(path):
            if line.startswith('sweep_traffic_bytes_per_add:'):
                return float(line.split(':', 1)[1])
    except OSError:
        pass
    return None


def large_lattice(length, runs, seed, steps=2, warmup=1):
    """The L = 1024 configuration: spanning_2d_grid(1024) (N = 1,048,576,
    M = 2,095,104), 1000 runs from Philox seed 7, union-find state in the
    HBM workspace (32-bit codes), 64-bit merge records, the whole per-step
    trace reduced on the device to exact sums (moments k = 3, 4 and all
    squares in 128/192-bit).  Returns the timing and the engine and exact
    sums of the last step (the input of the canonical comparison)."""
    import torch
    from pypercolate_amd import _graph
    from pypercolate_amd._engine import BondEngine, Pipeline
    eng = BondEngine(_graph.square_lattice(length))
    pipe = Pipeline(eng, runs, seed=seed)
    for _ in range(warmup):
        pipe.run()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        pipe.run()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    sweep = pipe.sweep_ms()
    s1 = pipe.sum1.cpu().numpy().view('uint64')
    ok = int(s1[0, eng.M, 0]) == runs * eng.N and int(s1[1, eng.M, 0]) == runs
    adds = float(runs) * eng.M
    info = {
        'workload': ('bond percolation, L={} square lattice (N={}, M={}), {} '
                     'runs from Philox seed {}: permute -> sweep (HBM state) '
                     '-> chunk states -> exact microcanonical sums').format(
                         length, eng.N, eng.M, runs, seed),
        'value': adds / (ms / 1e3), 'unit': 'bond additions/s',
        'ms_per_step': ms, 'steps': steps, 'warmup': warmup,
        'sweep_ms_per_step': sweep, 'sweep_launches': len(pipe.sweep_events),
        'sweep_adds_per_s': adds / (sweep / 1e3),
        'roofline_sweep': {'bound': 'hbm', 'alg_bytes_per_add': ALG_BYTES_PER_ADD_HBM,
                           'achieved': ALG_BYTES_PER_ADD_HBM * adds / (sweep / 1e3) / 1e9,
                           'peak': HBM_PEAK_GBS, 'unit': 'GB/s'},
        'sums_check': ok,
    }
    info['roofline_sweep']['frac'] = info['roofline_sweep']['achieved'] / HBM_PEAK_GBS
    per_add = profiled_traffic_per_add(length, runs)
    info['roofline_sweep']['traffic'] = (per_add * adds if per_add is not None
                                         else None)
    info['roofline_sweep']['edges'] = ('computed from the edge index'
                                       if eng.lattice_sweep else 'edge array')
    sums = (pipe.sum1, pipe.sum2)
    del pipe
    torch.cuda.empty_cache()
    return info, eng, sums, runs


def canonical_comparison(eng, sums, runs, num_p, reps=5):
    """The canonical average at every precision on real microcanonical data:
    the exact sums of `runs` runs on eng's lattice, 14 columns (mean and
    mean square of the 7 observables per step), num_p values of p in
    [0.3, 0.7].

    Forms: the plain FP64 GPU form (weight kernel, then the FP64 contraction
    on the vector ALUs or on the FP64 matrix cores), the FP32-input forms
    (weights and columns converted to FP32), and the fused forms that
    generate the weights inside the contraction (FP64 matrix cores, exact
    bf16 byte split).  Per form: median time of the contraction alone and of
    the whole canonical step (weights + conversions + contraction; HIP
    events), and the largest error per column against the plain FP64 VALU
    form, relative to sum_k w_k |x_k| (which is |Q| for these non-negative
    columns) floored at 1e-6 of that column's largest value, so results
    that only come from the far tails of the weights (spanning at p = 0.3)
    are compared absolutely at the floor.  A non-finite result (FP32
    overflow of the squared fourth moment, ~2^154) is reported as null."""
    import numpy as np
    import torch
    from pypercolate_amd import _graph, _native as nat
    from pypercolate_amd import _engine as E

    s1, s2 = sums
    mean, std = E.micro_moments(s1, s2, eng.M, runs)
    second = std * std * ((runs - 1.0) / runs) + mean * mean
    x = torch.cat([mean, second], dim=0).t().contiguous()   # (S, 14)
    ps = np.linspace(0.3, 0.7, num_p)
    lib = nat.load_library()
    stream = nat.stream_handle(torch)
    wt = E.binomial_weights(eng.M, ps)
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)

    def timed(fn):
        fn()
        times = []
        for _ in range(reps):
            t0.record()
            fn()
            t1.record()
            torch.cuda.synchronize()
            times.append(t0.elapsed_time(t1))
        return sorted(times)[len(times) // 2]

    def weights():
        nat.check(lib.pc_binomial_weights(eng.M, num_p, nat.ptr(wt_ps),
                                          nat.ptr(wt.start), nat.ptr(wt.length),
                                          wt.kmax, nat.ptr(wt.w), stream),
                  'pc_binomial_weights')
    wt_ps = torch.from_numpy(ps.copy()).cuda()
    weights_ms = timed(weights)
    wf32 = E.to_f32(wt.w)
    xf32 = E.to_f32(x)

    def convert():
        nat.check(lib.pc_convert_f64_to_f32(nat.ptr(wt.w), wt.w.numel(),
                                            nat.ptr(wf32), stream), 'convert')
        nat.check(lib.pc_convert_f64_to_f32(nat.ptr(x), x.numel(),
                                            nat.ptr(xf32), stream), 'convert')
    convert_ms = timed(convert)
    lens = wt.host_length.astype(np.float64)
    flops = 2.0 * lens.sum() * x.shape[1]
    mag = E.ContractPlan(x.abs(), wt, 'fp64').run().cpu().numpy()
    floor = 1e-6 * mag.max(axis=0, keepdims=True)
    scale = np.maximum(mag, np.where(floor > 0, floor, 1.0))
    out = {'workload': ('L={} (M={}), {} runs, {} p in [0.3, 0.7], 14 columns, '
                        'windows |k - Mp| <= 12 sigma + 48 (mean {:.0f} steps)'
                        ).format(int(round(eng.N ** 0.5)), eng.M, runs, num_p,
                                 lens.mean()),
           'weights_ms': weights_ms, 'convert_f32_ms': convert_ms,
           'weight_evaluations': float(lens.sum()),
           'algorithmic_flops': flops, 'forms': {}}
    ref = None
    plans = [(name, E.ContractPlan(x, wt, name), False)
             for name in ('fp64', 'fp64_mfma', 'fp32_mfma', 'bf16x3_mfma')]
    plans += [('fused_' + name, E.FusedPlan(x, ps, name), True)
              for name in ('fp64_mfma', 'split_bf16')]
    for name, plan, fused in plans:
        ms = timed(plan.launch if fused else plan.run)
        q = plan.run().cpu().numpy()
        if ref is None:
            ref = q
        with np.errstate(invalid='ignore', over='ignore'):
            err = (np.abs(q - ref) / scale).max(axis=0)
        finite = np.isfinite(err)
        total = ms if fused else ms + weights_ms + (
            convert_ms if name in ('fp32_mfma', 'bf16x3_mfma') else 0.0)
        out['forms'][name] = {
            'contract_ms': ms, 'canonical_ms': total,
            'tflops': flops / ms / 1e9,
            'max_rel_err_per_column': [float(e) if f else None
                                       for e, f in zip(err, finite)],
            'max_rel_err': float(err.max()) if finite.all() else None,
            'meets_1e-9': bool(finite.all() and err.max() <= 1e-9)}
    base = out['forms']['fp64']['canonical_ms']
    base_mfma = out['forms']['fp64_mfma']['canonical_ms']
    for name, f in out['forms'].items():
        f['speedup_vs_plain_fp64'] = base / f['canonical_ms']
        f['speedup_vs_plain_fp64_mfma'] = base_mfma / f['canonical_ms']
    ok = [(f['canonical_ms'], name) for name, f in out['forms'].items()
          if f['meets_1e-9']]
    best = min(ok)[1]
    out['best_at_1e-9'] = {'form': best,
                           # the comparator of the metric: the plain FP64
                           # form on the FP64 matrix cores (weight kernel +
                           # FP64 MFMA contraction)
                           'speedup': out['forms'][best]['speedup_vs_plain_fp64_mfma'],
                           'speedup_vs_plain_fp64_mfma':
                               out['forms'][best]['speedup_vs_plain_fp64_mfma'],
                           'speedup_vs_plain_fp64_valu':
                               out['forms'][best]['speedup_vs_plain_fp64']}
    return out


def per_run_canonical(length, runs, num_p, seed, warmup=1, reps=5):
    """The per-run (hpc) flavour of the canonical averages at L = 128:
    hpc.bond_canonical_averages_batch over `runs` device-permutation runs
    (Philox `seed`), num_p values of p in [0.45, 0.55]: per chunk of runs,
    sweep -> per-run observables over the union of the windows -> FP64
    matrix-core contraction of every run's 7 columns -> mean and M2 over
    the chunk's runs -> Chan merge.  Reports the whole call (wall clock,
    host merges included) and, on its own, the contraction of one chunk
    (HIP events, median of `reps`) against the FP64 matrix peak, with the
    algorithmic flops 2 * sum_p |window_p| * 7 * runs of the chunk.
    ``graph_ingest_ms`` is the host share of each call that reads the
    networkx graph into index arrays; ``ms_with_graph_arrays`` is the same
    call given the closed-form arrays of the lattice instead."""
    import torch
    from pypercolate_amd import hpc, percolate
    from pypercolate_amd import _engine as E
    from pypercolate_amd._engine import Weights
    pg = percolate.percolation_graph(percolate.spanning_2d_grid(length),
                                     spanning_cluster=True)
    ps = np.linspace(0.45, 0.55, num_p)
    # host share of every call: the networkx graph -> index arrays pass
    t = time.perf_counter()
    hpc._graph_arrays(pg['perc_graph'], True, pg['auxiliary_node_attributes'],
                      pg['auxiliary_edge_attributes'], pg['spanning_sides'])
    ingest_ms = (time.perf_counter() - t) * 1e3

    from pypercolate_amd import _graph
    ga = _graph.square_lattice(length)

    def call(**extra):
        return hpc.bond_canonical_averages_batch(ps=ps, num_runs=runs,
                                                 philox_seed=seed, **extra, **pg)

    def wall(**extra):
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = call(**extra)
        torch.cuda.synchronize()
        return out, (time.perf_counter() - t) * 1e3
    for _ in range(warmup):
        call()
    avg, ms = wall()
    avg_ga, ms_ga = wall(graph_arrays=ga)
    M = pg['num_edges']
    wt = E.binomial_weights(M, ps)
    g_lo = int(wt.host_start.min())
    g_len = int((wt.host_start + wt.host_length).max()) - g_lo
    shifted = Weights(M, ps, wt.start - g_lo, wt.length, wt.kmax, wt.w)
    shifted.host_start = wt.host_start - g_lo
    shifted.host_length = wt.host_length
    chunk = min(runs, 1024)
    x = torch.rand((g_len, 7 * chunk), dtype=torch.float64, device='cuda')
    plan = E.ContractPlan(x, shifted, 'fp64_mfma')
    plan.run()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    times = []
    for _ in range(reps):
        e0.record()
        plan.run()
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    cms = sorted(times)[len(times) // 2]
    flops = 2.0 * float(wt.host_length.sum()) * x.shape[1]
    ok = bool(avg['number_of_runs'][0] == runs and
              all(np.array_equal(avg[k], avg_ga[k]) for k in avg.dtype.names) and
              np.all((avg['percolation_probability_mean'] >= 0) &
                     (avg['percolation_probability_mean'] <= 1)))
    return {
        'workload': ('hpc.bond_canonical_averages_batch, L={} (N={}, M={}), {} '
                     'runs from Philox seed {}, {} p in [0.45, 0.55], 7 '
                     'columns per run').format(length, pg['num_nodes'], M,
                                               runs, seed, num_p),
        'ms': ms, 'graph_ingest_ms': ingest_ms,
        'ms_with_graph_arrays': ms_ga,
        'runs_per_s': runs / (ms / 1e3),
        'runs_per_s_with_graph_arrays': runs / (ms_ga / 1e3),
        'bond_additions_per_s': float(runs) * M / (ms / 1e3),
        'contraction_chunk': {
            'runs': chunk, 'columns': int(x.shape[1]), 'steps': g_len,
            'ms': cms, 'algorithmic_flops': flops,
            'tflops': flops / cms / 1e9, 'peak_tflops': 78.6,
            'frac': flops / cms / 1e9 / 78.6},
        'check': ok,
    }


def batch_bench(args):
    """Finite-size batch: L in {16, 32, 64, 128, 256, 512}, 2^26 / L^2 runs
    each (Philox seed 42), permute -> sweep -> reduce -> canonical averages
    (101 p in [0.4, 0.6]) for every L, with autotuned sweep knobs, one
    stream per L, recorded once as a HIP graph and replayed.  Beside it, the
    plain per-L GPU form: the same pipelines launched one L after the other
    on one stream, default knobs, no graph; its exact sums must be
    bit-identical to the batch's."""
    import torch
    from pypercolate_amd.batch import FiniteSizeBatch, default_runs
    lengths = [16, 32, 64, 128, 256, 512]
    runs = default_runs(lengths)
    ps = np.linspace(0.4, 0.6, 101)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)

    def timed(run):
        for _ in range(args.warmup):
            run()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(args.steps):
            run()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / args.steps

    plain = FiniteSizeBatch(lengths, runs, seed=42, ps=ps, tune=False)
    plain_ms = timed(plain.run_sequential)
    plain_res = plain.results()
    del plain
    torch.cuda.empty_cache()
    batch = FiniteSizeBatch(lengths, runs, seed=42, ps=ps)
    batch.capture()
    ms = timed(batch.run)
    res = batch.results()
    ok = all(int(r['sum1'][0, -1, 0]) == r['runs'] * L * L for L, r in res.items())
    exact = all(np.array_equal(res[L]['sum1'], plain_res[L]['sum1']) and
                np.array_equal(res[L]['sum2'], plain_res[L]['sum2'])
                for L in lengths)
    canon = max(float(np.abs(res[L]['canonical'] - plain_res[L]['canonical']).max())
                for L in lengths)
    del batch
    torch.cuda.empty_cache()
    # one persistent sweep launch over all L (device work queue), reductions
    # per L on the streams, recorded as a HIP graph
    pers = FiniteSizeBatch(lengths, runs, seed=42, ps=ps, persistent=True)
    pers.capture()
    pers_ms = timed(pers.run)
    pres = pers.results()
    pers_exact = all(np.array_equal(pres[L]['sum1'], plain_res[L]['sum1']) and
                     np.array_equal(pres[L]['sum2'], plain_res[L]['sum2']) and
                     np.array_equal(pres[L]['canonical'], plain_res[L]['canonical'])
                     for L in lengths)
    additions = pers.additions
    del pers
    torch.cuda.empty_cache()
    return {
        'metric': METRIC + ' [finite-size batch]',
        'value': additions / (ms / 1e3),
        'unit': 'bond additions/s',
        'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': ms, 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'u16/u32/u64/u128 + f64',
        'data': 'synthetic: Philox-keyed device permutations, seed 42',
        'config': {'workload': ('finite-size batch, L in {} with {} runs, '
                                'one HIP graph per step').format(lengths, runs),
                   'lengths': lengths, 'runs': runs,
                   'parallelism': 'single GPU, one stream per L'},
        'persistent': {'ms_per_step': pers_ms,
                       'value': additions / (pers_ms / 1e3),
                       'form': 'permutations of every L, then one persistent '
                               'sweep launch whose workgroups take the runs of '
                               'all L (largest first) from a device queue, '
                               'then per-L reductions on three streams; one '
                               'HIP graph',
                       'speedup_vs_streams': ms / pers_ms,
                       'sums_and_canonical_bit_exact_vs_plain': pers_exact},
        'plain_per_L': {'ms_per_step': plain_ms,
                        'value': additions / (plain_ms / 1e3),
                        'form': 'per-L launches in sequence on one stream, '
                                'default sweep knobs, no HIP graph',
                        'speedup_of_batch': plain_ms / ms,
                        'sums_bit_exact_vs_batch': exact,
                        'canonical_max_abs_diff': canon},
        'tuning': {L: r['tuning'] for L, r in res.items()},
        'sums_check': ok,
    }


def cpu_baseline(length, sample_runs, seed):
    """The oracle on a bounded sample of the workload: the C restatement of
    the sweep with the exact per-step sums over runs (nzo_bond_sums, no
    traces kept), OpenMP over runs, on host RandomState permutations (drawn
    once per batch of 256 runs, not timed)."""
    from oracle import oracle

    g = oracle.square_lattice_graph(length)
    M = g['num_edges']
    threads = min(int(os.environ.get('OMP_NUM_THREADS', 0)) or os.cpu_count() or 1, 16)
    batch = 256
    perms = oracle.randomstate_permutations(range(seed, seed + batch), M)
    t = 0.0
    done = 0
    sums = None
    while done < sample_runs:
        t0 = time.perf_counter()
        sums = oracle.bond_sums(g, perms, sums=sums, threads=threads)
        t += time.perf_counter() - t0
        done += batch
    return {
        'value': done * M / t,
        'unit': 'bond additions/s',
        'cores': threads,
        'kind': 'port',
        'sample': ('{} runs of L={} (host RandomState permutations, drawn '
                   'untimed), C oracle sweep with exact per-step sums '
                   '(128/192-bit), OpenMP over runs; {:.2f} s').format(done, length, t),
    }


def _pysweep_runs(job):
    from oracle import pysweep
    length, seeds = job
    for s in seeds:
        pysweep.lattice_run(length, s)
    return len(seeds)


def reference_class_baseline(length, single_runs, runs_per_proc):
    """The reference-class CPU path: a pure-Python restatement of the
    per-edge loop (oracle/pysweep.py: dict union-find, numpy uint64 moment
    vectors, per-step arrays), timed single-process on one core and with one
    process per core (the reference's jug-style parallelism).  Runs before
    the GPU is touched, so the worker processes are plain forks."""
    import multiprocessing as mp
    M = 2 * length * (length - 1)
    t0 = time.perf_counter()
    _pysweep_runs((length, list(range(single_runs))))
    t1 = time.perf_counter() - t0
    procs = min(os.cpu_count() or 1, 16)
    jobs = [(length, list(range(100 + i * runs_per_proc,
                                100 + (i + 1) * runs_per_proc)))
            for i in range(procs)]
    t0 = time.perf_counter()
    with mp.get_context('fork').Pool(procs) as pool:
        done = sum(pool.map(_pysweep_runs, jobs))
    tp = time.perf_counter() - t0
    return {
        'single_core': {'value': single_runs * M / t1,
                        'unit': 'bond additions/s', 'cores': 1,
                        'sample': '{} runs of L={}; {:.2f} s'.format(
                            single_runs, length, t1)},
        'all_cores': {'value': done * M / tp, 'unit': 'bond additions/s',
                      'cores': procs,
                      'sample': '{} processes x {} runs of L={}; {:.2f} s '
                                'including process start'.format(
                                    procs, runs_per_proc, length, tp)},
        'kind': 'port (pure Python, reference cost class)',
        'reference_published': {'value': 2.14e4, 'unit': 'bond additions/s',
                                'cores': 1,
                                'host': "i7-2620M, the reference's published figure"},
    }


def bitexact_mode(eng, R, steps, start, stop):
    """The L = 128 pipeline with numpy RandomState(r).permutation(M) for run r
    drawn on the device (MT19937 + the legacy shuffle), each step drawing
    the next step's permutations beside its last two sweeps and reductions
    (the headline loop draws them beside the last reductions)."""
    import torch
    from pypercolate_amd._engine import Pipeline
    exact = Pipeline(eng, R, seeds=np.arange(R), prefetch=True, prefetch_after=2)
    exact.run()
    torch.cuda.synchronize()
    start.record()
    for _ in range(steps):
        exact.run()
    stop.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / steps
    s1 = exact.sum1.cpu().numpy().view('uint64')
    res = {'value': float(R) * eng.M / (ms / 1e3), 'unit': 'bond additions/s',
           'ms_per_step': ms,
           'permutations': 'numpy RandomState(r).permutation(M), r = 0..{}, '
                           'on the device (MT19937 + legacy shuffle), the next '
                           'step\'s drawn beside the last two sweeps'.format(R - 1),
           'sums_check': bool(int(s1[0, eng.M, 0]) == R * eng.N and
                              int(s1[1, eng.M, 0]) == R)}
    del exact
    torch.cuda.empty_cache()
    return res


def main(argv=None):
    args = parse_args(argv)
    if args.gpus != 1:
        print('bench.py runs on one GPU only: this build has no multi-GPU '
              'path (the runs would shard across GPUs, with one all-reduce of '
              'the exact sums)', file=sys.stderr)
        return 2
    ref_class = None
    if not args.batch and args.cpu_sample_runs > 0:
        ref_class = reference_class_baseline(args.length, 8, 4)

    import torch
    from pypercolate_amd import _graph, _native as nat
    from pypercolate_amd._engine import BondEngine

    torch.cuda.set_device(0)
    if args.batch:
        out = batch_bench(args)
        print(json.dumps(out), flush=True)
        return 0 if out['sums_check'] else 1
    from pypercolate_amd._engine import Pipeline
    from pypercolate_amd.autotune import tune_sweep
    eng = BondEngine(_graph.square_lattice(args.length))
    tuning = tune_sweep(eng) if args.tune else None
    R, M, N, S = args.runs, eng.M, eng.N, eng.S
    dev = eng.device
    # each step also draws the next step's permutations once its sweeps have
    # ended, beside its last reductions (the first step draws its own)
    pipe = Pipeline(eng, R, seed=args.seed, prefetch=True, prefetch_after=0)
    sum1 = pipe.sum1

    for _ in range(args.warmup):
        pipe.run()
    torch.cuda.synchronize()

    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_wall = time.perf_counter()
    start.record()
    for _ in range(args.steps):
        pipe.run()
    stop.record()
    torch.cuda.synchronize()
    t_wall = time.perf_counter() - t_wall
    total_ms = start.elapsed_time(stop)
    ms_per_step = total_ms / args.steps
    # HIP-event time of the sweep launches of the last step, each on the
    # stream it runs on (two sweep streams: consecutive launches overlap, so
    # the average launch duration also counts the time a launch shares the
    # GPU with its neighbour; the span is first start to last end)
    sweep_step_ms = pipe.sweep_ms()
    sweep_span_ms = pipe.sweep_span_ms()
    launches = len(pipe.sweep_events)
    sweep_avg_ms = sweep_step_ms / max(launches, 1)

    # sanity on the last step's exact sums (not timed): at n = M every run
    # is one cluster of N nodes and spans
    s1 = sum1.cpu().numpy().view('uint64')
    ok = int(s1[0, M, 0]) == R * N and int(s1[1, M, 0]) == R

    if args.copy_calibration:
        a = torch.empty(1 << 28, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()

    adds = float(R) * M
    traffic_per_add = profiled_traffic_per_add(args.length, R)
    value = adds / (ms_per_step / 1e3)
    alg_bytes = ALG_BYTES_PER_ADD_LDS if N <= 16384 else ALG_BYTES_PER_ADD_HBM
    # algorithmic bytes of the step's sweep launches over their span: the
    # launches overlap on two streams, so one launch's own duration also
    # counts the time it shares the GPU with its neighbour
    achieved = alg_bytes * adds / (sweep_span_ms / 1e3) / 1e9
    out = {
        'metric': METRIC,
        'value': value,
        'unit': 'bond additions/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': ms_per_step,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'u16/u32/u64/u128',
        'data': ('synthetic: uniform device permutations (Philox-keyed '
                 'bucket shuffle), seed {}').format(args.seed),
        'config': {
            'workload': ('bond percolation, L={} square lattice '
                         '(N={}, M={}), {} runs per step: permute -> sweep '
                         '-> chunk states -> exact microcanonical sums').format(
                             args.length, N, M, R),
            'length': args.length, 'runs': R, 'nodes': N, 'edges': M,
            'parallelism': ('single GPU, one run per workgroup of {} '
                            'wavefront(s); permutations of all runs (one '
                            'launch; every step draws the next step\'s once '
                            'its sweeps have ended, beside its last '
                            'reductions), sub-batches of {} runs (one '
                            'residency), sweeps alternating between two '
                            'streams, reductions beside them on a third').format(
                                2 if N >= 4096 else 1, pipe.sub),
        },
        'roofline': {
            'kernel': 'sweep_lds_kernel' if N <= 16384 else 'sweep_global_kernel',
            'bound': 'hbm',
            # what actually limits it: the union-find state is in LDS, and
            # the resolver's dependent instruction chain (one resolver wave
            # per SIMD at 4 runs per CU) sets the pace, not a bandwidth
            'limiter': ('resolver dependent-instruction chain (LDS-resident '
                        'state, issue/latency bound)' if N <= 16384 else
                        'HBM lines of the union-find state'),
            'achieved': achieved,
            'peak': HBM_PEAK_GBS,
            'unit': 'GB/s',
            'frac': achieved / HBM_PEAK_GBS,
            'traffic': (None if traffic_per_add is None else
                        traffic_per_add * adds / max(launches, 1)),
            'traffic_per': 'launch of {} runs'.format(pipe.sub),
            'alg_bytes_per_add': alg_bytes,
            'achieved_from': ('algorithmic bytes of all sweep launches of a '
                              'step over their span (first start to last end, '
                              'HIP events on the sweep streams)'),
            'span_ms': sweep_span_ms,
            'launches_per_step': launches,
            'kernel_ms': sweep_avg_ms,
            'sweep_ms_per_step': sweep_step_ms,
            'achieved_per_launch_duration': (alg_bytes * adds / max(launches, 1)
                                             / (sweep_avg_ms / 1e3) / 1e9),
        },
        'tuning': tuning,
        'sums_check': ok,
        'wall_s': t_wall,
    }
    del pipe
    torch.cuda.empty_cache()
    # bit-exact mode: numpy RandomState(r).permutation(M) for run r, drawn
    # on the device (MT19937 + the legacy shuffle), same pipeline otherwise:
    # the reference's own permutations, so its traces are bit-identical
    if not args.no_bitexact:
        out['bitexact_mode'] = bitexact_mode(eng, R, args.steps, start, stop)
    if args.per_run:
        out['per_run_canonical'] = per_run_canonical(args.length, R, 1000,
                                                     args.seed)
    if args.canonical_length > 0:
        info, big, sums, runs = large_lattice(args.canonical_length,
                                              args.canonical_runs, seed=7)
        out['config_L{}'.format(args.canonical_length)] = info
        out['canonical'] = canonical_comparison(big, sums, runs,
                                                args.canonical_ps)
        del sums
    if args.cpu_sample_runs > 0:
        out['cpu_baseline'] = cpu_baseline(args.length, args.cpu_sample_runs,
                                           args.seed)
        out['cpu_baseline']['reference_class'] = ref_class
    print(json.dumps(out), flush=True)
    return 0 if ok else 1


if __name__ == '__main__':
    sys.exit(main())
