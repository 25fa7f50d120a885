"""Measurement-driven choice of the sweep's launch knobs.

For a graph, ``tune_sweep`` times the sweep on the device over a sample of
runs for every candidate
  * number of parallel conflict-free merge rounds per batch of 64 bonds,
  * union-find state placement (LDS for N <= 16384, or the HBM workspace),
  * one or two wavefronts per run (a finder beside the resolver),
and keeps the fastest.  The sample is one residency's worth of runs (the
HIP occupancy calculator's workgroups per CU times the number of CUs, at
most about 2^28 bonds of buffers), so the timing sees the concurrency of a
full launch, or the caller's own number of runs when that is smaller
(``max_runs``: a launch of a few runs per CU is bound by latency rather
than by the memory system, and may prefer other knobs).  None of the knobs
changes the results; ``tests/test_gpu_batch.py`` checks that the tuned
sweep is bit-identical to the default one.

Choices are cached per device and graph in a JSON file
($PYPERCOLATE_AMD_CACHE, default ~/.cache/pypercolate_amd/autotune.json).
"""

import hashlib
import json
import os

import numpy as np

from . import _native as nat

ROUND_CANDIDATES = (1, 2, 3, 4, 6, 8)


def cache_path():
    base = os.environ.get('PYPERCOLATE_AMD_CACHE') or os.path.join(
        os.path.expanduser('~'), '.cache', 'pypercolate_amd')
    return os.path.join(base, 'autotune.json')


def graph_key(graph):
    h = hashlib.sha1(graph.edges.tobytes())
    if graph.node_sides is not None:
        h.update(graph.node_sides.tobytes())
    return 'N{}-M{}-{}'.format(graph.num_nodes, graph.num_edges, h.hexdigest()[:16])


def _load_cache():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(cache):
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + '.tmp{}'.format(os.getpid())
        with open(tmp, 'w') as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def _time_sweep(engine, perm, records, span, reps=3):
    torch = engine.torch
    ws = None
    wb = engine.workspace_bytes(int(perm.shape[0]))
    if wb:
        ws = torch.empty((wb,), dtype=torch.uint8, device=engine.device)
    engine.sweep(perm, records=records, span_step=span, workspace=ws)
    times = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        engine.sweep(perm, records=records, span_step=span, workspace=ws)
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    return sorted(times)[len(times) // 2]


def tune_sweep(engine, sample_runs=None, rounds=ROUND_CANDIDATES, modes=None,
               use_cache=True, seed=0, max_runs=None):
    """Pick (merge_rounds, state_mode) for ``engine`` (a BondEngine), set
    them on it and return a dict with the choice and the timings.
    state_mode candidates are PC_STATE_LDS / PC_STATE_HBM, each with one
    and with two wavefronts per run (PC_SWEEP_*).  ``max_runs``: the most
    runs the caller sweeps in one launch; the sample is cut to it, and the
    cache keeps one choice per graph and cap."""
    torch = engine.torch
    device = torch.cuda.get_device_properties(engine.device)
    key = '{}|{}|{}'.format(device.gcnArchName, device.multi_processor_count,
                            graph_key(engine.graph))
    if max_runs:
        key += '|runs<={}'.format(int(max_runs))
    cache = _load_cache() if use_cache else {}
    if key in cache:
        choice = cache[key]
        engine.merge_rounds = int(choice['merge_rounds'])
        engine.state_mode = int(choice['state_mode'])
        return choice
    if modes is None:
        places = [nat.PC_STATE_HBM]
        if engine.N <= 16384:
            places = [nat.PC_STATE_LDS, nat.PC_STATE_HBM]
        modes = [p | w for p in places
                 for w in (nat.PC_SWEEP_ONE_WAVE, nat.PC_SWEEP_TWO_WAVES)]
    cus = device.multi_processor_count
    timings = {}
    best = None
    for mode in modes:
        occ = max(engine.occupancy(mode), 1)
        # one residency of runs, capped at about 2^28 bonds of buffers
        cap = max(64, (1 << 28) // max(engine.M, 1))
        runs = sample_runs or max(1, min(occ * cus, 8192, cap))
        if max_runs:
            runs = min(runs, int(max_runs))
        perm = engine.device_permutations(seed, runs)
        records = torch.empty((runs, max(engine.M, 1)), dtype=engine.rec_torch,
                              device=engine.device)
        span = torch.empty((runs,), dtype=torch.int32, device=engine.device)
        engine.state_mode = mode
        for r in rounds:
            engine.merge_rounds = r
            ms = _time_sweep(engine, perm, records, span)
            rate = runs * engine.M / ms * 1e3
            timings['{}/{}'.format(mode, r)] = rate
            if best is None or rate > best[0]:
                best = (rate, r, mode)
        del perm, records, span
    choice = {'merge_rounds': best[1], 'state_mode': best[2],
              'adds_per_s': best[0], 'candidates': timings}
    engine.merge_rounds, engine.state_mode = best[1], best[2]
    if use_cache:
        cache[key] = choice
        _store_cache(cache)
    return choice
