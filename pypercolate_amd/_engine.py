"""Device pipeline of the Newman-Ziff hot path on one GPU.

``BondEngine`` holds one graph in HBM and runs batches of runs through
libpercolate_hip: permutations (uploaded from numpy RandomState for bit
parity with the reference, or drawn on the device), the union-find sweep,
the per-run chunk states, and then either the exact microcanonical sums over
runs or per-run traces.  All buffers are torch tensors on the current
device; kernels run on torch's current stream.
"""

import ctypes

import numpy as np

from . import _graph
from . import _native as nat


class SweepResult(object):
    """Device outputs of one batch of runs."""

    def __init__(self, engine, num_runs, records, span_step, perm=None):
        self.engine = engine
        self.num_runs = num_runs
        self.records = records        # (R, M) int32/int64 tensor (u32/u64 bits)
        self.span_step = span_step    # (R,) int32 tensor (u32 bits)
        self.perm = perm              # (R, M) int32 tensor or None
        self._chunk_state = None

    @property
    def chunk_state(self):
        if self._chunk_state is None:
            self._chunk_state = self.engine.chunk_scan(self)
        return self._chunk_state

    @chunk_state.setter
    def chunk_state(self, value):
        self._chunk_state = value


class BondEngine(object):
    """One percolation graph resident on the GPU."""

    def __init__(self, graph, relabel=None):
        """relabel: store the union-find state of 2-D grid nodes in tiles
        (_graph.locality_order) -- by default for graphs whose state lives
        in HBM (N > 16384), where every random access moves a whole line.
        The results do not depend on it."""
        self.torch = torch = nat.torch_cuda()
        self.lib = nat.load_library()
        self.graph = graph
        self.N = graph.num_nodes
        self.M = graph.num_edges
        self.S = self.M + 1
        dev = torch.device('cuda', torch.cuda.current_device())
        self.device = dev
        if relabel is None:
            relabel = self.N > 16384
        order = _graph.locality_order(graph) if relabel else None
        self.relabelled = order is not None
        edges = graph.edges if order is None else order[graph.edges]
        sides = graph.node_sides
        if order is not None and sides is not None:
            sides = np.empty_like(graph.node_sides)
            sides[order] = graph.node_sides
        self.d_edges = torch.from_numpy(
            np.ascontiguousarray(edges, dtype=np.int32).reshape(-1)).to(dev)
        self.d_sides = (
            None if sides is None else torch.from_numpy(sides.copy()).to(dev)
        )
        self.span_first = 1 if graph.span_on_first_merge else 0
        # the closed-form square lattice: its edges follow from their index
        # in the sweep (pc_sweep_lattice), no edge array is read; with the
        # tiled relabelling only where the tiles are whole (L a multiple of 8)
        L = int(getattr(graph, 'lattice_length', 0) or 0)
        if L and order is not None and L % 8:
            L = 0
        self.lattice_length = L
        self.lattice_tiled = order is not None
        # used where the state is in HBM: there the random edge reads are
        # whole lines (L = 1024: 190 -> 168 ms per sweep of 1000 runs); with
        # LDS state the edge array sits in L2 and computing measured 1.4 %
        # slower (L = 128)
        self.lattice_sweep = bool(L) and self.N > 16384
        self.rec_bytes = self.lib.pc_record_bytes(self.N)
        self.rec_torch = torch.int32 if self.rec_bytes == 4 else torch.int64
        self.rec_numpy = nat.record_dtype(self.N)
        # sweep knobs (results do not depend on them); set by the autotuner:
        # merge rounds per batch and state placement (PC_STATE_*) OR-ed with
        # the wavefronts per run (PC_SWEEP_*)
        self.merge_rounds = 0
        self.state_mode = nat.PC_STATE_AUTO

    # -- permutations -----------------------------------------------------
    def upload_permutations(self, perms):
        """(R, M) host permutations -> device int32 tensor."""
        perms = np.asarray(perms)
        if perms.size and (perms.min() < 0 or perms.max() >= self.M):
            raise ValueError('permutation entries must be edge indices < M')
        perms = np.ascontiguousarray(perms, dtype=np.int32).reshape(-1, self.M)
        return self.torch.from_numpy(perms).to(self.device)

    def permutation_workspace(self, num_runs):
        """Device workspace of device_permutations for num_runs runs."""
        b = ctypes.c_uint64(0)
        nat.check(self.lib.pc_permute_philox_workspace_size(
            self.M, int(num_runs), ctypes.byref(b)),
            'pc_permute_philox_workspace_size')
        return (self.torch.empty((int(b.value),), dtype=self.torch.uint8,
                                 device=self.device) if b.value else None)

    def device_permutations(self, seed, num_runs, first_run=0, out=None,
                            workspace=None):
        """Philox-keyed uniform device permutations (pc_permute_philox) for
        runs first_run..first_run + num_runs - 1."""
        torch = self.torch
        if out is None:
            out = torch.empty((num_runs, self.M), dtype=torch.int32,
                              device=self.device)
        if workspace is None:
            workspace = self.permutation_workspace(num_runs)
        st = nat.stream_handle(torch)
        nat.check(self.lib.pc_permute_philox(
            int(seed) & 0xFFFFFFFFFFFFFFFF, int(first_run), int(num_runs),
            self.M, nat.ptr(out), self.M, nat.ptr(workspace),
            0 if workspace is None else int(workspace.numel()), st),
            'pc_permute_philox')
        return out

    def upload_seeds(self, seeds):
        """RandomState seeds (integers in [0, 2**32)) as a device uint32
        vector (an int32 tensor holding the bits)."""
        seeds = np.asarray(seeds, dtype=np.int64).reshape(-1)
        if seeds.size and (seeds.min() < 0 or seeds.max() > 0xFFFFFFFF):
            raise ValueError('RandomState seeds must be in [0, 2**32 - 1]')
        return self.torch.from_numpy(seeds.astype(np.uint32).view(np.int32)).to(
            self.device)

    def randomstate_permutations(self, seeds, out=None):
        """numpy RandomState(seed).permutation(M) for every seed, drawn on
        the device (bit-identical to numpy's legacy generator).  seeds: host
        integers, or a device vector from upload_seeds."""
        torch = self.torch
        d_seeds = seeds if torch.is_tensor(seeds) else self.upload_seeds(seeds)
        n = int(d_seeds.numel())
        if out is None:
            out = torch.empty((n, max(self.M, 1)), dtype=torch.int32,
                              device=self.device)
        nat.check(self.lib.pc_permute_randomstate(
            nat.ptr(d_seeds), n, self.M, nat.ptr(out), max(self.M, 1),
            nat.stream_handle(torch)), 'pc_permute_randomstate')
        return out

    # -- sweep ------------------------------------------------------------
    def workspace_bytes(self, num_runs, state_mode=None):
        b = ctypes.c_uint64(0)
        mode = self.state_mode if state_mode is None else state_mode
        nat.check(self.lib.pc_sweep_workspace_size(self.N, num_runs, mode,
                                                   ctypes.byref(b)),
                  'pc_sweep_workspace_size')
        return int(b.value)

    def occupancy(self, state_mode=None):
        """Resident sweep workgroups per CU (HIP occupancy calculator)."""
        v = ctypes.c_int32(0)
        mode = self.state_mode if state_mode is None else state_mode
        nat.check(self.lib.pc_sweep_occupancy(self.N, mode, ctypes.byref(v)),
                  'pc_sweep_occupancy')
        return int(v.value)

    def fuse_chunk_states(self):
        """Whether the sweep should write the chunk states itself (its finder
        wave folds the records one batch behind the resolver).  Measured on
        MI355X: with the union-find state in LDS the fold still lengthens the
        sweep by ~12 % (L = 128: 7.7 -> 8.7 ms for 10^4 runs), more than the
        separate pass (pc_chunk_scan) costs on another stream; with the state
        in HBM the sweep hides it and the separate pass (one wavefront
        walking each long run) is the slower choice."""
        placement = self.state_mode & ~nat.PC_SWEEP_WAVES_MASK
        lds = self.N <= 16384 and placement != nat.PC_STATE_HBM
        return not lds

    def sweep(self, d_perm, records=None, span_step=None, workspace=None,
              chunk_state=None, chunks=None):
        """Sweep the runs of d_perm (R, M).  With chunks (default: whether
        fuse_chunk_states()) the sweep also writes the chunk states (into
        chunk_state, or a new tensor), so the result needs no chunk_scan
        pass; otherwise they are computed on first use."""
        torch = self.torch
        R = int(d_perm.shape[0])
        if chunks is None:
            chunks = chunk_state is not None or self.fuse_chunk_states()
        if records is None:
            records = torch.empty((R, max(self.M, 1)), dtype=self.rec_torch,
                                  device=self.device)
        if span_step is None:
            span_step = torch.empty((R,), dtype=torch.int32, device=self.device)
        wb = self.workspace_bytes(R)
        if wb and workspace is None:
            workspace = torch.empty((wb,), dtype=torch.uint8, device=self.device)
        elif workspace is not None:
            wb = int(workspace.numel())
        if chunks and chunk_state is None:
            C = max((self.M + nat.PC_CHUNK - 1) // nat.PC_CHUNK, 1)
            chunk_state = torch.empty((R, C, nat.PC_STATE_WORDS),
                                      dtype=torch.int64, device=self.device)
        st = nat.stream_handle(torch)
        if self.lattice_sweep:
            nat.check(self.lib.pc_sweep_lattice(
                self.lattice_length, 1 if self.lattice_tiled else 0,
                nat.ptr(self.d_sides), self.span_first, nat.ptr(d_perm),
                max(self.M, 1), R, nat.ptr(records), max(self.M, 1),
                nat.ptr(span_step), nat.ptr(workspace), wb, self.merge_rounds,
                self.state_mode, nat.ptr(chunk_state) if chunks else None, st),
                'pc_sweep_lattice')
        else:
            nat.check(self.lib.pc_sweep(
                self.N, self.M, nat.ptr(self.d_edges), nat.ptr(self.d_sides),
                self.span_first, nat.ptr(d_perm), max(self.M, 1), R,
                nat.ptr(records), max(self.M, 1), nat.ptr(span_step),
                nat.ptr(workspace), wb, self.merge_rounds, self.state_mode,
                nat.ptr(chunk_state) if chunks else None, st),
                'pc_sweep')
        res = SweepResult(self, R, records, span_step, d_perm)
        if chunks:
            res.chunk_state = chunk_state
        return res

    def chunk_scan(self, res, out=None):
        torch = self.torch
        C = max((self.M + nat.PC_CHUNK - 1) // nat.PC_CHUNK, 1)
        if out is None:
            out = torch.empty((res.num_runs, C, nat.PC_STATE_WORDS),
                              dtype=torch.int64, device=self.device)
        st = nat.stream_handle(torch)
        nat.check(self.lib.pc_chunk_scan(
            self.N, self.M, res.num_runs, nat.ptr(res.records),
            max(self.M, 1), nat.ptr(out), st), 'pc_chunk_scan')
        return out

    # -- outputs ------------------------------------------------------------
    def microcanonical_sums(self, res, sum1=None, sum2=None, accumulate=False):
        """Exact per-step sums over runs: device tensors sum1 (7, S, 2) and
        sum2 (7, S, 3) of uint64 limbs (stored as int64)."""
        torch = self.torch
        if sum1 is None:
            sum1 = torch.empty((nat.PC_COLUMNS, self.S, 2), dtype=torch.int64,
                               device=self.device)
        if sum2 is None:
            sum2 = torch.empty((nat.PC_COLUMNS, self.S, 3), dtype=torch.int64,
                               device=self.device)
        st = nat.stream_handle(torch)
        nat.check(self.lib.pc_microcanonical_sums(
            self.N, self.M, res.num_runs, nat.ptr(res.records),
            max(self.M, 1), nat.ptr(res.span_step), nat.ptr(res.chunk_state),
            nat.ptr(sum1), nat.ptr(sum2), 1 if accumulate else 0, st),
            'pc_microcanonical_sums')
        return sum1, sum2

    def traces(self, res):
        """Per-run traces on the device: max (R, S) int32 (u32 bits),
        spanning (R, S) uint8, moments (R, S, 5) int64 (u64 bits)."""
        torch = self.torch
        R = res.num_runs
        mx = torch.empty((R, self.S), dtype=torch.int32, device=self.device)
        sp = torch.empty((R, self.S), dtype=torch.uint8, device=self.device)
        mom = torch.empty((R, self.S, 5), dtype=torch.int64, device=self.device)
        st = nat.stream_handle(torch)
        nat.check(self.lib.pc_expand_traces(
            self.N, self.M, R, nat.ptr(res.records), max(self.M, 1),
            nat.ptr(res.span_step), nat.ptr(res.chunk_state), nat.ptr(mx),
            nat.ptr(sp), nat.ptr(mom), st), 'pc_expand_traces')
        return mx, sp, mom

    def host_traces(self, res):
        """Per-run traces copied to numpy: max u32, spanning bool, moments u64."""
        mx, sp, mom = self.traces(res)
        return (mx.cpu().numpy().view(np.uint32), sp.cpu().numpy().astype(bool),
                mom.cpu().numpy().view(np.uint64))


def limbs_to_int(limbs):
    """(..., L) uint64 limb array (little-endian) -> object array of ints."""
    limbs = np.asarray(limbs).view(np.uint64)
    out = np.zeros(limbs.shape[:-1], dtype=object)
    for j in range(limbs.shape[-1] - 1, -1, -1):
        out = out * (1 << 64) + limbs[..., j].astype(object)
    return out


def host_sums(sum1, sum2):
    """Device sum tensors -> exact Python-int object arrays (7, S)."""
    return (limbs_to_int(sum1.cpu().numpy()),
            limbs_to_int(sum2.cpu().numpy()))


_PIPELINE_STREAMS = {}


def _pipeline_streams(torch, device):
    """The three streams every Pipeline on a device shares, created once
    per process: HIP deals streams to the hardware queues in creation order,
    so three fresh streams per pipeline would sooner or later share a queue
    with each other (or the caller's) and serialise."""
    key = torch.device(device).index
    if key not in _PIPELINE_STREAMS:
        with torch.cuda.device(device):
            _PIPELINE_STREAMS[key] = [torch.cuda.Stream() for _ in range(3)]
    return _PIPELINE_STREAMS[key]


class Pipeline(object):
    """R runs of one graph through permute -> sweep -> chunk states -> exact
    sums, in sub-batches on four streams so that the reduction of one
    sub-batch runs beside the sweep of the next (the sweep leaves most issue
    slots of a CU idle: its occupancy is bounded by LDS, four runs per CU at
    N = 16384).  Stream order: permutations on the caller's stream, sweeps
    alternating between two more (each waiting for its permutations; the
    next sub-batch's runs fill the CUs that the last runs of a launch leave
    idle), chunk states and sums on a fourth (each waiting for its sweep,
    accumulating into one pair of sum buffers).  Four streams in all, one
    per hardware queue of the process (HIP's default of four): streams that
    shared a queue would serialise.  HBM-state graphs get their chunk
    states from the sweep itself (BondEngine.fuse_chunk_states).  All
    buffers are allocated up front.

    Permutations: Philox-keyed (pc_permute_philox, seed) by default, or
    with seeds= (R integers) numpy RandomState(seeds[r]).permutation(M),
    bit-identical to the reference's draw, on the device.

    prefetch (permutations drawn ahead): each run() also draws the
    permutations of the next run() into a second buffer, on the caller's
    stream once the sweep of sub-batch ``-1 - prefetch_after`` has ended,
    so that they fill the CUs the last sweeps and the reductions leave free;
    the next run() starts with its sweeps (RandomState permutations too:
    10.4 ms per step of 10^4 runs at L = 128 with prefetch_after = 2,
    against 11.1 without; started earlier, they stretch the sweeps by about
    as much as they hide).  The first run() draws its own.  Every run()
    still draws one batch of permutations, the same ones (the permutations
    depend only on the seeds, or the seed and the run index)."""

    def __init__(self, engine, num_runs, seed=0, sub_batch=None, seeds=None,
                 permute_ahead=None, prefetch=False, prefetch_after=1):
        torch = self.torch = engine.torch
        self.engine = eng = engine
        self.R = int(num_runs)
        if self.R < 1:
            raise ValueError('a pipeline needs at least one run')
        self.seed = seed
        self.d_seeds = None
        if seeds is not None:
            self.d_seeds = eng.upload_seeds(seeds)
            if int(self.d_seeds.numel()) != self.R:
                raise ValueError('need one seed per run')
        dev = eng.device
        M, S = eng.M, eng.S
        if sub_batch is None:
            # one residency (runs the GPU holds at once): with the sweeps on
            # two streams, 9.3 ms per step of 10^4 runs at L = 128 against
            # 9.4 with 1.5 residencies, 9.5 with two and 10.6 with three
            # (an earlier, slower sweep; smaller first or last sub-batches
            # measured slower with the current one too)
            cus = torch.cuda.get_device_properties(dev).multi_processor_count
            sub_batch = max(eng.occupancy(), 1) * cus
        self.sub = max(1, min(int(sub_batch), self.R))
        # permutations of every run in one launch before the sweeps (both
        # kinds need most of a CU's LDS per run, so beside the sweeps they
        # wait for LDS the sweeps hold: 11.9 ms per step of 10^4 RandomState
        # runs at L = 128 against 18.0 per sub-batch), or per sub-batch
        self.permute_ahead = True if permute_ahead is None else bool(permute_ahead)
        C = max((M + nat.PC_CHUNK - 1) // nat.PC_CHUNK, 1)
        self.prefetch = bool(prefetch) and self.permute_ahead
        self.prefetch_after = max(0, int(prefetch_after))
        self.perms = [torch.empty((self.R, max(M, 1)), dtype=torch.int32, device=dev)
                      for _ in range(2 if self.prefetch else 1)]
        self.cur = 0
        self.prefetched = False
        self.perm = self.perms[0]
        self.records = torch.empty((self.R, max(M, 1)), dtype=eng.rec_torch,
                                   device=dev)
        self.span = torch.empty((self.R,), dtype=torch.int32, device=dev)
        self.state = torch.empty((self.R, C, nat.PC_STATE_WORDS),
                                 dtype=torch.int64, device=dev)
        self.sum1 = torch.empty((nat.PC_COLUMNS, S, 2), dtype=torch.int64,
                                device=dev)
        self.sum2 = torch.empty((nat.PC_COLUMNS, S, 3), dtype=torch.int64,
                                device=dev)
        wb = eng.workspace_bytes(self.sub)
        # one union-find workspace (HBM state) per sweep stream in use
        nsub = (self.R + self.sub - 1) // self.sub
        self.workspaces = [torch.empty((wb,), dtype=torch.uint8, device=dev)
                           if wb else None for _ in range(min(2, nsub))]
        self.streams = _pipeline_streams(torch, dev)
        self.sweep_events = []
        self.perm_ws = (eng.permutation_workspace(self.R if self.permute_ahead
                                                  else self.sub)
                        if seeds is None else None)

    def run(self):
        """Enqueue the whole batch; the caller's stream waits for it."""
        torch = self.torch
        eng = self.engine
        main = torch.cuda.current_stream()
        ss0, ss1, sr = self.streams
        sp = main
        start = torch.cuda.Event()
        start.record(main)
        for s in self.streams:
            s.wait_event(start)
        self.sweep_events = []
        first = True
        fused = eng.fuse_chunk_states()
        self.perm = self.perms[self.cur]
        if self.permute_ahead:
            with torch.cuda.stream(sp):
                if self.prefetched:
                    pass        # drawn by the previous run(), earlier on sp
                elif self.d_seeds is None:
                    eng.device_permutations(self.seed, self.R, out=self.perm,
                                            workspace=self.perm_ws)
                else:
                    eng.randomstate_permutations(self.d_seeds, out=self.perm)
                have_all = torch.cuda.Event()
                have_all.record(sp)
        nsub = (self.R + self.sub - 1) // self.sub
        for i, b0 in enumerate(range(0, self.R, self.sub)):
            n = min(self.sub, self.R - b0)
            ss = (ss0, ss1)[i & 1]
            perm = self.perm[b0:b0 + n]
            if self.permute_ahead:
                have_perm = have_all
            else:
                with torch.cuda.stream(sp):
                    if self.d_seeds is None:
                        eng.device_permutations(self.seed, n, first_run=b0, out=perm,
                                                workspace=self.perm_ws)
                    else:
                        eng.randomstate_permutations(self.d_seeds[b0:b0 + n], out=perm)
                    have_perm = torch.cuda.Event()
                    have_perm.record(sp)
            ss.wait_event(have_perm)
            with torch.cuda.stream(ss):
                e0 = torch.cuda.Event(enable_timing=True)
                e0.record(ss)
                res = eng.sweep(perm, records=self.records[b0:b0 + n],
                                span_step=self.span[b0:b0 + n],
                                workspace=self.workspaces[i & 1],
                                chunk_state=self.state[b0:b0 + n] if fused else None,
                                chunks=fused)
                e1 = torch.cuda.Event(enable_timing=True)
                e1.record(ss)
            self.sweep_events.append((e0, e1))
            sr.wait_event(e1)
            with torch.cuda.stream(sr):
                if not fused:
                    res.chunk_state = eng.chunk_scan(res, out=self.state[b0:b0 + n])
                eng.microcanonical_sums(res, sum1=self.sum1, sum2=self.sum2,
                                        accumulate=not first)
            first = False
            if self.prefetch and i == max(0, nsub - 1 - self.prefetch_after):
                # the next run()'s permutations, into the other buffer (the
                # run before this one, which read it, has ended)
                sp.wait_event(e1)
                with torch.cuda.stream(sp):
                    if self.d_seeds is None:
                        eng.device_permutations(self.seed, self.R,
                                                out=self.perms[1 - self.cur],
                                                workspace=self.perm_ws)
                    else:
                        eng.randomstate_permutations(self.d_seeds,
                                                     out=self.perms[1 - self.cur])
        if self.prefetch:
            self.cur = 1 - self.cur
            self.prefetched = True
        for s in self.streams:
            done = torch.cuda.Event()
            done.record(s)
            main.wait_event(done)
        return self.sum1, self.sum2

    def sweep_ms(self):
        """Summed HIP-event time of the sweeps of the last run() (launches
        on the two sweep streams overlap, so this exceeds their span)."""
        return sum(a.elapsed_time(b) for a, b in self.sweep_events)

    def sweep_span_ms(self):
        """HIP-event time from the start of the first sweep of the last
        run() to the end of its last sweep."""
        if not self.sweep_events:
            return 0.0
        return max(self.sweep_events[0][0].elapsed_time(b)
                   for _, b in self.sweep_events)


class Weights(object):
    """Windowed binomial weights for a set of p on the device."""

    def __init__(self, n, ps, start, length, kmax, w):
        self.n = n
        self.ps = ps            # host float64 (P,)
        self.start = start      # device int64 (P,)
        self.length = length    # device int64 (P,)
        self.host_start = None
        self.host_length = None
        self.kmax = kmax
        self.w = w              # device float64 (P, kmax)


def binomial_windows(n, ps):
    """Host windows (start, length) of the weights for every p."""
    lib = nat.load_library()
    ps = np.atleast_1d(np.asarray(ps, dtype=np.float64))
    start = np.empty(ps.size, dtype=np.int64)
    length = np.empty(ps.size, dtype=np.int64)
    a, b = ctypes.c_int64(), ctypes.c_int64()
    for i, p in enumerate(ps):
        nat.check(lib.pc_binomial_window(int(n), float(p), ctypes.byref(a),
                                         ctypes.byref(b)), 'pc_binomial_window')
        start[i], length[i] = a.value, b.value
    return start, length


def binomial_weights(n, ps, window=True):
    """Device weights B(n, k; p), normalised, for every p: over the window
    |k - np| <= 12 sigma + 48, or over all k = 0..n with window=False."""
    torch = nat.torch_cuda()
    lib = nat.load_library()
    ps = np.atleast_1d(np.asarray(ps, dtype=np.float64))
    if window:
        start, length = binomial_windows(n, ps)
    else:
        if ps.size and not ((ps >= 0) & (ps <= 1)).all():
            raise ValueError('p outside [0, 1]')
        start = np.zeros(ps.size, dtype=np.int64)
        length = np.full(ps.size, int(n) + 1, dtype=np.int64)
    kmax = int(length.max()) if ps.size else 1
    dev = torch.device('cuda', torch.cuda.current_device())
    d_ps = torch.from_numpy(ps.copy()).to(dev)
    d_start = torch.from_numpy(start).to(dev)
    d_len = torch.from_numpy(length).to(dev)
    w = torch.empty((max(ps.size, 1), kmax), dtype=torch.float64, device=dev)
    nat.check(lib.pc_binomial_weights(int(n), ps.size, nat.ptr(d_ps),
                                      nat.ptr(d_start), nat.ptr(d_len), kmax,
                                      nat.ptr(w), nat.stream_handle(torch)),
              'pc_binomial_weights')
    wt = Weights(n, ps, d_start, d_len, kmax, w)
    wt.host_start, wt.host_length = start, length
    return wt


def dense_weights(factors):
    """Caller-supplied convolution factors (length S) as a one-row Weights."""
    torch = nat.torch_cuda()
    f = np.ascontiguousarray(factors, dtype=np.float64).reshape(-1)
    dev = torch.device('cuda', torch.cuda.current_device())
    wt = Weights(f.size - 1, None,
                 torch.zeros(1, dtype=torch.int64, device=dev),
                 torch.full((1,), f.size, dtype=torch.int64, device=dev),
                 max(f.size, 1), torch.from_numpy(f.copy()).to(dev).reshape(1, -1))
    wt.host_start = np.zeros(1, dtype=np.int64)
    wt.host_length = np.array([f.size], dtype=np.int64)
    return wt


def to_f32(t):
    """FP64 device tensor -> FP32 copy (pc_convert_f64_to_f32)."""
    torch = nat.torch_cuda()
    t = t.contiguous()
    out = torch.empty(t.shape, dtype=torch.float32, device=t.device)
    nat.check(nat.load_library().pc_convert_f64_to_f32(
        nat.ptr(t), t.numel(), nat.ptr(out), nat.stream_handle(torch)),
        'pc_convert_f64_to_f32')
    return out


class ContractPlan(object):
    """Inputs of one contraction at one precision, converted once (FP32
    forms read FP32 copies of x and of the weights) with the workspace
    allocated: call run() as often as needed."""

    def __init__(self, x, weights, precision=nat.PC_PREC_FP64):
        torch = nat.torch_cuda()
        self.lib = nat.load_library()
        self.precision = nat.PRECISIONS.get(precision, precision)
        x = x.contiguous()
        self.S, self.C = int(x.shape[0]), int(x.shape[1])
        self.P = (int(weights.w.shape[0]) if weights.ps is None
                  else len(weights.ps))
        if self.P and self.S and (
                int(weights.host_start.min()) < 0 or
                int((weights.host_start + weights.host_length).max()) > self.S):
            raise ValueError('convolution window outside the {} steps'.format(self.S))
        fp32 = self.precision in (nat.PC_PREC_FP32_MFMA, nat.PC_PREC_BF16X3_MFMA)
        self.x = to_f32(x) if fp32 else x
        self.w = to_f32(weights.w) if fp32 else weights.w
        self.weights = weights
        b = ctypes.c_uint64(0)
        nat.check(self.lib.pc_contract_workspace_size(self.P, self.C,
                                                      self.precision,
                                                      ctypes.byref(b)),
                  'pc_contract_workspace_size')
        self.ws_bytes = int(b.value)
        self.ws = (torch.empty((self.ws_bytes,), dtype=torch.uint8,
                               device=x.device) if self.ws_bytes else None)
        self.out = torch.empty((self.P, self.C), dtype=torch.float64,
                               device=x.device)

    def run(self):
        torch = nat.torch_cuda()
        if self.P and self.S and self.C:
            wt = self.weights
            nat.check(self.lib.pc_canonical_contract(
                self.S, self.C, nat.ptr(self.x), self.C, self.P,
                nat.ptr(wt.start), nat.ptr(wt.length), nat.ptr(self.w),
                wt.kmax, nat.ptr(self.out), self.precision, nat.ptr(self.ws),
                self.ws_bytes, nat.stream_handle(torch)),
                'pc_canonical_contract')
        return self.out


def contract(x, weights, precision=nat.PC_PREC_FP64):
    """Canonical convolution of the columns of x (S, C) float64 device
    tensor with every weight row: (P, C) float64 device tensor."""
    return ContractPlan(x, weights, precision).run()


class FusedPlan(object):
    """Canonical averages of the columns of x ((n + 1, C) float64 device
    tensor, row k = state after k bonds) at every p, with the binomial
    weights generated inside the contraction (pc_canonical_fused):
    out[i, c] = sum_k B(n, k; p_i) x[k, c] / sum_k B(n, k; p_i) over the
    window of p_i.  The p-values are processed in sorted order (neighbouring
    windows share the tile's work) and the rows returned in the caller's
    order.  precision: 'fp64_mfma' or 'split_bf16'."""

    def __init__(self, x, ps, precision='fp64_mfma'):
        torch = nat.torch_cuda()
        self.lib = nat.load_library()
        self.precision = nat.PRECISIONS.get(precision, precision)
        if self.precision not in (nat.PC_PREC_FP64_MFMA, nat.PC_PREC_SPLIT_BF16):
            raise ValueError('fused canonical averages run in fp64_mfma or split_bf16')
        self.x = x.contiguous()
        self.S, self.C = int(self.x.shape[0]), int(self.x.shape[1])
        if self.S < 1:
            raise ValueError('x needs at least one step')
        self.n = self.S - 1
        ps = np.atleast_1d(np.asarray(ps, dtype=np.float64))
        self.P = ps.size
        order = np.argsort(ps, kind='stable')
        self.ps = ps
        start, length = binomial_windows(self.n, ps[order])
        dev = self.x.device
        self.d_ps = torch.from_numpy(ps[order].copy()).to(dev)
        self.d_start = torch.from_numpy(start).to(dev)
        self.d_len = torch.from_numpy(length).to(dev)
        self.unsort = torch.from_numpy(np.argsort(order).astype(np.int64)).to(dev)
        b = ctypes.c_uint64(0)
        nat.check(self.lib.pc_canonical_fused_workspace_size(
            self.P, self.C, self.precision, ctypes.byref(b)),
            'pc_canonical_fused_workspace_size')
        self.ws_bytes = int(b.value)
        self.ws = (torch.empty((self.ws_bytes,), dtype=torch.uint8, device=dev)
                   if self.ws_bytes else None)
        self.sorted_out = torch.empty((max(self.P, 1), self.C), dtype=torch.float64,
                                      device=dev)
        self.out = None

    def launch(self):
        """Enqueue the kernels (sorted-order result in self.sorted_out)."""
        torch = nat.torch_cuda()
        if self.P and self.C:
            nat.check(self.lib.pc_canonical_fused(
                self.n, self.C, nat.ptr(self.x), self.C, self.P,
                nat.ptr(self.d_ps), nat.ptr(self.d_start), nat.ptr(self.d_len),
                nat.ptr(self.sorted_out), self.precision, nat.ptr(self.ws),
                self.ws_bytes, nat.stream_handle(torch)), 'pc_canonical_fused')

    def run(self):
        self.launch()
        self.out = self.sorted_out[:self.P].index_select(0, self.unsort)
        return self.out


def canonical_fused(x, ps, precision='fp64_mfma'):
    """See FusedPlan: (P, C) float64 device tensor."""
    return FusedPlan(x, ps, precision).run()


def micro_ci(sum1, sum2, num_edges, num_runs, alpha):
    """The microcanonical statistics of every step with their confidence
    bounds, on the device (pc_microcanonical_ci): (21, S) float64 tensor,
    rows 0..2 max cluster size mean / lower / upper, 3..5 spanning, 6 + 3j
    .. 8 + 3j moment j.  The quantiles come from scipy on the host: Student's
    t at the interval's two ends (R - 1 degrees of freedom) and the beta
    quantiles of the spanning counts that occur."""
    import scipy.stats
    torch = nat.torch_cuda()
    lib = nat.load_library()
    S = num_edges + 1
    t_lo = t_hi = 0.0
    if num_runs > 1:
        t_lo, t_hi = scipy.stats.t.interval(1 - alpha, df=num_runs - 1)
    k = np.unique(sum1[1, :, 0].cpu().numpy().view(np.uint64))
    beta = np.full((num_runs + 1, 2), np.nan)
    q = scipy.stats.beta.ppf(np.array([alpha / 2, 1 - alpha / 2])[:, None],
                             k + 1, num_runs - k + 1)
    beta[k, 0] = q[0]
    beta[k, 1] = q[1]
    d_beta = torch.from_numpy(beta).to(sum1.device)
    out = torch.empty((21, S), dtype=torch.float64, device=sum1.device)
    nat.check(lib.pc_microcanonical_ci(
        num_edges, num_runs, nat.ptr(sum1), nat.ptr(sum2), float(t_lo),
        float(t_hi), nat.ptr(d_beta), nat.ptr(out), nat.stream_handle(torch)),
        'pc_microcanonical_ci')
    return out


def t_interval_quantiles(runs, alpha):
    """Per row, the Student-t quantiles scipy.stats.t.interval(1 - alpha,
    df=runs - 1) multiplies the scale by (its ppf at (1 -+ (1 - alpha)) / 2),
    evaluated once per distinct count; NaN where df <= 0."""
    import scipy.stats
    conf = 1 - alpha
    if not 0 <= conf <= 1:
        raise ValueError('alpha must be between 0 and 1 inclusive')
    dof = np.asarray(runs, dtype=np.float64).reshape(-1) - 1.0
    uniq, inv = np.unique(dof, return_inverse=True)
    q = np.full((2, uniq.size), np.nan)
    good = uniq > 0
    if good.any():
        q[:, good] = scipy.stats.t.ppf(
            np.array([(1.0 - conf) / 2, (1.0 + conf) / 2])[:, None], uniq[good])
    return q[0][inv], q[1][inv]


def finalize_columns(runs, mean, m2, div, alpha):
    """Finalized canonical averages on the device (pc_finalize_canonical):
    runs (P,) counts, mean and m2 (P, C) float64 device tensors, div (C,)
    divisors; returns a (4, P, C) float64 tensor of mean, std, lower and
    upper bound."""
    torch = nat.torch_cuda()
    lib = nat.load_library()
    P, C = mean.shape
    dev = mean.device
    q_lo, q_hi = t_interval_quantiles(runs, alpha)
    d_runs = torch.from_numpy(np.asarray(runs, dtype=np.uint32).reshape(-1)
                              .view(np.int32)).to(dev)
    d_q = torch.from_numpy(np.stack([q_lo, q_hi])).to(dev)
    d_div = torch.from_numpy(np.asarray(div, dtype=np.float64)).to(dev)
    mean = mean.contiguous()
    m2 = m2.contiguous()
    out = torch.empty((4, P, C), dtype=torch.float64, device=dev)
    nat.check(lib.pc_finalize_canonical(
        P, C, nat.ptr(d_runs), nat.ptr(mean), nat.ptr(m2), nat.ptr(d_div),
        nat.ptr(d_q[0]), nat.ptr(d_q[1]), nat.ptr(out),
        nat.stream_handle(torch)), 'pc_finalize_canonical')
    return out


def merge_moments(n_a, mean_a, m2_a, n_b, mean_b, m2_b, out=None):
    """Chan merge of two (n, mean, M2) sets on the device
    (pc_merge_moments): n_a, n_b (P,) float64 device tensors, the others
    (P, C); returns (mean, m2), written into out = (mean, m2) if given (which
    may be the a operands)."""
    torch = nat.torch_cuda()
    lib = nat.load_library()
    P, C = mean_a.shape
    if out is None:
        out = (torch.empty_like(mean_a), torch.empty_like(m2_a))
    nat.check(lib.pc_merge_moments(
        P, C, nat.ptr(n_a), nat.ptr(mean_a), nat.ptr(m2_a), nat.ptr(n_b),
        nat.ptr(mean_b), nat.ptr(m2_b), nat.ptr(out[0]), nat.ptr(out[1]),
        nat.stream_handle(torch)), 'pc_merge_moments')
    return out


def micro_moments(sum1, sum2, num_edges, num_runs):
    """Mean and sample std (ddof=1) per column and step: (7, S) float64."""
    torch = nat.torch_cuda()
    lib = nat.load_library()
    S = num_edges + 1
    mean = torch.empty((nat.PC_COLUMNS, S), dtype=torch.float64,
                       device=sum1.device)
    std = torch.empty_like(mean)
    nat.check(lib.pc_microcanonical_moments(
        num_edges, num_runs, nat.ptr(sum1), nat.ptr(sum2), nat.ptr(mean),
        nat.ptr(std), nat.stream_handle(torch)), 'pc_microcanonical_moments')
    return mean, std

