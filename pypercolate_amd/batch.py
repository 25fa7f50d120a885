"""Finite-size-scaling batch: many lattice sizes in one device program.

``FiniteSizeBatch(lengths, runs)`` holds, for every L, the square lattice
spanning_2d_grid(L), its buffers and its tuned sweep knobs, and runs for
each L the pipeline permute (Philox) -> sweep -> chunk states -> exact
microcanonical sums -> mean per step -> canonical averages at the given p
(FP64 on the matrix cores, binomial weights generated inside the
contraction).  The per-L pipelines are independent: they are issued on
three streams (default priority; the largest L, the next, and the rest in
turn), forked from and joined back into the caller's stream with events, so
the GPU overlaps the long large-L sweeps with the many short small-L ones.
``capture()`` records the whole program once as a HIP graph
(torch.cuda.CUDAGraph); ``run()`` then replays it with a single launch.  All buffers are allocated up front, so a
replay allocates nothing.

With ``persistent=True`` the sweeps of all sizes are one launch instead
(pc_sweep_multi): the permutations of every L first, then one persistent
kernel whose resident workgroups take the runs of all lattices, largest
first, from a device work queue, then the per-L reductions on the three
streams.

This is the batched form of the reference's jug study
(percolate/share/jugfile.py:166-259: for L in SYSTEM_DIMENSIONS, runs ->
reduce -> canonical averages).
"""

import ctypes

import numpy as np

from . import _graph
from . import _native as nat
from ._engine import (BondEngine, FusedPlan, SweepResult, _pipeline_streams,
                      micro_moments)
from .autotune import tune_sweep


class _LengthPlan(object):
    def __init__(self, length, runs, seed, ps, tune):
        self.length = length
        self.runs = runs
        self.seed = seed
        eng = self.engine = BondEngine(_graph.square_lattice(length))
        torch = eng.torch
        if tune:
            self.tuning = tune_sweep(eng, max_runs=runs)
        else:
            self.tuning = None
        dev = eng.device
        M, S = eng.M, eng.S
        C = max((M + nat.PC_CHUNK - 1) // nat.PC_CHUNK, 1)
        self.perm = torch.empty((runs, max(M, 1)), dtype=torch.int32, device=dev)
        self.records = torch.empty((runs, max(M, 1)), dtype=eng.rec_torch,
                                   device=dev)
        self.span = torch.empty((runs,), dtype=torch.int32, device=dev)
        self.state = torch.empty((runs, C, nat.PC_STATE_WORDS), dtype=torch.int64,
                                 device=dev)
        self.sum1 = torch.empty((nat.PC_COLUMNS, S, 2), dtype=torch.int64,
                                device=dev)
        self.sum2 = torch.empty((nat.PC_COLUMNS, S, 3), dtype=torch.int64,
                                device=dev)
        self.perm_ws = eng.permutation_workspace(runs)
        wb = eng.workspace_bytes(runs)
        self.workspace = (torch.empty((wb,), dtype=torch.uint8, device=dev)
                          if wb else None)
        self.ps = np.asarray(ps, dtype=np.float64)
        self.x = torch.empty((S, nat.PC_COLUMNS), dtype=torch.float64, device=dev)
        self.plan = (FusedPlan(self.x, self.ps, 'fp64_mfma')
                     if self.ps.size else None)
        self.mean = None
        self.std = None

    def permute(self):
        self.engine.device_permutations(self.seed, self.runs, out=self.perm,
                                        workspace=self.perm_ws)

    def launch(self):
        """Enqueue this L's pipeline on the current stream."""
        eng = self.engine
        self.permute()
        fused = eng.fuse_chunk_states()
        res = eng.sweep(self.perm, records=self.records, span_step=self.span,
                        workspace=self.workspace,
                        chunk_state=self.state if fused else None, chunks=fused)
        if not fused:
            res.chunk_state = eng.chunk_scan(res, out=self.state)
        self.reduce(res)

    def swept(self):
        """The sweep result of this L's buffers (after a persistent sweep,
        which writes records, span steps and chunk states)."""
        res = SweepResult(self.engine, self.runs, self.records, self.span,
                          self.perm)
        res.chunk_state = self.state
        return res

    def reduce(self, res):
        """Exact sums, mean per step and canonical averages of a swept L."""
        eng = self.engine
        eng.microcanonical_sums(res, sum1=self.sum1, sum2=self.sum2)
        self.mean, self.std = micro_moments(self.sum1, self.sum2, eng.M,
                                            self.runs)
        if self.plan is not None:
            self.x.copy_(self.mean.t())
            self.plan.run()

    def sweep_graph(self):
        """This L as a pc_sweep_graph of a persistent multi-graph sweep."""
        eng = self.engine
        stride = max(eng.M, 1)
        return nat.SweepGraph(
            num_nodes=eng.N, num_edges=eng.M, num_runs=self.runs,
            span_on_first_merge=eng.span_first, merge_rounds=eng.merge_rounds,
            # computed lattice edges for the HBM-state lattices (in the tiled
            # order the persistent sweep assumes there), as the engine uses
            lattice_length=(eng.lattice_length
                            if eng.lattice_sweep and eng.lattice_tiled else 0),
            d_edges=nat.ptr(eng.d_edges),
            d_node_sides=nat.ptr(eng.d_sides), d_perm=nat.ptr(self.perm),
            perm_stride=stride, d_records=nat.ptr(self.records),
            rec_stride=stride, d_span_step=nat.ptr(self.span),
            d_chunk_state=nat.ptr(self.state))


class FiniteSizeBatch(object):
    """See module docstring."""

    def __init__(self, lengths, runs, seed=42, ps=(), tune=True,
                 persistent=False):
        self.torch = torch = nat.torch_cuda()
        if np.isscalar(runs):
            runs = [int(runs)] * len(lengths)
        self.plans = [_LengthPlan(int(L), int(r), seed, ps, tune and not persistent)
                      for L, r in zip(lengths, runs)]
        # largest lattices first: their sweeps are the critical path.  Three
        # streams besides the caller's (one per hardware queue of the
        # process: streams that share a queue serialise): the largest L on
        # the first, the next on the second, all smaller ones in turn on
        # the third
        self.plans.sort(key=lambda p: -p.engine.M)
        self.streams = _pipeline_streams(torch, self.plans[0].engine.device)[
            :min(3, len(self.plans))]
        self.stream_of = [min(i, len(self.streams) - 1)
                          for i in range(len(self.plans))]
        self.graph = None
        self.persistent = bool(persistent)
        if self.persistent:
            self._prepare_multi()

    def _prepare_multi(self):
        """Graph table (largest L first), queue head and HBM-state slots of
        the persistent sweep."""
        torch = self.torch
        lib = self.plans[0].engine.lib
        dev = self.plans[0].engine.device
        n = len(self.plans)
        self.multi_graphs = (nat.SweepGraph * n)(
            *[p.sweep_graph() for p in self.plans])
        tb, wb = ctypes.c_uint64(0), ctypes.c_uint64(0)
        nat.check(lib.pc_sweep_multi_workspace_size(
            n, self.multi_graphs, ctypes.byref(tb), ctypes.byref(wb)),
            'pc_sweep_multi_workspace_size')
        self.multi_table = torch.empty((tb.value,), dtype=torch.uint8, device=dev)
        self.multi_ws = (torch.empty((wb.value,), dtype=torch.uint8, device=dev)
                         if wb.value else None)
        nat.check(lib.pc_sweep_multi_prepare(
            n, self.multi_graphs, nat.ptr(self.multi_table), tb.value),
            'pc_sweep_multi_prepare')

    def _launch_persistent(self):
        torch = self.torch
        lib = self.plans[0].engine.lib
        main = torch.cuda.current_stream()
        for plan in self.plans:
            plan.permute()
        nat.check(lib.pc_sweep_multi(
            len(self.plans), self.multi_graphs, nat.ptr(self.multi_table),
            int(self.multi_table.numel()), nat.ptr(self.multi_ws),
            0 if self.multi_ws is None else int(self.multi_ws.numel()),
            nat.stream_handle(torch)), 'pc_sweep_multi')
        fork = torch.cuda.Event()
        fork.record(main)
        joins = []
        for s in self.streams:
            s.wait_event(fork)
        for i, plan in enumerate(self.plans):
            stream = self.streams[self.stream_of[i]]
            with torch.cuda.stream(stream):
                plan.reduce(plan.swept())
                done = torch.cuda.Event()
                done.record(stream)
            joins.append(done)
        for done in joins:
            main.wait_event(done)

    def _launch_all(self):
        if self.persistent:
            self._launch_persistent()
            return
        torch = self.torch
        main = torch.cuda.current_stream()
        fork = torch.cuda.Event()
        fork.record(main)
        joins = []
        for s in self.streams:
            s.wait_event(fork)
        for i, plan in enumerate(self.plans):
            stream = self.streams[self.stream_of[i]]
            with torch.cuda.stream(stream):
                plan.launch()
                done = torch.cuda.Event()
                done.record(stream)
            joins.append(done)
        for done in joins:
            main.wait_event(done)

    def capture(self):
        """Record the whole batch as one HIP graph (after one eager run
        that warms up every kernel)."""
        torch = self.torch
        self._launch_all()
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            self._launch_all()
        self.graph = graph
        return graph

    def run(self):
        if self.graph is not None:
            self.graph.replay()
        else:
            self._launch_all()

    def run_sequential(self):
        """The plain per-L form: every L's pipeline in turn on the current
        stream (no per-L streams, no graph)."""
        for plan in self.plans:
            plan.launch()

    @property
    def additions(self):
        return float(sum(p.runs * p.engine.M for p in self.plans))

    def results(self):
        """Per L: exact sums (host limbs), mean per step and canonical
        averages of the 7 observables at ps (None without ps)."""
        out = {}
        for p in self.plans:
            out[p.length] = {
                'runs': p.runs,
                'sum1': p.sum1.cpu().numpy().view(np.uint64).copy(),
                'sum2': p.sum2.cpu().numpy().view(np.uint64).copy(),
                'mean': p.mean.cpu().numpy().copy(),
                'canonical': (p.plan.out.cpu().numpy().copy()
                              if p.plan is not None else None),
                'tuning': p.tuning,
            }
        return out


def default_runs(lengths, total=1 << 26):
    """Runs per L of the finite-size study: total / L^2."""
    return [max(1, total // (L * L)) for L in lengths]
