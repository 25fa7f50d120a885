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

This is the batched form of the reference's jug study
(percolate/share/jugfile.py:166-259: for L in SYSTEM_DIMENSIONS, runs ->
reduce -> canonical averages).
"""

import numpy as np

from . import _graph
from . import _native as nat
from ._engine import BondEngine, FusedPlan, _pipeline_streams, micro_moments
from .autotune import tune_sweep


class _LengthPlan(object):
    def __init__(self, length, runs, seed, ps, tune):
        self.length = length
        self.runs = runs
        self.seed = seed
        eng = self.engine = BondEngine(_graph.square_lattice(length))
        torch = eng.torch
        if tune:
            self.tuning = tune_sweep(eng)
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

    def launch(self):
        """Enqueue this L's pipeline on the current stream."""
        eng = self.engine
        eng.device_permutations(self.seed, self.runs, out=self.perm,
                                workspace=self.perm_ws)
        fused = eng.fuse_chunk_states()
        res = eng.sweep(self.perm, records=self.records, span_step=self.span,
                        workspace=self.workspace,
                        chunk_state=self.state if fused else None, chunks=fused)
        if not fused:
            res.chunk_state = eng.chunk_scan(res, out=self.state)
        eng.microcanonical_sums(res, sum1=self.sum1, sum2=self.sum2)
        self.mean, self.std = micro_moments(self.sum1, self.sum2, eng.M,
                                            self.runs)
        if self.plan is not None:
            self.x.copy_(self.mean.t())
            self.plan.run()


class FiniteSizeBatch(object):
    """See module docstring."""

    def __init__(self, lengths, runs, seed=42, ps=(), tune=True):
        self.torch = torch = nat.torch_cuda()
        if np.isscalar(runs):
            runs = [int(runs)] * len(lengths)
        self.plans = [_LengthPlan(int(L), int(r), seed, ps, tune)
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

    def _launch_all(self):
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
