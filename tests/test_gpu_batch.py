"""Finite-size batch and autotuner: the batched, graph-replayed program is
bit-identical to the unbatched per-L path, every replay equals the eager
run, and tuned sweep knobs never change the merge records."""
import numpy as np
import pytest

from pypercolate_amd import _graph
from pypercolate_amd import _native as nat
from pypercolate_amd._engine import BondEngine
from pypercolate_amd.autotune import tune_sweep
from pypercolate_amd.batch import FiniteSizeBatch

pytestmark = pytest.mark.gpu

LENGTHS = [8, 16, 32, 130]
RUNS = [96, 40, 24, 3]
PS = np.linspace(0.4, 0.6, 11)


def _unbatched_sums(L, runs, seed):
    eng = BondEngine(_graph.square_lattice(L))
    res = eng.sweep(eng.device_permutations(seed, runs))
    s1, s2 = eng.microcanonical_sums(res)
    return s1.cpu().numpy().view(np.uint64), s2.cpu().numpy().view(np.uint64)


def test_batch_graph_matches_unbatched(gpu, tmp_path, monkeypatch):
    monkeypatch.setenv('PYPERCOLATE_AMD_CACHE', str(tmp_path))
    batch = FiniteSizeBatch(LENGTHS, RUNS, seed=42, ps=PS, tune=True)
    batch.run()
    gpu.cuda.synchronize()
    eager = batch.results()
    for L, runs in zip(LENGTHS, RUNS):
        s1, s2 = _unbatched_sums(L, runs, 42)
        assert np.array_equal(eager[L]['sum1'], s1), L
        assert np.array_equal(eager[L]['sum2'], s2), L
        assert eager[L]['canonical'].shape == (PS.size, 7)
        assert np.isfinite(eager[L]['canonical']).all()
    batch.capture()
    for k in range(3):
        batch.run()
        gpu.cuda.synchronize()
        junk = gpu.ones(1 << 20, device='cuda').sum().item()  # eager work between
        assert junk == float(1 << 20)
        got = batch.results()
        for L in LENGTHS:
            for key in ('sum1', 'sum2', 'mean', 'canonical'):
                assert np.array_equal(got[L][key], eager[L][key]), (k, L, key)


def test_persistent_batch_matches_per_length(gpu):
    """One persistent sweep over lattices of every state kind (LDS codes with
    32-bit records at L = 8, 16; HBM codes with 32-bit records at L = 130,
    64-bit at L = 257) gives the per-L path's exact sums, means and
    canonical averages bit for bit, eagerly and on every HIP-graph replay."""
    lengths, runs = [8, 16, 130, 257], [96, 40, 5, 3]
    plain = FiniteSizeBatch(lengths, runs, seed=7, ps=PS, tune=False)
    plain.run_sequential()
    gpu.cuda.synchronize()
    want = plain.results()
    pers = FiniteSizeBatch(lengths, runs, seed=7, ps=PS, persistent=True)
    assert [p.length for p in pers.plans] == [257, 130, 16, 8]
    pers.run()
    gpu.cuda.synchronize()
    for L in lengths:
        got = pers.results()[L]
        for key in ('sum1', 'sum2', 'mean', 'canonical'):
            assert np.array_equal(got[key], want[L][key]), (L, key)
        plan = [p for p in pers.plans if p.length == L][0]
        ref = [p for p in plain.plans if p.length == L][0]
        assert np.array_equal(plan.records.cpu().numpy(), ref.records.cpu().numpy())
        assert np.array_equal(plan.span.cpu().numpy(), ref.span.cpu().numpy())
    pers.capture()
    for k in range(3):
        pers.run()
        gpu.cuda.synchronize()
        junk = gpu.ones(1 << 20, device='cuda').sum().item()
        assert junk == float(1 << 20)
        got = pers.results()
        for L in lengths:
            for key in ('sum1', 'sum2', 'mean', 'canonical'):
                assert np.array_equal(got[L][key], want[L][key]), (k, L, key)


def test_persistent_sweep_rejects_bad_tables(gpu):
    """Host-side checks of pc_sweep_multi: no graphs, a table that is too
    small, a missing output buffer."""
    import ctypes
    lib = nat.load_library()
    tb, wb = ctypes.c_uint64(0), ctypes.c_uint64(0)
    assert lib.pc_sweep_multi_workspace_size(0, None, ctypes.byref(tb),
                                             ctypes.byref(wb)) == -1
    batch = FiniteSizeBatch([8], [4], seed=1, persistent=True)
    g = batch.multi_graphs
    t = batch.multi_table
    assert lib.pc_sweep_multi(1, g, nat.ptr(t), 16, None, 0, None) == -1
    bad = (nat.SweepGraph * 1)(g[0])
    bad[0].d_span_step = None
    assert lib.pc_sweep_multi_workspace_size(1, bad, ctypes.byref(tb),
                                             ctypes.byref(wb)) == -1


@pytest.mark.parametrize('L', [16, 64, 130])
def test_tuned_knobs_do_not_change_records(gpu, tmp_path, monkeypatch, L):
    monkeypatch.setenv('PYPERCOLATE_AMD_CACHE', str(tmp_path))
    eng = BondEngine(_graph.square_lattice(L))
    perm = eng.device_permutations(3, 50)
    base = eng.sweep(perm)
    ref = base.records.cpu().numpy().copy(), base.span_step.cpu().numpy().copy()
    choice = tune_sweep(eng, sample_runs=64, rounds=(1, 3, 8))
    assert choice['merge_rounds'] in (1, 3, 8)
    places = [nat.PC_STATE_HBM] + ([nat.PC_STATE_LDS] if eng.N <= 16384 else [])
    modes = [p | w for p in places
             for w in (0, nat.PC_SWEEP_ONE_WAVE, nat.PC_SWEEP_TWO_WAVES)]
    for mode in modes:
        for r in (1, 5, 16):
            eng.state_mode, eng.merge_rounds = mode, r
            for chunks in (False, True):
                res = eng.sweep(perm, chunks=chunks)
                assert np.array_equal(res.records.cpu().numpy(), ref[0]), (mode, r)
                assert np.array_equal(res.span_step.cpu().numpy(), ref[1]), (mode, r)
                if chunks:
                    assert np.array_equal(res.chunk_state.cpu().numpy(),
                                          eng.chunk_scan(res).cpu().numpy()), (mode, r)
    # cached: a second tuning returns the stored choice
    assert tune_sweep(eng, sample_runs=64, rounds=(1, 3, 8)) == choice
