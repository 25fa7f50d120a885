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
