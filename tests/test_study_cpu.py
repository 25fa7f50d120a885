"""Study driver host logic: the reference study's seed stream and the
on-disk format (no GPU)."""
import numpy as np

from pypercolate_amd import hpc, study


def test_study_seeds_are_the_reference_stream():
    """percolate/share/jugfile.py:38-39,188,224: one RandomState(SEED),
    randint(2**32, size=NUMBER_OF_RUNS) per system size, in order."""
    rng = np.random.RandomState(201508061904 % 4294967296)
    expect = [rng.randint(4294967296, size=50) for _ in range(3)]
    got = study.study_seeds([8, 16, 32], 50)
    for L, e in zip([8, 16, 32], expect):
        np.testing.assert_array_equal(got[L], e)
    assert study.STUDY_SEED == 201508061904 % 4294967296


def test_results_round_trip(tmp_path):
    dtype = hpc.finalized_canonical_averages_dtype(spanning_cluster=True)
    res = {}
    for L in (8, 16):
        a = np.zeros(5, dtype=dtype)
        a['p'] = np.linspace(0.4, 0.6, 5)
        a['number_of_runs'] = L
        a['moments_mean'] = np.arange(25.0).reshape(5, 5) * L
        res[L] = a
    path = study.write_results(str(tmp_path / 'study'), res)
    back = study.read_results(path)
    assert sorted(back) == [8, 16]
    for L in res:
        assert back[L].dtype == dtype
        np.testing.assert_array_equal(back[L], res[L])


def test_checkpoint_names_depend_on_the_parameters():
    """A rerun with other p, seeds, task split or spanning flag writes and
    reads other checkpoint files (no stale task is reduced in)."""
    ps = np.linspace(0.4, 0.6, 5)
    seeds = np.arange(20)
    base = study.task_key(ps, seeds, 4, True)
    assert base == study.task_key(ps.copy(), seeds.copy(), 4, True)
    others = {study.task_key(ps[:4], seeds, 4, True),
              study.task_key(ps, seeds + 1, 4, True),
              study.task_key(ps, seeds, 5, True),
              study.task_key(ps, seeds, 4, False)}
    assert base not in others and len(others) == 4
