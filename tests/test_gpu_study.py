"""The study driver on the GPU: tasks + reduce equal one batch over all the
seeds, and a second run resumes from the task checkpoints."""
import numpy as np
import pytest

from pypercolate_amd import hpc, percolate, study

pytestmark = pytest.mark.gpu


def _fields(a):
    return [n for n in a.dtype.names if n not in ('number_of_runs',)]


def test_study_tasks_reduce_and_resume(gpu, tmp_path, monkeypatch):
    ps = np.linspace(0.4, 0.6, 11)
    ckpt = str(tmp_path / 'ckpt')
    out = str(tmp_path / 'study.npz')
    res = study.run_study(dimensions=(4, 6), runs=40, tasks=4, ps=ps,
                          checkpoint_dir=ckpt, output=out)
    seeds = study.study_seeds((4, 6), 40)
    for L in (4, 6):
        pg = percolate.percolation_graph(percolate.spanning_2d_grid(L))
        whole = hpc.bond_canonical_averages_batch(ps=ps, seeds=seeds[L], **pg)
        fin = hpc.finalize_canonical_averages(
            number_of_nodes=pg['num_nodes'], ps=ps, canonical_averages=whole,
            alpha=study.ALPHA)
        assert (res[L]['number_of_runs'] == 40).all()
        for name in _fields(fin):
            np.testing.assert_allclose(res[L][name], fin[name], rtol=1e-10,
                                       atol=1e-14, err_msg=name)
    # resume: every task is on disk, nothing is recomputed
    def boom(*a, **k):
        raise AssertionError('recomputed a finished task')
    monkeypatch.setattr(hpc, 'bond_canonical_averages_batch', boom)
    again = study.run_study(dimensions=(4, 6), runs=40, tasks=4, ps=ps,
                            checkpoint_dir=ckpt)
    back = study.read_results(out)
    for L in (4, 6):
        np.testing.assert_array_equal(again[L], res[L])
        np.testing.assert_array_equal(back[L], res[L])


def test_study_without_spanning(gpu):
    """spanning_cluster=False: the whole grid (auxiliary rows included) is
    the percolation graph; the study equals one batch over all the seeds."""
    ps = np.linspace(0.4, 0.6, 5)
    res = study.run_study(dimensions=(5,), runs=20, tasks=2, ps=ps,
                          spanning_cluster=False)
    seeds = study.study_seeds((5,), 20)
    pg = percolate.percolation_graph(percolate.spanning_2d_grid(5),
                                     spanning_cluster=False)
    whole = hpc.bond_canonical_averages_batch(ps=ps, seeds=seeds[5], **pg)
    fin = hpc.finalize_canonical_averages(
        number_of_nodes=pg['num_nodes'], ps=ps, canonical_averages=whole,
        alpha=study.ALPHA)
    assert 'percolation_probability_mean' not in fin.dtype.names
    for name in _fields(fin):
        np.testing.assert_allclose(res[5][name], fin[name], rtol=1e-10,
                                   atol=1e-14, err_msg=name)


def _oracle_study_size(L, seeds, ps, tasks, spanning):
    """The reference study for one size restated on the oracle
    (percolate/share/jugfile.py:57-135, 240-259): per seed the run's traces
    (RandomState(seed) permutation), its canonical statistics at every p
    (hpc.py:443-515), initialized (hpc.py:561-635) and reduced per task and
    then over tasks (hpc.py:638-702), finalized (hpc.py:752-834)."""
    import functools
    from oracle import oracle
    from tests.helpers import spanning_2d_grid
    g = oracle.percolation_graph_arrays(spanning_2d_grid(L), spanning)
    M = g['num_edges']
    pmf = [oracle.binomial_pmf(M, p) for p in ps]
    task_avgs = []
    for chunk in np.split(np.asarray(seeds), tasks):
        tr = oracle.bond_runs(g, oracle.randomstate_permutations(chunk, M))
        inits = []
        for r in range(len(chunk)):
            one = {k: v[r] for k, v in tr.items()}
            if not spanning:
                one.pop('has_spanning_cluster')
            inits.append(oracle.initialize_canonical_averages(
                [oracle.canonical_statistics(one, f) for f in pmf]))
        task_avgs.append(functools.reduce(oracle.bond_reduce, inits))
    total = functools.reduce(oracle.bond_reduce, task_avgs)
    return oracle.finalize_canonical_averages(g['num_nodes'], ps, total, study.ALPHA)


@pytest.mark.parametrize('spanning', [True, False])
def test_study_matches_oracle_map_reduce_finalize(gpu, spanning):
    """run_study (batched GPU tasks, bond_reduce over tasks, finalize) against
    the oracle's run-by-run map -> reduce -> finalize of the same seeds."""
    ps = np.linspace(0.35, 0.65, 9)
    res = study.run_study(dimensions=(5, 7), runs=36, tasks=3, ps=ps,
                          spanning_cluster=spanning)
    seeds = study.study_seeds((5, 7), 36)
    for L in (5, 7):
        want = _oracle_study_size(L, seeds[L], ps, 3, spanning)
        got = res[L]
        assert (got['number_of_runs'] == 36).all()
        np.testing.assert_array_equal(got['p'], ps)
        for key, val in want.items():
            if key in ('number_of_runs', 'p', 'alpha'):
                continue
            scale = np.abs(val).max() or 1.0
            np.testing.assert_allclose(got[key], val, rtol=1e-9,
                                       atol=1e-12 * scale, err_msg=key)
