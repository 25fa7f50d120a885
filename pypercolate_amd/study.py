"""Finite-size study driver: the shape of the reference's jug study
(percolate/share/jugfile.py:28-259) on one GPU, resumable.

For every system size L: draw the run seeds from one numpy RandomState as
the study does (``rng.randint(2**32, size=runs)``, jugfile.py:38-39,224),
split them into tasks (jugfile.py:227), compute every task's canonical
averages with one batched GPU map -> reduce
(``hpc.bond_canonical_averages_batch``: bond_run + bond_reduce of
jugfile.py:57-135), reduce the tasks with ``hpc.bond_reduce``
(jugfile.py:246-250), finalize them (jugfile.py:253-259) and write one
dataset per L named ``"{L}"`` (jugfile.py:138-156).

Each task's reduced averages are written to ``checkpoint_dir`` as it
finishes (``L{L}_task{t}_{key}.npy``, plain .npy with no pickled objects,
``key`` a digest of the parameters the task depends on); a study run again
with the same parameters and checkpoint directory loads the finished tasks
instead of recomputing them.

Output: HDF5 (``percolate_hpc.hdf5``, dataset ``"{L}"``, as the reference
writes it) when h5py is importable, otherwise an .npz archive with one
array per L under the same name.
"""

import hashlib
import os
from functools import reduce

import numpy as np
import scipy.stats

from . import _graph
from . import hpc
from . import percolate

UINT32_MAX = 4294967296
STUDY_SEED = 201508061904 % UINT32_MAX
ALPHA = 2 * scipy.stats.norm.cdf(-1.0)


def study_seeds(dimensions, runs, seed=STUDY_SEED):
    """The run seeds of every size, drawn in order from one RandomState."""
    rng = np.random.RandomState(seed=seed)
    return {L: rng.randint(UINT32_MAX, size=runs) for L in dimensions}


def task_key(ps, seeds, tasks, spanning_cluster):
    """Short digest of everything a task's result depends on (the p grid,
    the size's run seeds, the task split and spanning detection), part of
    every checkpoint's name: a study rerun with other parameters in the same
    directory computes its own tasks instead of loading stale ones."""
    h = hashlib.sha1()
    h.update(np.ascontiguousarray(ps, dtype=np.float64).tobytes())
    h.update(np.ascontiguousarray(seeds, dtype=np.int64).tobytes())
    h.update('{}:{}'.format(int(tasks), bool(spanning_cluster)).encode())
    return h.hexdigest()[:16]


def _task_path(checkpoint_dir, L, t, key):
    return os.path.join(checkpoint_dir, 'L{}_task{}_{}.npy'.format(L, t, key))


def run_study(dimensions=(8, 16, 32), runs=10000, tasks=100,
              ps=np.linspace(0.4, 0.6, num=40), spanning_cluster=True,
              seed=STUDY_SEED, alpha=ALPHA, checkpoint_dir=None,
              output=None, log=None):
    """Run the study; return {L: finalized canonical averages}.

    ``output``: path of the HDF5 file (h5py) or .npz archive to write, or
    None.  ``checkpoint_dir``: directory for per-task results (resume).
    """
    if runs % tasks:
        raise ValueError('runs must split evenly into tasks')
    ps = np.asarray(ps, dtype=np.float64)
    seeds = study_seeds(dimensions, runs, seed)
    if checkpoint_dir:
        os.makedirs(checkpoint_dir, exist_ok=True)
    results = {}
    for L in dimensions:
        graph = percolate.spanning_2d_grid(length=L)
        pg = percolate.percolation_graph(graph=graph,
                                         spanning_cluster=spanning_cluster)
        # the graph's index arrays, read once per size rather than per task
        ga = _graph.from_percolation_graph(**pg)
        reduced = []
        key = task_key(ps, seeds[L], tasks, spanning_cluster)
        want = np.dtype(hpc.canonical_averages_dtype(spanning_cluster))
        for t, chunk in enumerate(np.split(seeds[L], tasks)):
            path = _task_path(checkpoint_dir, L, t, key) if checkpoint_dir else None
            if path and os.path.exists(path):
                done = np.load(path, allow_pickle=False)
                if done.dtype != want or done.shape != ps.shape or \
                        (done['number_of_runs'] != len(chunk)).any():
                    raise ValueError('checkpoint {} does not belong to this '
                                     'study'.format(path))
                reduced.append(done)
                continue
            avg = hpc.bond_canonical_averages_batch(ps=ps, seeds=chunk,
                                                    graph_arrays=ga, **pg)
            if path:
                tmp = path + '.tmp.npy'
                np.save(tmp, avg, allow_pickle=False)
                os.replace(tmp, path)
            reduced.append(avg)
            if log:
                log('L={} task {}/{} done'.format(L, t + 1, tasks))
        total = reduce(hpc.bond_reduce, reduced)
        results[L] = hpc.finalize_canonical_averages(
            number_of_nodes=pg['num_nodes'], ps=ps, canonical_averages=total,
            alpha=alpha)
    if output:
        write_results(output, results)
    return results


def write_results(path, results):
    """One dataset per system size, named "{L}": HDF5 through h5py where it
    is importable (an existing dataset is an error, as in the reference),
    otherwise an .npz archive."""
    try:
        import h5py
    except ImportError:
        h5py = None
    if h5py is not None and not path.endswith('.npz'):
        with h5py.File(path, mode='a') as f:
            for L, arr in results.items():
                key = '{}'.format(L)
                if key in f:
                    raise RuntimeError('dataset {} exists in {}'.format(key, path))
                f.create_dataset(name=key, data=arr)
        return path
    if not path.endswith('.npz'):
        path = path + '.npz'
    np.savez(path, **{'{}'.format(L): arr for L, arr in results.items()})
    return path


def read_results(path):
    """{L: finalized canonical averages} from write_results' output."""
    if path.endswith('.npz'):
        with np.load(path, allow_pickle=False) as z:
            return {int(k): z[k] for k in z.files}
    import h5py
    with h5py.File(path, mode='r') as f:
        return {int(k): f[k][...] for k in f.keys()}
