"""GPU parity where the integers are wide: moments past 2**64 (128-bit
cluster-size moments, 192-bit sums of squares), the L = 1024 lattice, the
L = 128 configuration at its full 10^4 runs, and the pre-joined spanning
structure.  Integer results must be bit-exact against the oracle."""
import numpy as np
import pytest

from oracle import oracle
from pypercolate_amd import _graph
from pypercolate_amd._engine import BondEngine, host_sums, micro_moments
from tests.helpers import lattice_oracle_graph
from tests.test_oracle_exact import (BIG, N_CHAIN, SIZE_SETS, _chain_runs,
                                     _check_steps, chain_graph)

pytestmark = pytest.mark.gpu


def _limbs(t):
    return t.cpu().numpy().view(np.uint64)


def _chain_engine(n):
    return BondEngine(_graph.GraphArrays(n, np.stack(
        [np.arange(n - 1), np.arange(1, n)], axis=1)))


def test_chain_moments_past_2_64(gpu):
    """Chains whose non-largest clusters pass 2**16 nodes (HBM state,
    64-bit records): the traces' moments equal the oracle's mod 2**64, the
    exact sums over the three runs equal the oracle's 128/192-bit sums, and
    each run alone (R = 1, so the sum of x is x) reproduces the oracle's
    exact 128-bit moments at every step."""
    n = N_CHAIN
    g = chain_graph(n)
    perms = _chain_runs(n, SIZE_SETS, 3)
    expect = oracle.bond_runs(g, perms, exact=True)
    eng = _chain_engine(n)
    res = eng.sweep(eng.upload_permutations(perms))
    mx, sp, mom = eng.host_traces(res)
    np.testing.assert_array_equal(mx, expect['max_cluster_size'])
    np.testing.assert_array_equal(mom, expect['moments'])
    s1, s2 = eng.microcanonical_sums(res)
    o1, o2 = oracle.microcanonical_sums(expect, limbs=True)
    assert np.array_equal(_limbs(s1), o1)
    assert np.array_equal(_limbs(s2), o2)
    assert o1[2:, :, 1].any(), 'no moment passes 2**64'
    for r in range(perms.shape[0]):
        one = eng.sweep(eng.upload_permutations(perms[r:r + 1]))
        a1, _ = eng.microcanonical_sums(one)
        a1 = _limbs(a1)
        np.testing.assert_array_equal(a1[2:, :, 0], expect['moments'][r].T)
        np.testing.assert_array_equal(a1[2:, :, 1], expect['moments_hi'][r].T)


def test_chain_mean_and_std_from_wide_sums(gpu):
    """pc_microcanonical_moments on sums past 2**128: mean and sample
    standard deviation equal the exact rational values (Python ints) to
    1e-15 relative."""
    from fractions import Fraction
    n = N_CHAIN
    perms = _chain_runs(n, SIZE_SETS, 5)
    eng = _chain_engine(n)
    res = eng.sweep(eng.upload_permutations(perms))
    d1, d2 = eng.microcanonical_sums(res)
    mean, std = micro_moments(d1, d2, eng.M, perms.shape[0])
    mean, std = mean.cpu().numpy(), std.cpu().numpy()
    s1, s2 = host_sums(d1, d2)
    R = perms.shape[0]
    for c in range(7):
        for s in _check_steps(n):
            m = Fraction(int(s1[c, s]), R)
            var = Fraction(R * int(s2[c, s]) - int(s1[c, s]) ** 2, R * (R - 1))
            assert mean[c, s] == pytest.approx(float(m), rel=1e-15, abs=0)
            assert std[c, s] == pytest.approx(float(var) ** 0.5, rel=1e-15, abs=1e-300)
    assert max(int(v) for v in s2[6]) >= 1 << 128


def test_L1024_randomstate_runs_bit_exact(gpu):
    """The L = 1024 lattice (N = 1,048,576, M = 2,095,104: HBM state, tiled
    node labels, 64-bit records) with RandomState(r).permutation(M),
    r = 0..3: every run's trace equals the oracle's (moments mod 2**64), and
    the exact sums, over the four runs and over run 0 alone, equal the
    oracle's 128/192-bit sums of its exact moments."""
    L = 1024
    g = lattice_oracle_graph(L)
    perms = oracle.randomstate_permutations(range(4), g['num_edges'])
    expect = oracle.bond_runs(g, perms, threads=4, exact=True)
    eng = BondEngine(_graph.square_lattice(L))
    assert eng.relabelled
    d_perm = eng.upload_permutations(perms)
    res = eng.sweep(d_perm)
    mx, sp, mom = eng.host_traces(res)
    np.testing.assert_array_equal(mx, expect['max_cluster_size'])
    np.testing.assert_array_equal(sp, expect['has_spanning_cluster'])
    np.testing.assert_array_equal(mom, expect['moments'])
    del mx, sp, mom
    s1, s2 = eng.microcanonical_sums(res)
    o1, o2 = oracle.microcanonical_sums(expect, limbs=True)
    assert np.array_equal(_limbs(s1), o1)
    assert np.array_equal(_limbs(s2), o2)
    one = eng.sweep(d_perm[:1])
    a1, a2 = eng.microcanonical_sums(one)
    first = {k: v[:1] for k, v in expect.items()}
    b1, b2 = oracle.microcanonical_sums(first, limbs=True)
    assert np.array_equal(_limbs(a1), b1)
    assert np.array_equal(_limbs(a2), b2)


def test_L128_10000_randomstate_runs_bit_exact(gpu):
    """The L = 128 configuration at full size: 10^4 runs with
    RandomState(r).permutation(32512), r = 0..9999, through the pipeline
    (device MT19937 permutations, sub-batches on four streams); its exact
    sums equal the oracle's over the same permutations drawn by numpy on
    the host, swept in chunks of 256 runs."""
    import torch
    from pypercolate_amd._engine import Pipeline
    R = 10000
    eng = BondEngine(_graph.square_lattice(128))
    pipe = Pipeline(eng, R, seeds=np.arange(R))
    s1, s2 = pipe.run()
    torch.cuda.synchronize()
    g = lattice_oracle_graph(128)
    sums = None
    for first in range(0, R, 256):
        perms = oracle.randomstate_permutations(
            range(first, min(first + 256, R)), g['num_edges'])
        sums = oracle.bond_sums(g, perms, sums=sums, threads=16)
    assert np.array_equal(_limbs(s1), sums[0])
    assert np.array_equal(_limbs(s2), sums[1])
    # the device permutations of a few runs are numpy's
    got = pipe.perm[[0, 4321, R - 1]].cpu().numpy()
    for row, r in zip(got, (0, 4321, R - 1)):
        np.testing.assert_array_equal(row, np.random.RandomState(r).permutation(eng.M))


def _prejoined_graph():
    """A path of 12 nodes whose node 5 has auxiliary edges to both sides:
    the auxiliary structure alone joins the sides, so the reference reports
    a spanning cluster from the first merge on (percolate/hpc.py:217-218,
    269-274), also on merges far from node 5."""
    import networkx as nx
    gnx = nx.path_graph(12)
    gnx.add_node('a', span=0)
    gnx.add_node('b', span=1)
    gnx.add_edge('a', 5, span=0)
    gnx.add_edge('b', 5, span=1)
    return gnx


def test_span_on_first_merge(gpu):
    from pypercolate_amd.percolate import percolation_graph
    gnx = _prejoined_graph()
    pg = percolation_graph(gnx, spanning_cluster=True)
    ga = _graph.from_percolation_graph(**pg)
    assert ga.span_on_first_merge
    g = oracle.percolation_graph_arrays(gnx, True)
    perms = oracle.randomstate_permutations(range(16), g['num_edges'])
    expect = oracle.bond_runs(g, perms)
    assert not expect['has_spanning_cluster'][:, 0].any()
    assert expect['has_spanning_cluster'][:, 1:].all()
    eng = BondEngine(ga)
    res = eng.sweep(eng.upload_permutations(perms))
    mx, sp, mom = eng.host_traces(res)
    np.testing.assert_array_equal(sp, expect['has_spanning_cluster'])
    np.testing.assert_array_equal(mx, expect['max_cluster_size'])
    np.testing.assert_array_equal(mom, expect['moments'])
    s1, s2 = host_sums(*eng.microcanonical_sums(res))
    o1, o2 = oracle.microcanonical_sums(expect)
    assert (s1 == o1).all() and (s2 == o2).all()


def test_single_run_arrays_moments_past_2_64(gpu):
    """percolate.single_run_arrays reports the run's exact moments rounded
    to float64, also where they pass 2**64 (a chain of 400000 nodes near
    its end: a few clusters of ~10^5 nodes), against the oracle's exact
    128-bit moments of the same RandomState draw."""
    import networkx as nx
    from pypercolate_amd import percolate
    n = 400000
    np.random.seed(3)
    got = percolate.single_run_arrays(graph=nx.path_graph(n), spanning_cluster=False)
    np.random.seed(3)
    perm = np.random.permutation(n - 1)
    tr = oracle.bond_runs(chain_graph(n), perm[None, :], exact=True)
    exact = oracle.exact_moments(tr)[0]                      # (S, 5) ints
    np.testing.assert_array_equal(got['max_cluster_size'], tr['max_cluster_size'][0])
    want = np.array([[float(v) for v in row] for row in exact.T])
    np.testing.assert_allclose(got['moments'], want, rtol=2.3e-16, atol=0)
    assert max(int(v) for v in exact[:, 4]) >= 1 << 64
