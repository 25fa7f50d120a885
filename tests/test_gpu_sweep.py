"""GPU parity of the sweep pipeline (permute -> sweep -> chunk_scan ->
microcanonical sums / traces) against the oracle and the golden fixtures.
Integer results must be bit-exact."""
import numpy as np
import pytest

from oracle import oracle
from pypercolate_amd import _graph
from pypercolate_amd._engine import BondEngine, host_sums
from tests.helpers import grid_3x3_graph, spanning_2d_grid

pytestmark = pytest.mark.gpu


def _engine_for(nx_graph, spanning):
    from pypercolate_amd.percolate import percolation_graph
    pg = percolation_graph(nx_graph, spanning_cluster=spanning)
    ga = _graph.from_percolation_graph(**pg)
    return BondEngine(ga), oracle.percolation_graph_arrays(nx_graph, spanning)


def _check_traces(eng, res, expect, runs=None):
    mx, sp, mom = eng.host_traces(res)
    rows = slice(None) if runs is None else runs
    np.testing.assert_array_equal(mx, expect['max_cluster_size'][rows])
    np.testing.assert_array_equal(sp, expect['has_spanning_cluster'][rows])
    np.testing.assert_array_equal(mom, expect['moments'][rows])


def _check_sums(eng, res, expect):
    s1, s2 = host_sums(*eng.microcanonical_sums(res))
    o1, o2 = oracle.microcanonical_sums(expect)
    assert (s1 == o1).all(), 'sum of x differs'
    assert (s2 == o2).all(), 'sum of x**2 differs'


@pytest.mark.parametrize('spanning', [True, False])
def test_kat_3x3(gpu, golden, spanning):
    kat = golden('kat_3x3.json')
    eng, g = _engine_for(grid_3x3_graph(spanning, kat), spanning)
    perm = np.array([kat['permutation']])
    res = eng.sweep(eng.upload_permutations(perm))
    mx, sp, mom = eng.host_traces(res)
    assert mx[0].tolist() == kat['max_cluster_size']
    assert mom[0].tolist() == kat['moments']
    if spanning:
        assert sp[0].tolist() == kat['has_spanning_cluster']


@pytest.mark.parametrize('length', [2, 3, 4, 6])
def test_brute_force_fixtures(gpu, golden, length):
    bf = golden('brute_force.npz')
    eng, g = _engine_for(spanning_2d_grid(length), True)
    res = eng.sweep(eng.upload_permutations(bf['L{}_perm'.format(length)]))
    mx, sp, mom = eng.host_traces(res)
    np.testing.assert_array_equal(mx, bf['L{}_max'.format(length)])
    np.testing.assert_array_equal(mom, bf['L{}_moments'.format(length)])
    np.testing.assert_array_equal(sp, bf['L{}_spanning'.format(length)])


@pytest.mark.parametrize('spanning', [True, False])
def test_config_L16_40_runs(gpu, spanning):
    """spanning_2d_grid(16), 40 runs with RandomState(r) permutations."""
    eng, g = _engine_for(spanning_2d_grid(16), spanning)
    perms = oracle.randomstate_permutations(range(40), g['num_edges'])
    expect = oracle.bond_runs(g, perms)
    res = eng.sweep(eng.upload_permutations(perms))
    _check_traces(eng, res, expect)
    _check_sums(eng, res, expect)


def test_closed_form_lattice_equals_networkx(gpu):
    eng_nx, g = _engine_for(spanning_2d_grid(12), True)
    eng_cf = BondEngine(_graph.square_lattice(12))
    perms = oracle.randomstate_permutations(range(8), g['num_edges'])
    a = eng_nx.host_traces(eng_nx.sweep(eng_nx.upload_permutations(perms)))
    b = eng_cf.host_traces(eng_cf.sweep(eng_cf.upload_permutations(perms)))
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_random_graph_no_spanning(gpu):
    import networkx as nx
    gnx = nx.gnm_random_graph(300, 900, seed=11)
    eng, g = _engine_for(gnx, False)
    perms = oracle.randomstate_permutations(range(70), g['num_edges'])
    expect = oracle.bond_runs(g, perms)
    res = eng.sweep(eng.upload_permutations(perms))
    _check_traces(eng, res, expect)
    _check_sums(eng, res, expect)


@pytest.mark.parametrize('nodes,edges', [(6000, 18000), (20000, 30000)])
def test_random_graph_two_wavefronts(gpu, nodes, edges):
    """Random graphs past the two-wavefront threshold (4096 nodes), with the
    state in LDS and in HBM: traces and sums equal the oracle's."""
    import networkx as nx
    gnx = nx.gnm_random_graph(nodes, edges, seed=nodes)
    eng, g = _engine_for(gnx, False)
    perms = oracle.randomstate_permutations(range(24), g['num_edges'])
    expect = oracle.bond_runs(g, perms)
    res = eng.sweep(eng.upload_permutations(perms))
    _check_traces(eng, res, expect)
    _check_sums(eng, res, expect)


def test_L128_64_runs(gpu):
    """The L = 128 lattice (N = 16384, M = 32512), LDS state at its limit."""
    eng = BondEngine(_graph.square_lattice(128))
    g = oracle.percolation_graph_arrays(spanning_2d_grid(128), True)
    perms = oracle.randomstate_permutations(range(64), g['num_edges'])
    expect = oracle.bond_runs(g, perms)
    res = eng.sweep(eng.upload_permutations(perms))
    _check_sums(eng, res, expect)
    _check_traces(eng, res, expect)


@pytest.mark.parametrize('length', [130, 257])
def test_global_state(gpu, length):
    """N > 16384 keeps the union-find in HBM; N > 65536 uses 64-bit records."""
    eng = BondEngine(_graph.square_lattice(length))
    g = oracle.percolation_graph_arrays(spanning_2d_grid(length), True)
    perms = oracle.randomstate_permutations(range(6), g['num_edges'])
    expect = oracle.bond_runs(g, perms)
    res = eng.sweep(eng.upload_permutations(perms))
    _check_traces(eng, res, expect)
    _check_sums(eng, res, expect)


@pytest.mark.parametrize('M', [1, 2, 12, 33, 64, 65, 480, 4096, 4097, 32512,
                               32768, 32769, 65536, 65537, 130560, 2095104])
def test_device_permutation_matches_oracle(gpu, M):
    """pc_permute_philox equals the oracle's restatement of the bucket
    shuffle bit for bit: the whole run as one bucket in the 2-wave (M <=
    4096) and the 8-wave group kernel (up to 32768: the L = 128 lattice),
    one sub-bucket or several, then several buckets through count, scan and
    scatter (from 32769 edges; several chunks of those passes at the
    L = 256 and L = 1024 lattices; the count and scatter blocks of 11 runs
    span two groups of eight runs, the second padded)."""
    eng = BondEngine(_graph.GraphArrays(M + 1, np.stack(
        [np.arange(M), np.arange(1, M + 1)], axis=1)))
    runs = 2 if M >= 1000000 else (11 if M > 32768 else 3)
    d = eng.device_permutations(seed=0x1234567890AB, num_runs=runs, first_run=5)
    got = d.cpu().numpy()
    want = oracle.bucket_permutations(0x1234567890AB, 5, runs, M)
    for r in range(runs):
        np.testing.assert_array_equal(got[r], want[r], err_msg='run {}'.format(r))


def test_device_permutation_L128_is_permutation(gpu):
    eng = BondEngine(_graph.square_lattice(128))
    d = eng.device_permutations(seed=1234, num_runs=32)
    srt = d.sort(dim=1).values.cpu().numpy()
    np.testing.assert_array_equal(srt, np.broadcast_to(np.arange(eng.M), srt.shape))


def test_full_size_properties(gpu):
    """L = 128, 10^4 runs from device permutations: size-independent checks
    of the exact sums (end state, identities between columns)."""
    eng = BondEngine(_graph.square_lattice(128))
    R = 10000
    res = eng.sweep(eng.device_permutations(seed=1234, num_runs=R))
    s1, s2 = host_sums(*eng.microcanonical_sums(res))
    N, M = eng.N, eng.M
    # step 0 and step M
    assert s1[0, 0] == R and s1[0, M] == R * N
    assert s1[1, 0] == 0 and s1[1, M] == R
    assert all(s1[2 + k, M] == 0 for k in range(5))
    assert all(s1[2 + k, 0] == R * (N - 1) for k in range(5))
    # moments[1] = N - max for every run and step
    assert (s1[3] == R * N - s1[0]).all()
    # spanning is 0/1: sum of squares equals sum
    assert (s2[1] == s1[1]).all()
    # max cluster size and spanning count never decrease
    assert (np.diff(s1[0].astype(float)) >= 0).all()
    assert (np.diff(s1[1].astype(float)) >= 0).all()
    # Cauchy-Schwarz: R * sum x^2 >= (sum x)^2
    for c in range(7):
        assert (R * s2[c] - s1[c] * s1[c] >= 0).all()


def test_pipeline_matches_single_launch(gpu):
    """Sub-batches on three streams (reduction beside the next sweep) give
    the same exact sums as one launch over all runs."""
    from pypercolate_amd._engine import Pipeline
    eng = BondEngine(_graph.square_lattice(32))
    R = 3000
    res = eng.sweep(eng.device_permutations(9, R))
    ref = [t.cpu().numpy() for t in eng.microcanonical_sums(res)]
    pipe = Pipeline(eng, R, seed=9, sub_batch=700)
    for _ in range(2):  # a second run must restart the accumulation
        s1, s2 = pipe.run()
        gpu.cuda.synchronize()
        assert np.array_equal(s1.cpu().numpy(), ref[0])
        assert np.array_equal(s2.cpu().numpy(), ref[1])


@pytest.mark.parametrize('L,after', [(32, 1), (32, 0), (32, 9), (130, 1)])
def test_pipeline_prefetch_matches_single_launch(gpu, L, after):
    """Permutations of the next run() drawn during this one (prefetch):
    every run() gives the one-launch exact sums, also with other work on
    the caller's stream between runs and for the HBM-state sweep."""
    from pypercolate_amd._engine import Pipeline
    eng = BondEngine(_graph.square_lattice(L))
    R = 3000 if L == 32 else 300
    res = eng.sweep(eng.device_permutations(9, R))
    ref = [t.cpu().numpy() for t in eng.microcanonical_sums(res)]
    pipe = Pipeline(eng, R, seed=9, sub_batch=R // 4 + 1, prefetch=True,
                    prefetch_after=after)
    assert pipe.prefetch
    for k in range(4):
        s1, s2 = pipe.run()
        junk = gpu.ones(1 << 20, device='cuda').sum().item()
        assert junk == float(1 << 20)
        gpu.cuda.synchronize()
        assert np.array_equal(s1.cpu().numpy(), ref[0]), k
        assert np.array_equal(s2.cpu().numpy(), ref[1]), k
        assert np.array_equal(pipe.perm.cpu().numpy(), res.perm.cpu().numpy()), k


@pytest.mark.parametrize('L,relabel', [(1, False), (2, False), (3, False),
                                       (8, False), (16, False), (37, False),
                                       (128, False), (8, True), (136, True),
                                       (264, True), (130, False)])
def test_lattice_edges_equal_edge_list(gpu, L, relabel):
    """pc_sweep_lattice (endpoints computed from the edge index, row-major
    or tiled node order) gives pc_sweep's records, span steps and chunk
    states over the closed-form lattice's edge array, in LDS and HBM state
    (L = 136, 264 tiled; L = 130 row-major in HBM)."""
    ga = _graph.square_lattice(L)
    eng = BondEngine(ga, relabel=relabel)
    assert eng.relabelled == relabel
    assert eng.lattice_length == L and eng.lattice_tiled == relabel
    assert eng.lattice_sweep == (L * L > 16384)
    runs = 6 if L < 200 else 3
    perm = eng.device_permutations(31, runs)
    eng.lattice_sweep = True
    lat = eng.sweep(perm, chunks=True)
    eng.lattice_sweep = False
    ref = eng.sweep(perm, chunks=True)
    M = eng.M   # L = 1 has no edges: nothing is written
    np.testing.assert_array_equal(lat.records.cpu().numpy()[:, :M],
                                  ref.records.cpu().numpy()[:, :M])
    np.testing.assert_array_equal(lat.span_step.cpu().numpy(), ref.span_step.cpu().numpy())
    if M:
        np.testing.assert_array_equal(lat.chunk_state.cpu().numpy(),
                                      ref.chunk_state.cpu().numpy())


def test_lattice_sweep_rejects_bad_sizes(gpu):
    from pypercolate_amd import _native as nat
    lib = nat.load_library()
    one = gpu.zeros(8, dtype=gpu.int32, device='cuda')
    args = (None, 0, nat.ptr(one), 1, 1, nat.ptr(one), 1, nat.ptr(one),
            None, 0, 0, 0, None, None)
    assert lib.pc_sweep_lattice(0, 0, *args) == -1
    assert lib.pc_sweep_lattice(12, 1, *args) == -1       # tiles need L % 8 == 0
    assert lib.pc_sweep_lattice(20000, 0, *args) == -1


@pytest.mark.parametrize('ahead,prefetch', [(True, False), (False, False),
                                            (True, True)])
def test_pipeline_bit_exact_mode_matches_oracle(gpu, ahead, prefetch):
    """The pipeline with device RandomState permutations (seeds = run
    index, the L = 16 configuration) gives the oracle's exact sums over
    numpy's permutations, drawn all before the sweeps, per sub-batch, or
    ahead for the next run() (three run() calls, the later ones on
    prefetched permutations)."""
    from pypercolate_amd._engine import Pipeline
    eng = BondEngine(_graph.square_lattice(16))
    g = oracle.percolation_graph_arrays(spanning_2d_grid(16), True)
    tr = oracle.bond_runs(g, oracle.randomstate_permutations(range(40), g['num_edges']))
    ref1, ref2 = oracle.microcanonical_sums(tr)
    pipe = Pipeline(eng, 40, seeds=np.arange(40), sub_batch=16, permute_ahead=ahead,
                    prefetch=prefetch, prefetch_after=1)
    assert pipe.prefetch == prefetch
    for _ in range(3 if prefetch else 1):
        s1, s2 = host_sums(*pipe.run())
        assert (s1 == ref1).all() and (s2 == ref2).all()


@pytest.mark.parametrize('M', [1, 2, 3, 12, 63, 64, 65, 129, 480, 4097, 32512,
                               32768, 32769, 65536, 70000])
def test_randomstate_permutations_bit_exact(gpu, M):
    """Device MT19937 + legacy shuffle == numpy RandomState(seed).permutation:
    up to 32768 edges the swap targets drawn first (64-word blocks settled
    at once) and the swaps applied all at once (per-position step lists,
    pointer jumping); above, the shuffle walked in sequence (LDS array up
    to 65536 edges, the output row above)."""
    eng = BondEngine(_graph.GraphArrays(M + 1, np.stack(
        [np.arange(M), np.arange(1, M + 1)], axis=1)))
    seeds = [0, 1, 42, 12345, 4294967295, 2 ** 31 + 7]
    got = eng.randomstate_permutations(seeds).cpu().numpy()
    for r, seed in enumerate(seeds):
        np.testing.assert_array_equal(
            got[r], np.random.RandomState(seed).permutation(M), err_msg=str(seed))


def test_randomstate_kat_3x3(gpu, golden):
    """The reference's golden permutation (test_percolate.py:141-142)."""
    kat = golden('kat_3x3.json')
    eng = BondEngine(_graph.square_lattice(3))
    got = eng.randomstate_permutations([kat['seed']]).cpu().numpy()[0]
    assert got.tolist() == kat['permutation']


@pytest.mark.parametrize('length,mode', [(16, 'auto'), (128, 'auto'), (64, 'hbm'),
                                         (130, 'auto'), (257, 'auto')])
def test_sweep_chunk_states_equal_chunk_scan(gpu, length, mode):
    """The chunk states the sweep writes itself are bit-identical to the
    separate pass over the records (pc_chunk_scan): LDS state, HBM state
    with 32-bit records, HBM state with 64-bit records."""
    from pypercolate_amd import _native as nat
    eng = BondEngine(_graph.square_lattice(length))
    if mode == 'hbm':
        eng.state_mode = nat.PC_STATE_HBM
    res = eng.sweep(eng.device_permutations(77, 96), chunks=True)
    fused = res.chunk_state.cpu().numpy()
    scanned = eng.chunk_scan(res).cpu().numpy()
    assert np.array_equal(fused, scanned)


@pytest.mark.parametrize('length', [130, 257])
def test_relabelled_state_gives_same_records(gpu, length):
    """HBM-state graphs store the union-find codes in 4 x 8 tiles of the
    grid (BondEngine relabel): records, span steps and chunk states equal
    those of the graph's own labels."""
    g = _graph.square_lattice(length)
    a, b = BondEngine(g, relabel=True), BondEngine(g, relabel=False)
    assert a.relabelled and not b.relabelled
    perm = a.device_permutations(5, 40)
    ra, rb = a.sweep(perm, chunks=True), b.sweep(perm, chunks=True)
    assert np.array_equal(ra.records.cpu().numpy(), rb.records.cpu().numpy())
    assert np.array_equal(ra.span_step.cpu().numpy(), rb.span_step.cpu().numpy())
    assert np.array_equal(ra.chunk_state.cpu().numpy(), rb.chunk_state.cpu().numpy())


@pytest.mark.parametrize('kind,length', [('grid', 2), ('grid', 5), ('grid', 9),
                                         ('grid', 33), ('chain', 2), ('chain', 3),
                                         ('chain', 200)])
@pytest.mark.parametrize('placement', ['lds', 'hbm'])
def test_two_wavefronts_small_graphs(gpu, kind, length, placement):
    """The two-wavefront sweep forced onto small graphs, whose only batch
    or last batch is partial: traces and sums equal the oracle's, and the
    chunk states it writes equal the separate pass."""
    from pypercolate_amd import _native as nat
    from tests.helpers import spanning_1d_chain
    if kind == 'grid':
        ga, nxg = _graph.square_lattice(length), spanning_2d_grid(length)
    else:
        ga, nxg = _graph.linear_chain(length), spanning_1d_chain(length)
    g = oracle.percolation_graph_arrays(nxg, True)
    eng = BondEngine(ga)
    eng.state_mode = nat.PC_SWEEP_TWO_WAVES | (
        nat.PC_STATE_LDS if placement == 'lds' else nat.PC_STATE_HBM)
    perms = oracle.randomstate_permutations(range(12), g['num_edges'])
    expect = oracle.bond_runs(g, perms)
    res = eng.sweep(eng.upload_permutations(perms), chunks=True)
    _check_traces(eng, res, expect)
    _check_sums(eng, res, expect)
    assert np.array_equal(res.chunk_state.cpu().numpy(),
                          eng.chunk_scan(res).cpu().numpy())


@pytest.mark.parametrize('kind', ['star', 'complete', 'wheel'])
def test_hbm_state_hub_graphs(gpu, kind):
    """HBM state forced onto graphs whose merges keep hitting one root (a
    star with extra rim edges, a complete graph, a wheel): the root the
    finder hands over is changed by the resolver in nearly every batch, so
    the resolver's stamp test and re-reads, the star merges and the folder
    wave are all exercised.  Traces and sums equal the oracle's, and the
    chunk states equal the separate pass."""
    import networkx as nx
    from pypercolate_amd import _native as nat
    if kind == 'star':
        gnx = nx.star_graph(3000)
        gnx.add_edges_from((i, i + 1) for i in range(1, 3000, 7))
    elif kind == 'complete':
        gnx = nx.complete_graph(90)
    else:
        gnx = nx.wheel_graph(2500)
    eng, g = _engine_for(gnx, False)
    eng.state_mode = nat.PC_STATE_HBM
    perms = oracle.randomstate_permutations(range(10), g['num_edges'])
    expect = oracle.bond_runs(g, perms)
    res = eng.sweep(eng.upload_permutations(perms), chunks=True)
    _check_traces(eng, res, expect)
    _check_sums(eng, res, expect)
    assert np.array_equal(res.chunk_state.cpu().numpy(),
                          eng.chunk_scan(res).cpu().numpy())


@pytest.mark.parametrize('chunks', [True, False])
def test_hbm_sweep_with_fewer_slots_than_runs(gpu, chunks):
    """A union-find workspace of 7 run slots for 40 runs: each workgroup
    sweeps several runs in turn, clearing its slot, claim and stamp tables
    between them.  Records, span steps and chunk states equal those of a
    workspace with a slot per run."""
    eng = BondEngine(_graph.square_lattice(130))
    assert eng.N > 16384
    perm = eng.device_permutations(11, 40)
    full = eng.sweep(perm, chunks=chunks)
    ws = gpu.empty((7 * eng.N * 4,), dtype=gpu.uint8, device='cuda')
    part = eng.sweep(perm, workspace=ws, chunks=chunks)
    assert np.array_equal(part.records.cpu().numpy(), full.records.cpu().numpy())
    assert np.array_equal(part.span_step.cpu().numpy(), full.span_step.cpu().numpy())
    if chunks:
        assert np.array_equal(part.chunk_state.cpu().numpy(),
                              full.chunk_state.cpu().numpy())
