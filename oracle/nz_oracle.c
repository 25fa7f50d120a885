/*
 * TEST INFRASTRUCTURE ONLY.  CPU oracle for the Newman-Ziff bond-percolation
 * sweep.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load this library, and only as the checker; the product path
 * (pypercolate_amd) never links, loads or calls it.
 *
 * It restates, step for step, the reference's single-run sweep
 *   percolate/hpc.py:194-307   (bond_sample_states: uint64 moments)
 * including its bookkeeping of the largest cluster by root
 * (hpc.py:245-246, 282-305) and its second union-find over the auxiliary
 * spanning nodes (hpc.py:223-243, 269-274).  The union-find restates
 * networkx.utils.union_find.UnionFind (a third-party dependency of the
 * reference, unpinned in requirements.txt, networkx 1.10 for the published
 * numbers): full path compression on lookup, union by weight, ties broken
 * towards the larger key as in nx 1.x's `max((weight, root))`.
 *
 * Parity status: pinned by the reference's own 3x3 known-answer test
 * (test_percolate.py:132-192), the two tutorial traces
 * (docs/tutorial-bond-square-lattice.ipynb), brute-force connected
 * components on small lattices (the strategy of test_hpc.py:250-292) and the
 * closed-form end states; see tests/test_oracle_sweep.py.
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -shared).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* networkx UnionFind.__getitem__: walk to the root, then point every node
 * on the path at the root. */
static int64_t uf_find(int64_t *parent, int64_t x)
{
    int64_t root = x;
    while (parent[root] != root)
        root = parent[root];
    while (parent[x] != root) {
        int64_t next = parent[x];
        parent[x] = root;
        x = next;
    }
    return root;
}

/* networkx UnionFind.union(*objects): find every root, pick the heaviest
 * (nx 1.x: max over (weight, root) pairs) and hang all other roots below it. */
static void uf_union(int64_t *parent, uint64_t *weight, const int64_t *objs,
                     int count)
{
    int64_t roots[3];
    int64_t heaviest = -1;
    for (int i = 0; i < count; ++i) {
        roots[i] = uf_find(parent, objs[i]);
        if (heaviest < 0 || weight[roots[i]] > weight[heaviest] ||
            (weight[roots[i]] == weight[heaviest] && roots[i] > heaviest))
            heaviest = roots[i];
    }
    for (int i = 0; i < count; ++i) {
        if (roots[i] != heaviest) {
            /* a root may appear twice among objs; only hang it once */
            if (parent[roots[i]] == heaviest)
                continue;
            weight[heaviest] += weight[roots[i]];
            parent[roots[i]] = heaviest;
        }
    }
}

static void powers_u64(uint64_t base, uint64_t out[5])
{
    /* numpy: `weight ** np.arange(5, dtype='uint64')`, wrapping mod 2**64 */
    uint64_t p = 1;
    for (int k = 0; k < 5; ++k) {
        out[k] = p;
        p *= base;
    }
}

/*
 * One run.  Perc nodes are 0..num_nodes-1 in perc_graph.nodes() order, so
 * node 0 is the reference's initial `max_cluster_root` (hpc.py:246).
 * Auxiliary nodes (spanning detection) are num_nodes..num_nodes+num_aux-1;
 * aux_side[i] is 0 or 1, the position of the node's 'span' value in
 * `spanning_sides`.  aux_edges are pairs in the combined index space with
 * aux_edge_side their side position.  perm[n] is the edge index drawn at
 * step n+1 (hpc.py:206, 254).
 *
 * Outputs have num_edges+1 rows; row 0 is the n = 0 state (hpc.py:213-218);
 * out_edge[0] is left 0 (undefined in the reference, hpc.py:362-363).
 * Returns 0, or -1 if an auxiliary side has no nodes.
 */
int nzo_bond_run(int64_t num_nodes, int64_t num_edges, const int64_t *edges,
                 int spanning, int64_t num_aux, const int32_t *aux_side,
                 int64_t num_aux_edges, const int64_t *aux_edges,
                 const int32_t *aux_edge_side, const int64_t *perm,
                 uint32_t *out_edge, uint8_t *out_span, uint32_t *out_max,
                 uint64_t *out_moments)
{
    const int64_t total = num_nodes + (spanning ? num_aux : 0);
    int64_t *parent = (int64_t *)malloc(sizeof(int64_t) * (size_t)(total > 0 ? total : 1));
    uint64_t *weight = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(total > 0 ? total : 1));
    int64_t *sparent = NULL;
    uint64_t *sweight = NULL;
    int64_t side_roots[2] = {0, 0};
    int rc = 0;

    uint32_t max_size = 1;
    uint64_t moments[5];
    uint8_t span = 0;
    for (int k = 0; k < 5; ++k)
        moments[k] = (uint64_t)(num_nodes - 1);

    out_edge[0] = 0;
    out_span[0] = 0;
    out_max[0] = max_size;
    memcpy(out_moments, moments, sizeof(moments));

    for (int64_t i = 0; i < num_nodes; ++i) {
        parent[i] = i;
        weight[i] = 1;
    }

    if (spanning) {
        sparent = (int64_t *)malloc(sizeof(int64_t) * (size_t)total);
        sweight = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)total);
        for (int64_t i = 0; i < total; ++i) {
            sparent[i] = i;
            sweight[i] = 1;
        }
        /* hpc.py:229-236: merge all auxiliary nodes of each side */
        for (int side = 0; side < 2; ++side) {
            int64_t first = -1;
            for (int64_t a = 0; a < num_aux; ++a) {
                if (aux_side[a] != side)
                    continue;
                if (first < 0) {
                    first = num_nodes + a;
                } else {
                    int64_t pair[2] = {first, num_nodes + a};
                    uf_union(sparent, sweight, pair, 2);
                }
            }
            if (first < 0) {
                rc = -1;
                goto done;
            }
            side_roots[side] = uf_find(sparent, first);
        }
        /* hpc.py:238-239: union(side_root, *edge) for every auxiliary edge */
        for (int64_t e = 0; e < num_aux_edges; ++e) {
            int64_t objs[3] = {side_roots[aux_edge_side[e]], aux_edges[2 * e],
                               aux_edges[2 * e + 1]};
            uf_union(sparent, sweight, objs, 3);
        }
        /* hpc.py:241-243 */
        side_roots[0] = uf_find(sparent, side_roots[0]);
        side_roots[1] = uf_find(sparent, side_roots[1]);
    }

    int64_t max_cluster_root = 0; /* hpc.py:246: first perc node */

    for (int64_t n = 0; n < num_edges; ++n) {
        const int64_t edge_index = perm[n];
        const int64_t u = edges[2 * edge_index], v = edges[2 * edge_index + 1];
        int64_t roots[2] = {uf_find(parent, u), uf_find(parent, v)};
        uint64_t weights[2] = {weight[roots[0]], weight[roots[1]]};

        if (roots[0] != roots[1]) {
            uf_union(parent, weight, roots, 2);
            if (spanning) {
                uf_union(sparent, sweight, roots, 2);
                span = uf_find(sparent, side_roots[0]) ==
                       uf_find(sparent, side_roots[1]);
            }
            const int64_t root = uf_find(parent, u);
            const uint64_t w = weight[root];
            uint64_t pw[5];

            /* hpc.py:283-286 */
            for (int i = 0; i < 2; ++i) {
                if (roots[i] == max_cluster_root)
                    continue;
                powers_u64(weights[i], pw);
                for (int k = 0; k < 5; ++k)
                    moments[k] -= pw[k];
            }
            if (roots[0] == max_cluster_root || roots[1] == max_cluster_root) {
                /* hpc.py:288-291: merged with the largest cluster */
                max_cluster_root = root;
                max_size = (uint32_t)w;
            } else if ((uint64_t)max_size >= w) {
                /* hpc.py:294-297 */
                powers_u64(w, pw);
                for (int k = 0; k < 5; ++k)
                    moments[k] += pw[k];
            } else {
                /* hpc.py:298-305: the merged cluster overtakes */
                max_cluster_root = root;
                powers_u64(max_size, pw);
                for (int k = 0; k < 5; ++k)
                    moments[k] += pw[k];
                max_size = (uint32_t)w;
            }
        }
        out_edge[n + 1] = (uint32_t)edge_index;
        out_span[n + 1] = span;
        out_max[n + 1] = max_size;
        memcpy(out_moments + 5 * (n + 1), moments, sizeof(moments));
    }

done:
    free(parent);
    free(weight);
    free(sparent);
    free(sweight);
    return rc;
}

/* Many runs, one per OpenMP thread: perms is [runs][num_edges], outputs are
 * [runs][num_edges+1] (moments [runs][num_edges+1][5]). */
int nzo_bond_runs(int64_t runs, int64_t num_nodes, int64_t num_edges,
                  const int64_t *edges, int spanning, int64_t num_aux,
                  const int32_t *aux_side, int64_t num_aux_edges,
                  const int64_t *aux_edges, const int32_t *aux_edge_side,
                  const int64_t *perms, uint32_t *out_edge, uint8_t *out_span,
                  uint32_t *out_max, uint64_t *out_moments, int threads)
{
    int rc = 0;
    const int64_t S = num_edges + 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads) reduction(| : rc)
    for (int64_t r = 0; r < runs; ++r) {
        rc |= nzo_bond_run(num_nodes, num_edges, edges, spanning, num_aux,
                           aux_side, num_aux_edges, aux_edges, aux_edge_side,
                           perms + r * num_edges, out_edge + r * S,
                           out_span + r * S, out_max + r * S,
                           out_moments + 5 * r * S);
    }
    return rc;
}

/*
 * Per-step sums over runs of the seven integer observables and of their
 * squares, exactly: column order max_cluster_size, has_spanning_cluster,
 * moments[0..4].  sum1 is [7][S][2] (u128 as two little-endian u64 limbs),
 * sum2 is [7][S][3] (u192).  Inputs as produced by nzo_bond_runs.
 */
void nzo_microcanonical_sums(int64_t runs, int64_t S, const uint8_t *span,
                             const uint32_t *max_size, const uint64_t *moments,
                             uint64_t *sum1, uint64_t *sum2)
{
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < S; ++s) {
        u128 acc1[7];
        u128 acc2lo[7];
        uint64_t acc2hi[7];
        for (int c = 0; c < 7; ++c) {
            acc1[c] = 0;
            acc2lo[c] = 0;
            acc2hi[c] = 0;
        }
        for (int64_t r = 0; r < runs; ++r) {
            uint64_t x[7];
            x[0] = max_size[r * S + s];
            x[1] = span[r * S + s];
            for (int k = 0; k < 5; ++k)
                x[2 + k] = moments[(r * S + s) * 5 + k];
            for (int c = 0; c < 7; ++c) {
                acc1[c] += x[c];
                u128 sq = (u128)x[c] * x[c];
                u128 before = acc2lo[c];
                acc2lo[c] += sq;
                if (acc2lo[c] < before)
                    acc2hi[c] += 1;
            }
        }
        for (int c = 0; c < 7; ++c) {
            sum1[(c * S + s) * 2 + 0] = (uint64_t)acc1[c];
            sum1[(c * S + s) * 2 + 1] = (uint64_t)(acc1[c] >> 64);
            sum2[(c * S + s) * 3 + 0] = (uint64_t)acc2lo[c];
            sum2[(c * S + s) * 3 + 1] = (uint64_t)(acc2lo[c] >> 64);
            sum2[(c * S + s) * 3 + 2] = acc2hi[c];
        }
    }
}
