/*
 * TEST INFRASTRUCTURE ONLY.  CPU oracle for the Newman-Ziff bond-percolation
 * sweep.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load this library, and only as the checker; the product path
 * (pypercolate_amd) never links, loads or calls it.
 *
 * It restates, step for step, the reference's single-run sweep
 *   percolate/hpc.py:194-307   (bond_sample_states: uint64 moments, carried
 *                              here exactly in 128 bits)
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

static void powers_u128(uint64_t base, u128 out[5])
{
    /* `weight ** np.arange(5)` exactly; the reference's uint64 numpy
     * arithmetic (hpc.py:286,297,302) is this reduced mod 2**64 */
    u128 p = 1;
    for (int k = 0; k < 5; ++k) {
        out[k] = p;
        p *= base;
    }
}

/*
 * Where a run's per-step state goes: per-step trace rows (edge, spanning
 * flag, largest cluster, moments as low and optionally high 64-bit words),
 * or exact per-step sums over runs (see nzo_bond_sums).
 */
struct step_sink {
    uint32_t *edge;
    uint8_t *span;
    uint32_t *max;
    uint64_t *mom_lo;
    uint64_t *mom_hi;   /* NULL: high words not wanted */
    uint64_t *sum1;     /* non-NULL: accumulate [7][S][2] sums instead */
    uint64_t *sum2;     /* [7][S][3] */
    int64_t S;
    int overflow;
};

/* acc[0..limbs) += v (little-endian 64-bit limbs); sets *ovf on carry out */
static void add_limbs(uint64_t *acc, const uint64_t *v, int limbs, int *ovf)
{
    unsigned carry = 0;
    for (int j = 0; j < limbs; ++j) {
        const uint64_t a = acc[j];
        const uint64_t s = a + v[j];
        const unsigned c1 = s < a;
        const uint64_t t = s + carry;
        const unsigned c2 = t < s;
        acc[j] = t;
        carry = c1 | c2;
    }
    if (carry)
        *ovf = 1;
}

/* x**2 for x < 2**128 as four 64-bit limbs */
static void square_u128(u128 x, uint64_t out[4])
{
    const uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
    const u128 ll = (u128)lo * lo, lh = (u128)lo * hi, hh = (u128)hi * hi;
    /* x^2 = ll + 2 lh 2^64 + hh 2^128 */
    uint64_t r[4] = {(uint64_t)ll, (uint64_t)(ll >> 64), (uint64_t)hh,
                     (uint64_t)(hh >> 64)};
    const uint64_t m[4] = {0, (uint64_t)lh, (uint64_t)(lh >> 64), 0};
    int ovf = 0;
    add_limbs(r, m, 4, &ovf);
    add_limbs(r, m, 4, &ovf);
    memcpy(out, r, sizeof r);
}

static void sink_step(struct step_sink *k, int64_t s, uint32_t edge,
                      uint8_t span, uint32_t max_size, const u128 mom[5])
{
    if (k->sum1) {
        u128 x[7];
        x[0] = max_size;
        x[1] = span;
        for (int j = 0; j < 5; ++j)
            x[2 + j] = mom[j];
        for (int c = 0; c < 7; ++c) {
            const uint64_t v1[2] = {(uint64_t)x[c], (uint64_t)(x[c] >> 64)};
            uint64_t sq[4];
            add_limbs(k->sum1 + (c * k->S + s) * 2, v1, 2, &k->overflow);
            square_u128(x[c], sq);
            if (sq[3])
                k->overflow = 1;
            add_limbs(k->sum2 + (c * k->S + s) * 3, sq, 3, &k->overflow);
        }
        return;
    }
    k->edge[s] = edge;
    k->span[s] = span;
    k->max[s] = max_size;
    for (int j = 0; j < 5; ++j) {
        k->mom_lo[5 * s + j] = (uint64_t)mom[j];
        if (k->mom_hi)
            k->mom_hi[5 * s + j] = (uint64_t)(mom[j] >> 64);
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
 * The moments are carried exactly in 128 bits (every graph with fewer than
 * 2**32 nodes: sum of size**4 < 2**128).  Their low 64 bits are the
 * reference's uint64 values, which wrap mod 2**64 (numpy uint64 arithmetic,
 * hpc.py:286,297,302): reduction mod 2**64 commutes with the additions,
 * subtractions and products of the bookkeeping.
 *
 * Per-step rows: num_edges+1 of them; row 0 is the n = 0 state
 * (hpc.py:213-218); its edge is 0 (undefined in the reference,
 * hpc.py:362-363).  Returns 0, or -1 if an auxiliary side has no nodes.
 */
static int bond_run(int64_t num_nodes, int64_t num_edges, const int64_t *edges,
                    int spanning, int64_t num_aux, const int32_t *aux_side,
                    int64_t num_aux_edges, const int64_t *aux_edges,
                    const int32_t *aux_edge_side, const int64_t *perm64,
                    const int32_t *perm32, struct step_sink *sink)
{
    const int64_t total = num_nodes + (spanning ? num_aux : 0);
    int64_t *parent = (int64_t *)malloc(sizeof(int64_t) * (size_t)(total > 0 ? total : 1));
    uint64_t *weight = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(total > 0 ? total : 1));
    int64_t *sparent = NULL;
    uint64_t *sweight = NULL;
    int64_t side_roots[2] = {0, 0};
    int rc = 0;

    uint32_t max_size = 1;
    u128 moments[5];
    uint8_t span = 0;
    for (int k = 0; k < 5; ++k)
        moments[k] = (u128)(num_nodes - 1);

    sink_step(sink, 0, 0, span, max_size, moments);

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
        const int64_t edge_index = perm64 ? perm64[n] : (int64_t)perm32[n];
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
            u128 pw[5];

            /* hpc.py:283-286 */
            for (int i = 0; i < 2; ++i) {
                if (roots[i] == max_cluster_root)
                    continue;
                powers_u128(weights[i], pw);
                for (int k = 0; k < 5; ++k)
                    moments[k] -= pw[k];
            }
            if (roots[0] == max_cluster_root || roots[1] == max_cluster_root) {
                /* hpc.py:288-291: merged with the largest cluster */
                max_cluster_root = root;
                max_size = (uint32_t)w;
            } else if ((uint64_t)max_size >= w) {
                /* hpc.py:294-297 */
                powers_u128(w, pw);
                for (int k = 0; k < 5; ++k)
                    moments[k] += pw[k];
            } else {
                /* hpc.py:298-305: the merged cluster overtakes */
                max_cluster_root = root;
                powers_u128(max_size, pw);
                for (int k = 0; k < 5; ++k)
                    moments[k] += pw[k];
                max_size = (uint32_t)w;
            }
        }
        sink_step(sink, n + 1, (uint32_t)edge_index, span, max_size, moments);
    }

done:
    free(parent);
    free(weight);
    free(sparent);
    free(sweight);
    return rc;
}

/* One run into trace rows (see bond_run); out_moments_hi may be NULL. */
int nzo_bond_run(int64_t num_nodes, int64_t num_edges, const int64_t *edges,
                 int spanning, int64_t num_aux, const int32_t *aux_side,
                 int64_t num_aux_edges, const int64_t *aux_edges,
                 const int32_t *aux_edge_side, const int64_t *perm,
                 uint32_t *out_edge, uint8_t *out_span, uint32_t *out_max,
                 uint64_t *out_moments, uint64_t *out_moments_hi)
{
    struct step_sink k = {out_edge, out_span, out_max, out_moments,
                          out_moments_hi, NULL, NULL, num_edges + 1, 0};
    return bond_run(num_nodes, num_edges, edges, spanning, num_aux, aux_side,
                    num_aux_edges, aux_edges, aux_edge_side, perm, NULL, &k);
}

/* Many runs, one per OpenMP thread: perms is [runs][num_edges], outputs are
 * [runs][num_edges+1] (moments [runs][num_edges+1][5], low words; the high
 * words of the exact values go to out_moments_hi unless it is NULL). */
int nzo_bond_runs(int64_t runs, int64_t num_nodes, int64_t num_edges,
                  const int64_t *edges, int spanning, int64_t num_aux,
                  const int32_t *aux_side, int64_t num_aux_edges,
                  const int64_t *aux_edges, const int32_t *aux_edge_side,
                  const int64_t *perms, uint32_t *out_edge, uint8_t *out_span,
                  uint32_t *out_max, uint64_t *out_moments,
                  uint64_t *out_moments_hi, int threads)
{
    int rc = 0;
    const int64_t S = num_edges + 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads) reduction(| : rc)
    for (int64_t r = 0; r < runs; ++r) {
        rc |= nzo_bond_run(num_nodes, num_edges, edges, spanning, num_aux,
                           aux_side, num_aux_edges, aux_edges, aux_edge_side,
                           perms + r * num_edges, out_edge + r * S,
                           out_span + r * S, out_max + r * S,
                           out_moments + 5 * r * S,
                           out_moments_hi ? out_moments_hi + 5 * r * S : NULL);
    }
    return rc;
}

/*
 * Per-step sums over runs of the seven integer observables and of their
 * squares, exactly: column order max_cluster_size, has_spanning_cluster,
 * moments[0..4] (moments = moments_lo + 2**64 moments_hi; moments_hi may be
 * NULL: zero).  sum1 is [7][S][2] (u128 as two little-endian u64 limbs),
 * sum2 is [7][S][3] (u192).  Inputs as produced by nzo_bond_runs.  Returns
 * 0, or -1 if a sum does not fit its limbs.
 */
int nzo_microcanonical_sums(int64_t runs, int64_t S, const uint8_t *span,
                            const uint32_t *max_size, const uint64_t *moments,
                            const uint64_t *moments_hi, uint64_t *sum1,
                            uint64_t *sum2)
{
    int ovf_any = 0;
    memset(sum1, 0, sizeof(uint64_t) * 2 * 7 * (size_t)S);
    memset(sum2, 0, sizeof(uint64_t) * 3 * 7 * (size_t)S);
#pragma omp parallel for schedule(static) reduction(| : ovf_any)
    for (int64_t s = 0; s < S; ++s) {
        struct step_sink k = {NULL, NULL, NULL, NULL, NULL, sum1, sum2, S, 0};
        for (int64_t r = 0; r < runs; ++r) {
            u128 mom[5];
            for (int j = 0; j < 5; ++j) {
                const int64_t at = (r * S + s) * 5 + j;
                mom[j] = (u128)moments[at] |
                         (moments_hi ? (u128)moments_hi[at] << 64 : 0);
            }
            sink_step(&k, s, 0, span[r * S + s], max_size[r * S + s], mom);
        }
        ovf_any |= k.overflow;
    }
    return ovf_any ? -1 : 0;
}

/*
 * Exact per-step sums over runs straight from the sweep, without traces:
 * the runs of perms ([runs][num_edges] int32) are swept (OpenMP over runs,
 * each thread accumulating into its own [7][S][2] / [7][S][3] buffers) and
 * the sums added to sum1 / sum2 (accumulate != 0) or written (accumulate ==
 * 0).  Returns 0, -1 if an auxiliary side has no nodes, -2 if a sum does
 * not fit its limbs, -3 if memory runs out.
 */
int nzo_bond_sums(int64_t runs, int64_t num_nodes, int64_t num_edges,
                  const int64_t *edges, int spanning, int64_t num_aux,
                  const int32_t *aux_side, int64_t num_aux_edges,
                  const int64_t *aux_edges, const int32_t *aux_edge_side,
                  const int32_t *perms, uint64_t *sum1, uint64_t *sum2,
                  int accumulate, int threads)
{
    const int64_t S = num_edges + 1;
    const size_t n1 = 2 * 7 * (size_t)S, n2 = 3 * 7 * (size_t)S;
    int rc = 0, ovf = 0;
    if (threads < 1)
        threads = 1;
    if (threads > runs)
        threads = runs > 0 ? (int)runs : 1;
    if (!accumulate) {
        memset(sum1, 0, sizeof(uint64_t) * n1);
        memset(sum2, 0, sizeof(uint64_t) * n2);
    }
#pragma omp parallel num_threads(threads) reduction(| : rc, ovf)
    {
        uint64_t *a1 = (uint64_t *)calloc(n1, sizeof(uint64_t));
        uint64_t *a2 = (uint64_t *)calloc(n2, sizeof(uint64_t));
        if (!a1 || !a2) {
            rc |= 4;
        } else {
            struct step_sink k = {NULL, NULL, NULL, NULL, NULL, a1, a2, S, 0};
#pragma omp for schedule(dynamic, 1)
            for (int64_t r = 0; r < runs; ++r)
                if (bond_run(num_nodes, num_edges, edges, spanning, num_aux,
                             aux_side, num_aux_edges, aux_edges,
                             aux_edge_side, NULL, perms + r * num_edges, &k))
                    rc |= 1;
            ovf |= k.overflow;
#pragma omp critical
            {
                int o = 0;
                for (size_t i = 0; i < n1; i += 2)
                    add_limbs(sum1 + i, a1 + i, 2, &o);
                for (size_t i = 0; i < n2; i += 3)
                    add_limbs(sum2 + i, a2 + i, 3, &o);
                ovf |= o;
            }
        }
        free(a1);
        free(a2);
    }
    if (rc & 4)
        return -3;
    if (rc & 1)
        return -1;
    return ovf ? -2 : 0;
}
