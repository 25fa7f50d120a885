/*
 * percolate_hip.h -- C ABI of libpercolate_hip.so, the MI355X (gfx950)
 * implementation of pypercolate's Newman-Ziff bond-percolation hot path.
 *
 * Conventions
 *   - Every entry point returns an int status: PC_OK (0) or a negative
 *     PC_E* code; pc_last_error() then holds a thread-local message.  No
 *     exception or abort crosses the ABI.
 *   - Pointers named d_* are device (HBM) pointers owned by the caller; the
 *     library allocates nothing behind them.  `stream` is a hipStream_t
 *     (NULL = the default stream); work is enqueued asynchronously on it.
 *   - Graph: perc nodes are 0..num_nodes-1, edges are int32 pairs
 *     d_edges[2*e], d_edges[2*e+1].  d_node_sides[i] is 0, 1 (touches the
 *     first spanning side) or 2 (touches the second), or d_node_sides is NULL
 *     when spanning clusters are not detected.  span_on_first_merge = 1 when
 *     the auxiliary structure alone already joins both sides (the reference
 *     then reports a spanning cluster from the first merge on).
 *   - Steps: step s = 0..num_edges is the state after s bonds.  perm[s-1] is
 *     the edge index added at step s.
 *   - Merge records: one per step s >= 1, at index s-1.  0 = the edge joined
 *     two nodes of one cluster; otherwise the sizes (a, b) of the two
 *     clusters it merged, packed a | b << 16 in a uint32 when
 *     num_nodes <= 65536 and a | b << 32 in a uint64 above
 *     (pc_record_bytes()).
 *   - Spanning: span_step[r] is the first step with a spanning cluster, or
 *     0xFFFFFFFF if none.
 */
#ifndef PERCOLATE_HIP_H
#define PERCOLATE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PC_OK 0
#define PC_EINVAL (-1)    /* bad argument (maps to ValueError)              */
#define PC_ENOMEM (-2)    /* workspace too small (maps to MemoryError)      */
#define PC_EHIP (-3)      /* HIP runtime error (maps to RuntimeError)       */
#define PC_EOVERFLOW (-4) /* size outside the supported range (ValueError)  */

#define PC_NO_SPAN 0xFFFFFFFFu
#define PC_CHUNK 64       /* steps per chunk-state entry                    */
#define PC_STATE_WORDS 8  /* uint64 words per chunk-state entry             */
#define PC_COLUMNS 7      /* max, spanning, moments 0..4                    */

/* ABI revision of this header. */
int pc_abi_version(void);

/* Message for the last failing call on this thread ("" if none). */
const char *pc_last_error(void);

/* Bytes per merge record for a graph of num_nodes nodes (4 or 8). */
int pc_record_bytes(uint32_t num_nodes);

/*
 * Device permutations (performance mode).  Fills d_perm[r * perm_stride + i]
 * for r < num_runs, i < num_edges with a uniform random permutation of
 * 0..num_edges-1 for run first_run + r, keyed by (seed, run): a
 * Rao-Sandelius shuffle on Philox4x32-10 words: runs of more than 32768
 * edges are first grouped into buckets of 2048-4096 by random labels in a
 * stable counting sort; every bucket (the whole run up to 32768 edges) is
 * shuffled in LDS by a second level of labels into sub-buckets of about 64
 * and Fisher-Yates with unbiased bounded draws.  Exactly uniform given
 * uniform words.  Replaces the host-side numpy
 * RandomState(seed).permutation(num_edges) of percolate/hpc.py:195,206 when
 * bit-exactness with numpy is not required.  d_workspace holds
 * pc_permute_philox_workspace_size bytes (per-run bucket counts; none, and
 * d_workspace may be NULL, up to 32768 edges).  num_edges <= 2^23 and
 * first_run + num_runs <= 2^32, else PC_EOVERFLOW.
 */
int pc_permute_philox(uint64_t seed, uint64_t first_run, uint32_t num_runs,
                      uint32_t num_edges, int32_t *d_perm, int64_t perm_stride,
                      void *d_workspace, uint64_t workspace_bytes,
                      void *stream);

/*
 * pc_sweep for the square lattice of side `length` (1..16384) without an
 * edge array: each bond's endpoints are computed from its index, in the
 * edge order of the closed-form lattice (rows r = 0..L-1 of nodes
 * (r, c); in every row but the last, node (r, c)'s edge down to (r + 1, c),
 * then right to (r, c + 1) when c < L - 1; the last row's L - 1 edges to the
 * right) -- networkx's order for spanning_2d_grid(L)'s percolation graph,
 * percolate/percolate.py:899-965.  Node (r, c) is r L + c, or with `tiled`
 * (L a multiple of 8) ((r / 4)(L / 8) + c / 8) 32 + (r mod 4) 8 + c mod 8,
 * 4 x 8 tiles of one 128-byte line each.  N = L^2, M = 2 L (L - 1); the
 * other arguments and outputs are pc_sweep's.  Records equal pc_sweep's
 * over the same edges.
 */
int pc_sweep_lattice(uint32_t length, int32_t tiled,
                     const uint8_t *d_node_sides, int32_t span_on_first_merge,
                     const int32_t *d_perm, int64_t perm_stride,
                     uint32_t num_runs, void *d_records, int64_t rec_stride,
                     uint32_t *d_span_step, void *d_workspace,
                     uint64_t workspace_bytes, uint32_t merge_rounds,
                     uint32_t state_mode, uint64_t *d_chunk_state, void *stream);

/*
 * One graph of a persistent multi-graph sweep (pc_sweep_multi): the
 * arguments pc_sweep takes for one graph, with chunk states always written.
 */
typedef struct pc_sweep_graph {
    uint32_t num_nodes, num_edges, num_runs;
    int32_t span_on_first_merge;
    uint32_t merge_rounds;          /* 0: the library default */
    uint32_t lattice_length;        /* 0: the edges are d_edges; L: the square
                                       lattice of side L, edges computed as
                                       pc_sweep_lattice does (tiled order when
                                       num_nodes > 16384), d_edges unused */
    const int32_t *d_edges;         /* [num_edges][2] */
    const uint8_t *d_node_sides;    /* [num_nodes] spanning sides, or NULL */
    const int32_t *d_perm;          /* [num_runs][perm_stride] */
    int64_t perm_stride;
    void *d_records;                /* uint32 (num_nodes <= 65536) or uint64 */
    int64_t rec_stride;
    uint32_t *d_span_step;          /* [num_runs] */
    uint64_t *d_chunk_state;        /* [num_runs][ceil(num_edges / 64)][8] */
} pc_sweep_graph;

/*
 * Persistent sweep of several graphs in one launch: the per-system-size loop
 * of the reference's finite-size study (percolate/share/jugfile.py:178-253,
 * each size's runs through percolate/hpc.py:249-307) as one device work
 * queue.  Every run of every graph is a queue item, numbered in the order of
 * graphs[] (list the largest graphs first: their runs are the longest); each
 * resident workgroup takes the next item with an atomic on a queue head,
 * sweeps it (LDS or HBM union-find state by the graph's size) and repeats
 * until the queue is empty.  Outputs per graph are pc_sweep's records, span
 * steps and chunk states.
 * pc_sweep_multi_workspace_size gives the bytes of the device table (graph
 * descriptors and the queue head) and of the HBM-state workspace (one slot
 * per resident workgroup); pc_sweep_multi_prepare writes the table (a
 * synchronous copy, once per set of buffers); pc_sweep_multi clears the
 * queue head with a kernel and launches the sweep (kernels only, so it may
 * be captured in a HIP graph).  The table must be 256-byte aligned.
 */
int pc_sweep_multi_workspace_size(uint32_t num_graphs,
                                  const pc_sweep_graph *graphs,
                                  uint64_t *table_bytes,
                                  uint64_t *workspace_bytes);
int pc_sweep_multi_prepare(uint32_t num_graphs, const pc_sweep_graph *graphs,
                           void *d_table, uint64_t table_bytes);
int pc_sweep_multi(uint32_t num_graphs, const pc_sweep_graph *graphs,
                   void *d_table, uint64_t table_bytes, void *d_workspace,
                   uint64_t workspace_bytes, void *stream);

/* Device workspace bytes pc_permute_philox needs. */
int pc_permute_philox_workspace_size(uint32_t num_edges, uint32_t num_runs,
                                     uint64_t *bytes);

/* Where pc_sweep keeps the union-find state of a run. */
#define PC_STATE_AUTO 0 /* LDS when num_nodes <= 16384, else HBM             */
#define PC_STATE_LDS 1  /* 16-bit codes in LDS (num_nodes <= 16384)          */
#define PC_STATE_HBM 2  /* 32-bit codes in the HBM workspace                 */

/*
 * Wavefronts per run, OR-ed into state_mode of pc_sweep, pc_sweep_occupancy
 * and pc_sweep_workspace_size (neither: the library picks, two from 4096
 * nodes on or with the state in HBM).  Two: a finder wavefront finds the
 * roots of the next batch of bonds while the resolver merges this one; with
 * the state in HBM and d_chunk_state given, a third folds the records into
 * the chunk states.  Results do not depend on it.
 */
#define PC_SWEEP_ONE_WAVE 0x10
#define PC_SWEEP_TWO_WAVES 0x20
#define PC_SWEEP_WAVES_MASK 0x30

/*
 * Bit-exact device permutations: d_perm[r * perm_stride + i] =
 * numpy.random.RandomState(d_seeds[r]).permutation(num_edges)[i], the
 * permutation the reference draws per run (percolate/hpc.py:195,206), from
 * MT19937 and numpy's legacy shuffle on the GPU.  Up to 32768 edges in two
 * kernels (the swap targets of every step from one wavefront per run, then
 * every swap at once from one workgroup of 1024 threads and 4 bytes of LDS
 * per edge); above, the swaps in sequence, one run per wavefront.  The
 * output row is also the work space between the two kernels.
 */
int pc_permute_randomstate(const uint32_t *d_seeds, uint32_t num_runs,
                           uint32_t num_edges, int32_t *d_perm,
                           int64_t perm_stride, void *stream);

/* Device workspace bytes pc_sweep needs (0 when the state is in LDS). */
int pc_sweep_workspace_size(uint32_t num_nodes, uint32_t num_runs,
                            uint32_t state_mode, uint64_t *bytes);

/*
 * The Newman-Ziff sweep of num_runs independent runs, one per workgroup:
 * replaces the per-edge loop of percolate/hpc.py:249-307 (and its twin
 * percolate/percolate.py:298-356).  Reads d_perm[r * perm_stride + n],
 * writes merge records d_records[r * rec_stride + n] (n < num_edges) and
 * d_span_step[r].  Union-find state lives in LDS or in d_workspace
 * (pc_sweep_workspace_size bytes) per state_mode (PC_STATE_*).
 * merge_rounds is the number of parallel conflict-free merge rounds per
 * batch of 64 bonds before the in-order remainder (0 = library default);
 * results do not depend on it or on state_mode (PC_STATE_* | PC_SWEEP_*).  With d_chunk_state
 * non-NULL the sweep also writes the chunk states of pc_chunk_scan (same
 * layout, d_chunk_state[(r * C + c) * 8 + w]), so that pass is not needed.
 */
int pc_sweep(uint32_t num_nodes, uint32_t num_edges, const int32_t *d_edges,
             const uint8_t *d_node_sides, int32_t span_on_first_merge,
             const int32_t *d_perm, int64_t perm_stride, uint32_t num_runs,
             void *d_records, int64_t rec_stride, uint32_t *d_span_step,
             void *d_workspace, uint64_t workspace_bytes,
             uint32_t merge_rounds, uint32_t state_mode,
             uint64_t *d_chunk_state, void *stream);

/* Resident sweep workgroups (runs) per compute unit for a graph size. */
int pc_sweep_occupancy(uint32_t num_nodes, uint32_t state_mode,
                       int32_t *blocks_per_cu);

/*
 * Per-run state at the start of every chunk of PC_CHUNK steps, from the
 * merge records: d_chunk_state[(r * C + c) * 8 + w], C = ceil(num_edges/64),
 * words {max cluster size, merges, S2, S3 lo, S3 hi, S4 lo, S4 hi, 0} after
 * c * 64 bonds, where Sk is the sum over all clusters of size**k.
 */
int pc_chunk_scan(uint32_t num_nodes, uint32_t num_edges, uint32_t num_runs,
                  const void *d_records, int64_t rec_stride,
                  uint64_t *d_chunk_state, void *stream);

/*
 * Microcanonical reduction over runs, exact: for every step s = 0..num_edges
 * and column k (max_cluster_size, has_spanning_cluster, moments[0..4]) the
 * sum over runs of x (d_sum1[(k * S + s) * 2 + limb], uint128) and of x**2
 * (d_sum2[(k * S + s) * 3 + limb], uint192), S = num_edges + 1.  Replaces the
 * per-step gathers of percolate/percolate.py:862-891 feeding
 * _microcanonical_average_* (percolate.py:575-705).  Moments are exact
 * integers (their low 64 bits are the reference's uint64 values).  With
 * accumulate != 0 the sums are added to the buffers' contents (batches of
 * runs reduced in turn); otherwise the buffers are cleared first.
 */
int pc_microcanonical_sums(uint32_t num_nodes, uint32_t num_edges,
                           uint32_t num_runs, const void *d_records,
                           int64_t rec_stride, const uint32_t *d_span_step,
                           const uint64_t *d_chunk_state, uint64_t *d_sum1,
                           uint64_t *d_sum2, int32_t accumulate, void *stream);

/*
 * Per-run traces, the arrays of percolate/hpc.py:310-404
 * (bond_microcanonical_statistics): d_max_size[r * S + s] (uint32),
 * d_spanning[r * S + s] (uint8) and d_moments[(r * S + s) * 5 + k] (uint64,
 * the reference's mod 2**64 values), S = num_edges + 1.
 */
int pc_expand_traces(uint32_t num_nodes, uint32_t num_edges, uint32_t num_runs,
                     const void *d_records, int64_t rec_stride,
                     const uint32_t *d_span_step, const uint64_t *d_chunk_state,
                     uint32_t *d_max_size, uint8_t *d_spanning,
                     uint64_t *d_moments, void *stream);

/* Precisions of pc_canonical_contract (element type of d_x and d_w). */
#define PC_PREC_FP64 0        /* FP64 inputs, FP64 FMA on the vector ALUs    */
#define PC_PREC_FP64_MFMA 1   /* FP64 inputs, v_mfma_f64_16x16x4f64          */
#define PC_PREC_FP32_MFMA 2   /* FP32 inputs, v_mfma_f32_16x16x4f32, FP32
                                 accumulator added into FP64 every 8 MFMAs  */
#define PC_PREC_BF16X3_MFMA 3 /* FP32 inputs split into three bf16 terms,
                                 six products on v_mfma_f32_16x16x32_bf16  */
#define PC_PREC_SPLIT_BF16 4  /* pc_canonical_fused only: weights and columns
                                 cut into four bytes (exact in bf16), 13
                                 byte products on v_mfma_f32_16x16x32_bf16
                                 with exact FP32 sums per 64 steps          */

/*
 * Convolution window of the binomial weights B(n, k; p): k in
 * [*start, *start + *len), the k with |k - n p| <= 12 sigma + 48 (each tail
 * outside holds at most exp(-72) of the mass).  Host-only, no GPU work.
 */
int pc_binomial_window(uint32_t n, double p, int64_t *start, int64_t *len);

/*
 * Binomial weights B(n, k; p) for num_p values p = d_ps[i] over the windows
 * (d_start[i], d_len[i]): d_w[i * kmax + j] = B(n, d_start[i] + j; p),
 * normalised to sum 1 over the window, 0 for j >= d_len[i].  Replaces the
 * ratio recursion of percolate/percolate.py:1067-1109 (_binomial_pmf) with
 * Loader's saddle-point form in FP64.
 */
int pc_binomial_weights(uint32_t n, uint32_t num_p, const double *d_ps,
                        const int64_t *d_start, const int64_t *d_len,
                        int64_t kmax, double *d_w, void *stream);

/* Device workspace bytes pc_canonical_contract needs at a precision. */
int pc_contract_workspace_size(uint32_t num_p, uint32_t cols, int32_t precision,
                               uint64_t *bytes);

/*
 * Canonical averages, the convolution of percolate/percolate.py:1196-1221
 * and percolate/hpc.py:496-512: d_out[i * cols + c] =
 * sum_j d_w[i * kmax + j] * d_x[(d_start[i] + j) * x_stride + c] for
 * c < cols <= 32 and rows < num_steps, at the given precision (PC_PREC_*;
 * d_x and d_w are double for the FP64 forms, float otherwise).  The MFMA
 * forms sum fixed slices of the steps into d_workspace and add them in a
 * fixed order (deterministic).
 */
int pc_canonical_contract(uint32_t num_steps, uint32_t cols, const void *d_x,
                          int64_t x_stride, uint32_t num_p,
                          const int64_t *d_start, const int64_t *d_len,
                          const void *d_w, int64_t kmax, double *d_out,
                          int32_t precision, void *d_workspace,
                          uint64_t workspace_bytes, void *stream);

/*
 * Per-run observables over a window of steps as FP64 columns, the inputs of
 * the per-run canonical statistics (percolate/hpc.py:443-515):
 * d_out[(s - g_lo) * (R * C) + r * C + c] for s in [g_lo, g_lo + g_len),
 * C = 7 ([has_spanning_cluster,] max_cluster_size, moments[0..4]) with
 * spanning, 6 without.  Moments are their uint64 values as in
 * bond_microcanonical_statistics.
 */
int pc_window_traces(uint32_t num_nodes, uint32_t num_edges, uint32_t num_runs,
                     const void *d_records, int64_t rec_stride,
                     const uint32_t *d_span_step, const uint64_t *d_chunk_state,
                     int64_t g_lo, int64_t g_len, int32_t spanning,
                     double *d_out, void *stream);

/*
 * (mean, M2) over runs of per-run canonical statistics
 * d_q[p * (R * cols) + r * cols + c]: the number_of_runs / *_mean / *_m2 of
 * reduce(bond_reduce, map(bond_initialize_canonical_averages, ...))
 * (percolate/hpc.py:561-702), two-pass in a fixed order.  cols <= 8
 * (7 observables with spanning, 6 without); more is PC_EINVAL.
 */
int pc_run_moments(const double *d_q, uint32_t num_p, uint32_t num_runs,
                   uint32_t cols, double *d_mean, double *d_m2, void *stream);

/* Device workspace bytes pc_canonical_fused needs at a precision. */
int pc_canonical_fused_workspace_size(uint32_t num_p, uint32_t cols,
                                      int32_t precision, uint64_t *bytes);

/*
 * Canonical averages with the binomial weights generated inside the
 * contraction, the whole of _binomial_pmf + the convolution of
 * percolate/percolate.py:1067-1109,1196-1221 (and hpc.py:496-512) without a
 * weight matrix: d_out[i * cols + c] = sum_k B(n, k; p_i) X[k][c] /
 * sum_k B(n, k; p_i), k over the window (d_start[i], d_len[i]) of
 * pc_binomial_window, X[k][c] = d_x[k * x_stride + c] for k = 0..num_edges
 * (n = num_edges).  Weights: Loader's form where a lane's run of
 * consecutive k enters the window, then the exact ratio recursion.
 * precision: PC_PREC_FP64_MFMA (FP64 matrix cores) or PC_PREC_SPLIT_BF16
 * (exact byte split on the bf16 matrix cores).  Deterministic; windows of
 * nearby p share work, so sorted d_ps run fastest.
 */
int pc_canonical_fused(uint32_t num_edges, uint32_t cols, const double *d_x,
                       int64_t x_stride, uint32_t num_p, const double *d_ps,
                       const int64_t *d_start, const int64_t *d_len,
                       double *d_out, int32_t precision, void *d_workspace,
                       uint64_t workspace_bytes, void *stream);

/* Scratch bytes per GPU of pc_canonical_fused_noalloc (32 MiB). */
#define PC_NOALLOC_SCRATCH_BYTES (32u << 20)

/*
 * pc_canonical_fused without a workspace argument and without allocating:
 * no counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The partial sums go to a fixed scratch array of
 * PC_NOALLOC_SCRATCH_BYTES per GPU, which the library holds until the
 * process ends; a problem whose pc_canonical_fused_workspace_size exceeds it
 * fails with PC_ENOMEM (every problem of at most 4096 p and 64 columns fits;
 * at 21 columns, up to 63550 p).  Results are
 * bit-identical to pc_canonical_fused.  Every call on one GPU uses that GPU's
 * one array: calls that overlap on one GPU (other streams or host threads)
 * race and must be serialised by the caller.
 */
int pc_canonical_fused_noalloc(uint32_t num_edges, uint32_t cols,
                               const double *d_x, int64_t x_stride,
                               uint32_t num_p, const double *d_ps,
                               const int64_t *d_start, const int64_t *d_len,
                               double *d_out, int32_t precision, void *stream);

/* FP64 -> FP32 conversion of n elements (inputs of the FP32 precisions). */
int pc_convert_f64_to_f32(const double *d_in, int64_t n, float *d_out,
                          void *stream);

/*
 * Mean and sample standard deviation (ddof = 1) over num_runs runs of each
 * column and step, from pc_microcanonical_sums: d_mean[k * S + s],
 * d_std[k * S + s] (FP64; the variance numerator is formed exactly).  The
 * statistics of percolate/percolate.py:575-705 before the confidence
 * intervals.
 */
int pc_microcanonical_moments(uint32_t num_edges, uint32_t num_runs,
                              const uint64_t *d_sum1, const uint64_t *d_sum2,
                              double *d_mean, double *d_std, void *stream);

/*
 * The reduction of canonical averages on the device: percolate/hpc.py:638-702
 * (bond_reduce, through simoa.stats.online_variance), the pairwise merge of
 * Chan, Golub & LeVeque of two sets of (n, mean, M2) for num_ps rows of
 * num_columns statistics (element p * num_columns + c; the counts d_n_a[p],
 * d_n_b[p] as doubles): mean = mean_a + delta * (n_b / n), M2 = M2_a + M2_b
 * + delta * delta * (n_a * n_b / n), delta = mean_b - mean_a, n = n_a + n_b.
 * d_mean_out and d_m2_out may be d_mean_a and d_m2_a.
 */
int pc_merge_moments(uint32_t num_ps, uint32_t num_columns, const double *d_n_a,
                     const double *d_mean_a, const double *d_m2_a,
                     const double *d_n_b, const double *d_mean_b,
                     const double *d_m2_b, double *d_mean_out, double *d_m2_out,
                     void *stream);

/*
 * Finalized canonical averages on the device: percolate/hpc.py:752-834
 * (finalize_canonical_averages) for num_ps rows of num_columns statistics,
 * element i = p * num_columns + c of d_mean and d_m2 (the mean and the sum
 * of squared deviations over d_runs[p] runs).  d_out[k * P * C + i],
 * k = 0..3: the mean and the sample standard deviation sqrt(M2 / (n - 1)),
 * each divided by d_div[c] (the number of nodes for the percolation
 * strength and the moments, 1 otherwise), and the Student-t interval
 * bounds d_q_lo[p] * (std / sqrt(n)) + mean and d_q_hi[p] * ... + mean,
 * each product and sum rounded separately as scipy.stats.t.interval does
 * (the caller supplies t.ppf of (1 -+ (1 - alpha)) / 2 at n - 1 degrees of
 * freedom); NaN where scipy's ppf refuses its arguments (n < 2, a scale
 * that is not positive, a NaN mean).
 */
int pc_finalize_canonical(uint32_t num_ps, uint32_t num_columns,
                          const uint32_t *d_runs, const double *d_mean,
                          const double *d_m2, const double *d_div,
                          const double *d_q_lo, const double *d_q_hi,
                          double *d_out, void *stream);

/*
 * The microcanonical statistics of every step with their confidence bounds,
 * from pc_microcanonical_sums: the epilogue of percolate/percolate.py:557-705
 * (_microcanonical_average_spanning_cluster, _max_cluster_size, _moments)
 * on the device.  d_out[c * S + s], S = num_edges + 1, 21 columns:
 * 0..2 max cluster size mean, lower, upper bound; 3..5 spanning posterior
 * mean (k + 1) / (R + 2) and the bounds d_beta[2k], d_beta[2k + 1]
 * (k = runs with a spanning cluster; d_beta holds the caller's beta
 * quantiles for k = 0..R); 6 + 3j .. 8 + 3j moment j = 0..4 mean, lower,
 * upper bound.  A bound is mean + (std / sqrt(R)) * t_lo (t_hi), rounded
 * like the host's scale * ppf + loc, where R > 1 and std > 0, else the
 * mean; t_lo, t_hi are the caller's Student-t quantiles (R - 1 degrees of
 * freedom) of the interval.
 */
int pc_microcanonical_ci(uint32_t num_edges, uint32_t num_runs,
                         const uint64_t *d_sum1, const uint64_t *d_sum2,
                         double t_lo, double t_hi, const double *d_beta,
                         double *d_out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* PERCOLATE_HIP_H */
