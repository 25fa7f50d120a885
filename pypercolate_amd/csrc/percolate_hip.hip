// libpercolate_hip: MI355X (gfx950, wave64, CDNA4) kernels for the
// Newman-Ziff bond-percolation hot path, behind the C ABI declared in
// include/percolate_hip.h.
//
// Pipeline for R runs on a graph with N nodes and M edges:
//   permute (Philox-keyed bucket shuffle, or MT19937 + numpy's shuffle for
//   bit-exact RandomState permutations) -> sweep (union-find, one run per
//   workgroup of one or two wavefronts, merge records) -> chunk states
//   (per-run state every 64
//   steps: two passes over the records, or written by the sweep itself) ->
//   microcanonical sums (exact Sigma x, Sigma x^2 over runs per step) or
//   expand (per-run traces) -> canonical averages (binomial convolution on
//   the matrix cores, the weights generated inside the contraction).
//
// The observables of one run at step s (largest cluster, spanning flag,
// moments k = 0..4 of all clusters but the largest) depend only on the set
// of bonds occupied after s steps.  With Sk(s) = sum over all clusters of
// size**k they are exact prefix functions of the merge records (a, b):
//   max(s) = max(1, max over merges <= s of a + b)
//   #clusters(s) = N - merges(s)
//   Sk(s) = N + sum over merges <= s of (a + b)**k - a**k - b**k
//   moments_k(s) = Sk(s) - max(s)**k      (k = 0: #clusters - 1)
// which equals the reference's incremental bookkeeping
// (percolate/hpc.py:282-305) whichever largest cluster it excludes on ties.
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <type_traits>
#include <vector>

#include "percolate_hip.h"
#include "stirlerr_table.h"

typedef unsigned __int128 u128;

// ---------------------------------------------------------------------------
// Error reporting
// ---------------------------------------------------------------------------
static thread_local char g_last_error[512];

static int pc_fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof g_last_error, fmt, ap);
    va_end(ap);
    return code;
}

#define PC_HIP_CHECK(expr)                                                     \
    do {                                                                       \
        hipError_t e_ = (expr);                                                \
        if (e_ != hipSuccess)                                                  \
            return pc_fail(PC_EHIP, "%s failed: %s", #expr,                    \
                           hipGetErrorString(e_));                             \
    } while (0)

#define PC_LAUNCH_CHECK(name)                                                  \
    do {                                                                       \
        hipError_t e_ = hipGetLastError();                                     \
        if (e_ != hipSuccess)                                                  \
            return pc_fail(PC_EHIP, "launch of %s failed: %s", name,           \
                           hipGetErrorString(e_));                             \
    } while (0)

// ---------------------------------------------------------------------------
// Wavefront helpers (wave64).  Scans use DPP row shifts inside each row of
// 16 lanes and the GFX9 row broadcasts 15 / 31 across rows.
// ---------------------------------------------------------------------------
#define DPP_ROW_SHR(n) (0x110 + (n))
#define DPP_ROW_BCAST15 0x142
#define DPP_ROW_BCAST31 0x143

template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ uint32_t dpp32(uint32_t v)
{
    // lanes without a source (row edge or masked row) read 0, the identity
    // of every scan below (sum and unsigned max)
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK,
                                                 0xf, false);
}

template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ uint64_t dpp64(uint64_t v)
{
    return (uint64_t)dpp32<CTRL, ROW_MASK>((uint32_t)v) |
           ((uint64_t)dpp32<CTRL, ROW_MASK>((uint32_t)(v >> 32)) << 32);
}

template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ u128 dpp128(u128 v)
{
    return (u128)dpp64<CTRL, ROW_MASK>((uint64_t)v) |
           ((u128)dpp64<CTRL, ROW_MASK>((uint64_t)(v >> 64)) << 64);
}

template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ uint32_t dpp_any(uint32_t v) { return dpp32<CTRL, ROW_MASK>(v); }
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ uint64_t dpp_any(uint64_t v) { return dpp64<CTRL, ROW_MASK>(v); }
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ u128 dpp_any(u128 v) { return dpp128<CTRL, ROW_MASK>(v); }

// Inclusive prefix sum over the 64 lanes.
template <typename T>
__device__ __forceinline__ T wave_scan_add(T x)
{
    x += dpp_any<DPP_ROW_SHR(1)>(x);
    x += dpp_any<DPP_ROW_SHR(2)>(x);
    x += dpp_any<DPP_ROW_SHR(4)>(x);
    x += dpp_any<DPP_ROW_SHR(8)>(x);
    x += dpp_any<DPP_ROW_BCAST15, 0xa>(x);
    x += dpp_any<DPP_ROW_BCAST31, 0xc>(x);
    return x;
}

// Inclusive prefix maximum over the 64 lanes (unsigned, identity 0).
__device__ __forceinline__ uint32_t wave_scan_max(uint32_t x)
{
    x = max(x, dpp32<DPP_ROW_SHR(1)>(x));
    x = max(x, dpp32<DPP_ROW_SHR(2)>(x));
    x = max(x, dpp32<DPP_ROW_SHR(4)>(x));
    x = max(x, dpp32<DPP_ROW_SHR(8)>(x));
    x = max(x, dpp32<DPP_ROW_BCAST15, 0xa>(x));
    x = max(x, dpp32<DPP_ROW_BCAST31, 0xc>(x));
    return x;
}

__device__ __forceinline__ uint64_t umax64(uint64_t a, uint64_t b) { return a > b ? a : b; }

// The first active lane's value, as a wave-uniform (scalar) value.
__device__ __forceinline__ uint32_t uniform(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}


__device__ __forceinline__ uint32_t read_lane(uint32_t v, uint32_t lane)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)lane);
}
__device__ __forceinline__ uint64_t read_lane(uint64_t v, uint32_t lane)
{
    return (uint64_t)read_lane((uint32_t)v, lane) |
           ((uint64_t)read_lane((uint32_t)(v >> 32), lane) << 32);
}
__device__ __forceinline__ u128 read_lane(u128 v, uint32_t lane)
{
    return (u128)read_lane((uint64_t)v, lane) |
           ((u128)read_lane((uint64_t)(v >> 64), lane) << 64);
}

// Multi-word atomic adds: each word's carry-out is recovered from the old
// value the atomic returns, so concurrent adders compose exactly.
__device__ __forceinline__ void atomic_add_u128(uint64_t *p, u128 v)
{
    uint64_t lo = (uint64_t)v, hi = (uint64_t)(v >> 64);
    if (lo) {
        unsigned long long old = atomicAdd((unsigned long long *)p, lo);
        if (old + lo < old)
            hi += 1;
    }
    if (hi)
        atomicAdd((unsigned long long *)(p + 1), hi);
}

__device__ __forceinline__ void atomic_add_u192(uint64_t *p, uint64_t w0,
                                                uint64_t w1, uint64_t w2)
{
    if (w0) {
        unsigned long long old = atomicAdd((unsigned long long *)p, w0);
        if (old + w0 < old) {
            w1 += 1;
            if (w1 == 0)
                w2 += 1;
        }
    }
    if (w1) {
        unsigned long long old = atomicAdd((unsigned long long *)(p + 1), w1);
        if (old + w1 < old)
            w2 += 1;
    }
    if (w2)
        atomicAdd((unsigned long long *)(p + 2), w2);
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), the counter-based
// generator of the device permutations.
// ---------------------------------------------------------------------------
struct u32x4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0,
                                               uint32_t k1)
{
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        if (i) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        c = u32x4{(uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1,
                  (uint32_t)(p0 >> 32) ^ c.w ^ k1, (uint32_t)p0};
    }
    return c;
}

// ---------------------------------------------------------------------------
// Device permutations (performance mode): a Rao-Sandelius shuffle keyed by
// Philox4x32-10.  Every element draws a uniform random bucket label; the
// elements are grouped by label in a stable counting sort (one radix pass
// over random keys), and every bucket is shuffled uniformly: a second level
// of random labels, then Fisher-Yates with unbiased bounded draws inside
// each sub-bucket.  With independent uniform words the result is an exactly
// uniform permutation: iid labels plus a uniform order inside each label
// class is a sort by iid continuous keys.
//
// Layout (the oracle's nz_shuffle.c restates it; perm[i] = edge index added
// at step i + 1):
//   - M <= SHUF_WHOLE: the whole run is one bucket (lb1 = 0), shuffled by
//     one workgroup straight from the identity.  Otherwise 2^lb1 buckets,
//     lb1 the least with M <= SHUF_BUCKET << lb1; label of element i = the
//     top lb1 bits of key(1, 0, i).
//   - Bucket b of n elements, in their grouped (increasing) order: in runs
//     of several buckets, n > SHUF_GROUP_CAP (odds below e^-1500 at these
//     sizes): plain Fisher-Yates in HBM with 32-bit draws from ctr
//     (u, run, b, 4 << 28).  Otherwise
//     2^lb2 sub-buckets (the least lb2 with n <= SHUF_SUB << lb2), label of
//     position j = the top lb2 bits of key(2, b, j), stable grouping, then
//     Fisher-Yates of each sub-bucket (start st, size c) with 16-bit draws:
//     the halves of ctr (st, run, b, 3 << 28 | u) in turn, a rejected word
//     replaced by the next half of ctr (st, run, b, 5 << 28 | v).  Every
//     accepted draw so takes exactly one word of the first stream, and the
//     lanes of a wave, which start their sub-buckets together, refill it
//     together.
//   - key(tag, b, j) = 16-bit half (j >> 6) & 7 (low half first) of
//     Philox(ctr ((j >> 9) << 6 | (j & 63), run, b, tag << 28)): one Philox
//     call per lane labels 8 positions 64 apart, so a wave labels 512
//     consecutive positions per call and visits them in order.
//   - Bounded draw on [0, r) from a w-bit word x: m = x * r, rejected while
//     (m mod 2^w) < (2^w - r) mod r (Lemire), result m >> w.
// Kernels: for several buckets count (labels per bucket and chunk of
// SHUF_CHUNK elements) -> scan (per run) -> scatter (stable grouping into
// the output row); then group (one workgroup per bucket, in LDS).
// ---------------------------------------------------------------------------
#define SHUF_BUCKET 4096u
#define SHUF_GROUP_CAP 8192u
#define SHUF_WHOLE 32768u
#define SHUF_SUB 64u
#define SHUF_SMALL_CAP 4096u  // one-bucket runs up to this size: 2-wave groups
#define SHUF_CHUNK 4096u
#define SHUF_MAX_EDGES (1u << 23)
#define SHUF_W 4

__device__ __forceinline__ uint32_t half16(const u32x4 &w, uint32_t h)
{
    const uint32_t wd = (h & 4u) ? ((h & 2u) ? w.w : w.z) : ((h & 2u) ? w.y : w.x);
    return (wd >> ((h & 1u) * 16u)) & 0xFFFFu;
}

__device__ __forceinline__ u32x4 shuf_philox(uint32_t x, uint32_t run, uint32_t b,
                                             uint32_t w, uint32_t k0, uint32_t k1)
{
    return philox4x32_10(u32x4{x, run, b, w}, k0, k1);
}

// Lanes (of the active ones in `valid`) whose label equals this lane's, for a
// label of `bits` bits: one ballot per bit.  Per bit, s is all ones where
// the lane's bit is set, and each 32-bit half of the mask keeps the lanes
// whose ballot bit equals s (one three-input bit operation per half).
__device__ __forceinline__ uint64_t match_label(uint32_t label, uint32_t bits,
                                                uint64_t valid)
{
    uint32_t lo = (uint32_t)valid, hi = (uint32_t)(valid >> 32);
    for (uint32_t i = 0; i < bits; ++i) {
        const uint32_t s = (uint32_t)((int32_t)(label << (31u - i)) >> 31);
        const uint64_t ones = __ballot(s != 0u);
        lo &= ~((uint32_t)ones ^ s);
        hi &= ~((uint32_t)(ones >> 32) ^ s);
    }
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask)
{
    const uint32_t lane = threadIdx.x & 63u;
    return (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
}

// Lemire's rejection threshold (2^w - r) mod r for w = 16 and w = 32, called
// only when a draw's low part is below r.  The empty asm hides r from the
// optimiser, which would otherwise hoist the integer division out of the
// rare branch and pay it on every draw.
__device__ __forceinline__ uint32_t lemire_threshold16(uint32_t r)
{
    asm volatile("" : "+v"(r));
    return (65536u - r) % r;
}
__device__ __forceinline__ uint32_t lemire_threshold32(uint32_t r)
{
    asm volatile("" : "+v"(r));
    return (0u - r) % r;
}

// A rejected 16-bit draw (m = x * r with m mod 2^16 < r) redrawn from the
// sub-bucket's replacement stream until accepted; v counts the replacement
// words taken.
__device__ __forceinline__ uint32_t shuffle_redraw(uint32_t m, uint32_t r,
                                                   uint32_t st, uint32_t run,
                                                   uint32_t b, uint32_t k0,
                                                   uint32_t k1, uint32_t &v)
{
    const uint32_t thr = lemire_threshold16(r);
    while ((m & 0xFFFFu) < thr) {
        const u32x4 w = shuf_philox(st, run, b, (5u << 28) | (v >> 3), k0, k1);
        const uint32_t h = v & 7u;
        const uint32_t wd = (h & 4u) ? ((h & 2u) ? w.w : w.z) : ((h & 2u) ? w.y : w.x);
        m = ((wd >> ((h & 1u) * 16u)) & 0xFFFFu) * r;
        ++v;
    }
    return m;
}

// Fisher-Yates of a[0..c) with the 16-bit draw streams of sub-bucket start
// st.  The Philox words are kept as scalars: a struct indexed by the running
// half counter would be demoted to a private (scratch or LDS) array.
template <typename E>
__device__ __forceinline__ void shuffle_sub(E *a, uint32_t c, uint32_t st,
                                            uint32_t run, uint32_t b, uint32_t k0,
                                            uint32_t k1)
{
    uint32_t w0 = 0, w1 = 0, w2 = 0, w3 = 0;
    uint32_t h = 8u, u = 0u, v = 0u;
    for (uint32_t j = c - 1u; j >= 1u; --j) {
        const uint32_t r = j + 1u;
        if (h == 8u) {
            const u32x4 w = shuf_philox(st, run, b, (3u << 28) | u, k0, k1);
            w0 = w.x;
            w1 = w.y;
            w2 = w.z;
            w3 = w.w;
            ++u;
            h = 0u;
        }
        const uint32_t wd = (h & 4u) ? ((h & 2u) ? w3 : w2) : ((h & 2u) ? w1 : w0);
        // 16-bit word times r <= 2^15 + 1: a 24-bit multiply is exact
        uint32_t m = __umul24((wd >> ((h & 1u) * 16u)) & 0xFFFFu, r);
        ++h;
        if (__builtin_expect((m & 0xFFFFu) < r, 0))
            m = shuffle_redraw(m, r, st, run, b, k0, k1, v);
        const uint32_t k = m >> 16;
        const E t = a[j];
        a[j] = a[k];
        a[k] = t;
    }
}

// Fisher-Yates of a bucket of n > SHUF_GROUP_CAP elements in HBM, 32-bit
// draws (one lane).
__device__ __noinline__ void shuffle_wide(int32_t *a, uint32_t n, uint32_t run,
                                          uint32_t b, uint32_t k0, uint32_t k1)
{
    uint32_t w0 = 0, w1 = 0, w2 = 0, w3 = 0;
    uint32_t h = 4u, u = 0u;
    for (uint32_t j = n - 1u; j >= 1u; --j) {
        const uint32_t r = j + 1u;
        uint64_t m;
        for (;;) {
            if (h == 4u) {
                const u32x4 w = shuf_philox(u, run, b, 4u << 28, k0, k1);
                w0 = w.x;
                w1 = w.y;
                w2 = w.z;
                w3 = w.w;
                ++u;
                h = 0u;
            }
            const uint32_t x = h == 0u ? w0 : h == 1u ? w1 : h == 2u ? w2 : w3;
            ++h;
            m = (uint64_t)x * r;
            const uint32_t lo = (uint32_t)m;
            if (__builtin_expect(lo < r, 0) && lo < lemire_threshold32(r))
                continue;
            break;
        }
        const uint32_t k = (uint32_t)(m >> 32);
        const int32_t t = a[j];
        a[j] = a[k];
        a[k] = t;
    }
}

// Bucket labels of a chunk of elements: the wave's groups of 512 positions
// within the chunk, and its range.
struct ShufChunk {
    uint32_t g0, g1;  // the wave's group range (global group indices)
    uint32_t lo, hi;  // the chunk's element range
};

__device__ __forceinline__ ShufChunk shuf_chunk(uint32_t c, uint32_t M)
{
    const uint32_t wave = threadIdx.x >> 6;
    ShufChunk k;
    k.lo = c * SHUF_CHUNK;
    k.hi = min(M, k.lo + SHUF_CHUNK);
    const uint32_t first = k.lo / 512u, G = (k.hi - k.lo + 511u) / 512u;
    k.g0 = first + wave * G / SHUF_W;
    k.g1 = first + (wave + 1u) * G / SHUF_W;
    return k;
}

// Run and chunk of a count or scatter block: block g * 8 * chunks + 8 c + j
// is chunk c of run 8 g + j of the launch's rn runs, so the chunks of a run
// go to one XCD (workgroups are dealt to the 8 XCDs in turn) in chunk order.
// Both kernels write short pieces next to those of the neighbouring chunks
// (a count per bucket, a segment per bucket), and the lines they share fill
// up in that XCD's L2 instead of going out part-written.  False for the
// padding blocks of a last group of fewer than 8 runs.
__device__ __forceinline__ bool shuf_block(uint32_t chunks, uint32_t rn,
                                           uint32_t &rl, uint32_t &c)
{
    const uint32_t per_group = 8u * chunks;
    rl = blockIdx.x / per_group * 8u + blockIdx.x % 8u;
    c = blockIdx.x % per_group / 8u;
    return rl < rn;
}

__host__ __forceinline__ uint32_t shuf_blocks(uint32_t chunks, uint32_t rn)
{
    return (rn + 7u) / 8u * 8u * chunks;
}

// Count kernel: counts[(r * nb1 + b) * chunks + c] = elements of chunk c of
// run r with bucket label b.
__global__ __launch_bounds__(64 * SHUF_W) void shuffle_count_kernel(
    uint32_t k0, uint32_t k1, uint32_t first_run, uint32_t M, uint32_t lb1,
    uint32_t chunks, uint32_t r0, uint32_t rn, uint32_t *__restrict__ counts)
{
    extern __shared__ uint32_t cnt[];  // 2^lb1
    const uint32_t tid = threadIdx.x, lane = tid & 63u, nb1 = 1u << lb1;
    uint32_t rl, c;
    if (!shuf_block(chunks, rn, rl, c))
        return;
    const uint32_t r = r0 + rl, run = first_run + r;
    for (uint32_t i = tid; i < nb1; i += 64u * SHUF_W)
        cnt[i] = 0u;
    __syncthreads();
    const ShufChunk ch = shuf_chunk(c, M);
    for (uint32_t g = ch.g0; g < ch.g1; ++g) {
        const u32x4 key = lb1 ? shuf_philox((g << 6) | lane, run, 0u, 1u << 28, k0, k1)
                              : u32x4{0, 0, 0, 0};
#pragma unroll
        for (uint32_t k = 0; k < 8u; ++k) {
            const uint32_t i = g * 512u + k * 64u + lane;
            if (i < ch.hi)
                atomicAdd(&cnt[lb1 ? half16(key, k) >> (16u - lb1) : 0u], 1u);
        }
    }
    __syncthreads();
    for (uint32_t b = tid; b < nb1; b += 64u * SHUF_W)
        counts[((uint64_t)r * nb1 + b) * chunks + c] = cnt[b];
}

// Scan kernel, one workgroup per run: counts become exclusive prefixes in
// (bucket major, chunk minor) order, starts[r * (nb1 + 1) + b] the first
// position of bucket b (starts[.. + nb1] = M).  Each wave reads whole
// buckets' chunk counts, 64 at a time (coalesced); bucket totals are
// scanned in LDS in between.
__global__ __launch_bounds__(256) void shuffle_scan_kernel(
    uint32_t M, uint32_t lb1, uint32_t chunks, uint32_t r0,
    uint32_t *__restrict__ counts, uint32_t *__restrict__ starts)
{
    extern __shared__ uint32_t tot[];  // 2^lb1 bucket totals, then bases
    __shared__ uint32_t wsum[4];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t r = r0 + blockIdx.x, nb1 = 1u << lb1;
    uint32_t *q = counts + (uint64_t)r * nb1 * chunks;
    uint32_t *st = starts + (uint64_t)r * (nb1 + 1u);
    for (uint32_t b = wave; b < nb1; b += 4u) {
        uint32_t acc = 0;
        for (uint32_t c = lane; c < chunks; c += 64u)
            acc += q[(uint64_t)b * chunks + c];
        acc = wave_scan_add(acc);
        if (lane == 63u)
            tot[b] = acc;
    }
    __syncthreads();
    {
        const uint32_t per = (nb1 + 255u) / 256u;
        const uint32_t a = min(tid * per, nb1), e = min(a + per, nb1);
        uint32_t sum = 0;
        for (uint32_t b = a; b < e; ++b)
            sum += tot[b];
        const uint32_t inc = wave_scan_add(sum);
        if (lane == 63u)
            wsum[wave] = inc;
        __syncthreads();
        uint32_t acc = inc - sum;
        for (uint32_t w = 0; w < wave; ++w)
            acc += wsum[w];
        for (uint32_t b = a; b < e; ++b) {
            const uint32_t t = tot[b];
            tot[b] = acc;
            st[b] = acc;
            acc += t;
        }
        if (tid == 0u)
            st[nb1] = M;
    }
    __syncthreads();
    for (uint32_t b = wave; b < nb1; b += 4u) {
        uint32_t carry = tot[b];
        for (uint32_t c0 = 0; c0 < chunks; c0 += 64u) {
            const uint32_t c = c0 + lane;
            const uint32_t v = c < chunks ? q[(uint64_t)b * chunks + c] : 0u;
            const uint32_t inc = wave_scan_add(v);
            if (c < chunks)
                q[(uint64_t)b * chunks + c] = carry + inc - v;
            carry += read_lane(inc, 63u);
        }
    }
}

// Scatter kernel: the elements of chunk c of run r, sorted by bucket label
// in LDS (stable), then written bucket by bucket with consecutive threads on
// consecutive addresses (whole cache lines, instead of one scattered 4-byte
// store per element); blocks laid out by shuf_block.
__global__ __launch_bounds__(64 * SHUF_W) void shuffle_scatter_kernel(
    uint32_t k0, uint32_t k1, uint32_t first_run, uint32_t M, uint32_t lb1,
    uint32_t chunks, uint32_t r0, uint32_t rn, const uint32_t *__restrict__ counts,
    int32_t *__restrict__ perm, int64_t stride)
{
    __shared__ int32_t sval[SHUF_CHUNK];
    __shared__ uint16_t slab[SHUF_CHUNK];
    __shared__ uint32_t wsum[SHUF_W];
    extern __shared__ uint32_t dyn[];  // [SHUF_W][nb1] bases, lstart, gbase
    const uint32_t T = 64u * SHUF_W, tid = threadIdx.x, lane = tid & 63u,
                   wave = tid >> 6;
    const uint32_t nb1 = 1u << lb1;
    uint32_t rl, c;
    if (!shuf_block(chunks, rn, rl, c))
        return;
    const uint32_t r = r0 + rl, run = first_run + r;
    uint32_t *base = dyn, *lstart = dyn + SHUF_W * nb1, *gbase = lstart + nb1;
    uint32_t *my = base + wave * nb1;
    for (uint32_t i = tid; i < SHUF_W * nb1; i += T)
        base[i] = 0u;
    __syncthreads();
    const ShufChunk ch = shuf_chunk(c, M);
    for (uint32_t g = ch.g0; g < ch.g1; ++g) {
        const u32x4 key = lb1 ? shuf_philox((g << 6) | lane, run, 0u, 1u << 28, k0, k1)
                              : u32x4{0, 0, 0, 0};
#pragma unroll
        for (uint32_t k = 0; k < 8u; ++k) {
            const uint32_t i = g * 512u + k * 64u + lane;
            if (i < ch.hi)
                atomicAdd(&my[lb1 ? half16(key, k) >> (16u - lb1) : 0u], 1u);
        }
    }
    __syncthreads();
    // local bucket starts (exclusive over buckets of the chunk's totals),
    // per-wave local bases, and the chunk's global base of each bucket
    {
        const uint32_t per = (nb1 + T - 1u) / T;
        const uint32_t a = min(tid * per, nb1), e = min(a + per, nb1);
        uint32_t sum = 0;
        for (uint32_t b = a; b < e; ++b)
            for (uint32_t w = 0; w < (uint32_t)SHUF_W; ++w)
                sum += base[w * nb1 + b];
        const uint32_t inc = wave_scan_add(sum);
        if (lane == 63u)
            wsum[wave] = inc;
        __syncthreads();
        uint32_t acc = inc - sum;
        for (uint32_t w = 0; w < wave; ++w)
            acc += wsum[w];
        for (uint32_t b = a; b < e; ++b) {
            lstart[b] = acc;
            gbase[b] = counts[((uint64_t)r * nb1 + b) * chunks + c];
            for (uint32_t w = 0; w < (uint32_t)SHUF_W; ++w) {
                const uint32_t n = base[w * nb1 + b];
                base[w * nb1 + b] = acc;
                acc += n;
            }
        }
    }
    __syncthreads();
    for (uint32_t g = ch.g0; g < ch.g1; ++g) {
        const u32x4 key = lb1 ? shuf_philox((g << 6) | lane, run, 0u, 1u << 28, k0, k1)
                              : u32x4{0, 0, 0, 0};
#pragma unroll
        for (uint32_t k = 0; k < 8u; ++k) {
            const uint32_t i = g * 512u + k * 64u + lane;
            const bool ok = i < ch.hi;
            const uint64_t valid = __ballot(ok);
            const uint32_t bkt = lb1 ? half16(key, k) >> (16u - lb1) : 0u;
            const uint64_t peers = match_label(bkt, lb1, valid);
            if (ok) {
                const uint32_t rank = lanes_below(peers);
                const uint32_t at = my[bkt];
                sval[at + rank] = (int32_t)i;
                slab[at + rank] = (uint16_t)bkt;
                if (rank == 0u)
                    my[bkt] = at + (uint32_t)__popcll(peers);
            }
        }
    }
    __syncthreads();
    int32_t *row = perm + (int64_t)r * stride;
    const uint32_t n = ch.hi - ch.lo;
    for (uint32_t t = tid; t < n; t += T) {
        const uint32_t b = slab[t];
        row[gbase[b] + (t - lstart[b])] = sval[t];
    }
}

// Words per wave of the sub-bucket counters (16-bit, in pairs) for groups of
// at most cap elements.
__host__ __device__ __forceinline__ uint32_t shuf_hist_words(uint32_t cap)
{
    uint32_t lb = 0;
    while ((SHUF_SUB << lb) < cap)
        ++lb;
    return ((1u << lb) + 1u) / 2u;
}

// One bucket a[0..n), n <= the launch's cap, shuffled by a workgroup of W
// waves: sub-bucket labels, stable grouping of the local indices 0..n-1 in
// ent[] (16-bit), Fisher-Yates of each sub-bucket (one thread each), then
// a[t] = old a[ent[t]] (all loads in flight together, all stores after a
// barrier), or a[t] = ent[t] when a is the identity (IDENT: the whole run
// is one bucket).  hist holds W * hw words, wsum W; GPW bounds the groups
// of 512 positions a wave labels.
template <int W, int GPW, bool IDENT>
__device__ void shuffle_group_lds(uint32_t n, uint32_t b, uint32_t run,
                                  uint32_t k0, uint32_t k1, int32_t *a,
                                  uint16_t *ent, uint32_t *hist, uint32_t hw,
                                  uint32_t *wsum)
{
    const uint32_t T = 64u * W, tid = threadIdx.x, wave = tid >> 6,
                   lane = tid & 63u;
    uint32_t lb = 0;
    while ((SHUF_SUB << lb) < n)
        ++lb;
    const uint32_t nb = 1u << lb;
    const uint32_t G = (n + 511u) / 512u;
    const uint32_t g0 = wave * G / W, g1 = (wave + 1u) * G / W;
    uint32_t *my = hist + wave * hw;

    for (uint32_t i = tid; i < W * hw; i += T)
        hist[i] = 0u;
    __syncthreads();

    // labels of this wave's positions (8 per Philox call, kept for the
    // grouping below), counted per wave
    u32x4 keys[GPW];
#pragma unroll
    for (uint32_t gi = 0; gi < (uint32_t)GPW; ++gi) {
        const uint32_t g = g0 + gi;
        keys[gi] = (lb && g < g1)
                       ? shuf_philox((g << 6) | lane, run, b, 2u << 28, k0, k1)
                       : u32x4{0, 0, 0, 0};
        if (g < g1) {
#pragma unroll
            for (uint32_t k = 0; k < 8u; ++k) {
                const uint32_t j = g * 512u + k * 64u + lane;
                if (j < n) {
                    const uint32_t s = lb ? half16(keys[gi], k) >> (16u - lb) : 0u;
                    atomicAdd(&my[s >> 1], 1u << ((s & 1u) * 16u));
                }
            }
        }
    }
    __syncthreads();

    // exclusive prefix over (sub-bucket major, wave minor); each thread owns
    // whole words (both 16-bit counters of a sub-bucket pair, all waves)
    {
        const uint32_t pairs = (nb + 1u) / 2u;
        const uint32_t per = (pairs + T - 1u) / T;
        const uint32_t p0 = min(tid * per, pairs), p1 = min(p0 + per, pairs);
        uint32_t tot = 0;
        for (uint32_t p = p0; p < p1; ++p)
            for (uint32_t w = 0; w < (uint32_t)W; ++w) {
                const uint32_t word = hist[w * hw + p];
                tot += (word & 0xFFFFu) + (word >> 16);
            }
        const uint32_t inc = wave_scan_add(tot);
        if (lane == 63u)
            wsum[wave] = inc;
        __syncthreads();
        uint32_t run_ = inc - tot;
        for (uint32_t w = 0; w < wave; ++w)
            run_ += wsum[w];
        for (uint32_t p = p0; p < p1; ++p) {
            // sub-bucket 2p (low halves) of every wave, then 2p + 1 (high)
            uint32_t lo_total = 0;
            for (uint32_t w = 0; w < (uint32_t)W; ++w)
                lo_total += hist[w * hw + p] & 0xFFFFu;
            uint32_t a0 = run_, a1 = run_ + lo_total;
            for (uint32_t w = 0; w < (uint32_t)W; ++w) {
                const uint32_t word = hist[w * hw + p];
                hist[w * hw + p] = a0 | (a1 << 16);
                a0 += word & 0xFFFFu;
                a1 += word >> 16;
            }
            run_ = a1;
        }
    }
    __syncthreads();

    // stable grouping of the local indices: positions in increasing order,
    // rank among the lanes of the same label, the label's first lane
    // advances the wave's base
#pragma unroll
    for (uint32_t gi = 0; gi < (uint32_t)GPW; ++gi) {
        const uint32_t g = g0 + gi;
        if (g < g1) {
#pragma unroll
            for (uint32_t k = 0; k < 8u; ++k) {
                const uint32_t j = g * 512u + k * 64u + lane;
                const bool ok = j < n;
                const uint64_t valid = __ballot(ok);
                const uint32_t s = lb ? half16(keys[gi], k) >> (16u - lb) : 0u;
                const uint64_t peers = match_label(s, lb, valid);
                if (ok) {
                    const uint32_t rank = lanes_below(peers);
                    const uint32_t sh = (s & 1u) * 16u;
                    const uint32_t at = (my[s >> 1] >> sh) & 0xFFFFu;
                    ent[at + rank] = (uint16_t)j;
                    if (rank == 0u)
                        atomicAdd(&my[s >> 1], (uint32_t)__popcll(peers) << sh);
                }
            }
        }
    }
    __syncthreads();

    // Fisher-Yates per sub-bucket: after the grouping the last wave's base
    // of sub-bucket s is the end of s
    const uint32_t *last = hist + (W - 1) * hw;
    for (uint32_t s = tid; s < nb; s += T) {
        const uint32_t en = (last[s >> 1] >> ((s & 1u) * 16u)) & 0xFFFFu;
        const uint32_t st =
            s ? (last[(s - 1u) >> 1] >> (((s - 1u) & 1u) * 16u)) & 0xFFFFu : 0u;
        if (en - st >= 2u)
            shuffle_sub(ent + st, en - st, st, run, b, k0, k1);
    }
    __syncthreads();

    if (IDENT) {
        for (uint32_t t = tid; t < n; t += T)
            a[t] = (int32_t)ent[t];
        return;
    }
    // a[t] = old a[ent[t]]: every load issued before any store of the
    // workgroup (the empty asm makes each thread wait for its loads before
    // the barrier)
    constexpr uint32_t GATHER = IDENT ? 1u : SHUF_GROUP_CAP / (64u * W);
    int32_t v[GATHER];
#pragma unroll
    for (uint32_t i = 0; i < GATHER; ++i) {
        const uint32_t t = tid + i * T;
        v[i] = t < n ? a[ent[t]] : 0;
    }
#pragma unroll
    for (uint32_t i = 0; i < GATHER; ++i)
        asm volatile("" ::"v"(v[i]));
    __syncthreads();
#pragma unroll
    for (uint32_t i = 0; i < GATHER; ++i) {
        const uint32_t t = tid + i * T;
        if (t < n)
            a[t] = v[i];
    }
}

// Group kernel: bucket b = blockIdx.x of run r0 + blockIdx.y, in place in
// the output row; cap bounds the buckets' sizes handled in LDS (the run's
// size when IDENT: one bucket, the identity before the shuffle).  Dynamic
// LDS: shuf_group_lds_bytes(W, cap).
template <int W, bool IDENT>
__global__ __launch_bounds__(64 * W) void shuffle_group_kernel(
    uint32_t k0, uint32_t k1, uint32_t first_run, uint32_t lb1, uint32_t r0,
    uint32_t cap, const uint32_t *__restrict__ starts,
    int32_t *__restrict__ perm, int64_t stride)
{
    extern __shared__ uint32_t gdyn[];
    constexpr uint32_t CAP =
        IDENT ? (W == 2 ? SHUF_SMALL_CAP : SHUF_WHOLE) : SHUF_GROUP_CAP;
    constexpr int GPW = (int)(CAP / (512u * W));
    const uint32_t hw = shuf_hist_words(cap);
    uint32_t *hist = gdyn, *wsum = gdyn + W * hw;
    uint16_t *ent = (uint16_t *)(wsum + W);
    const uint32_t b = blockIdx.x, r = r0 + blockIdx.y;
    const uint32_t run = first_run + r;
    uint32_t o = 0, n = cap;
    if (!IDENT) {
        const uint32_t *st = starts + (uint64_t)r * ((1u << lb1) + 1u);
        o = st[b];
        n = st[b + 1u] - o;
    }
    int32_t *a = perm + (int64_t)r * stride + o;
    if (n <= 1u) {
        if (IDENT && n == 1u && threadIdx.x == 0u)
            a[0] = 0;
        return;
    }
    if (!IDENT && n > SHUF_GROUP_CAP) {
        if (threadIdx.x == 0u)
            shuffle_wide(a, n, run, b, k0, k1);
        return;
    }
    shuffle_group_lds<W, GPW, IDENT>(n, b, run, k0, k1, a, ent, hist, hw, wsum);
}

static size_t shuf_group_lds_bytes(uint32_t W, uint32_t cap)
{
    return (size_t)(W * shuf_hist_words(cap) + W) * 4u + (((size_t)cap * 2u + 3u) & ~(size_t)3u);
}

// ---------------------------------------------------------------------------
// Bit-exact device permutations: numpy's legacy RandomState(seed).permutation
// (the permutations the reference draws, percolate/hpc.py:195,206).
// MT19937 (Matsumoto & Nishimura, 1998) seeded with init_genrand(seed), as
// RandomState does for an integer seed; permutation(n) shuffles arange(n)
// from the end: for i = n-1 .. 1, j is the first 32-bit draw whose low bits
// under the smallest all-ones mask >= i are <= i, then a[i] and a[j] swap.
// One run per wavefront: the 624-word twist is computed by all lanes in four
// phases that respect the recurrence's order, the draws are consumed on
// uniform (scalar) control flow and lane 0 performs the swaps, in LDS
// (16-bit entries, n <= 65536) or in the output row.
// ---------------------------------------------------------------------------
#define MT_N 624
#define MT_SHIFT 397
#define MT_LDS_MAX 65536u

__device__ __forceinline__ uint32_t mt_next_word(uint32_t cur, uint32_t nxt,
                                                 uint32_t far)
{
    const uint32_t y = (cur & 0x80000000u) | (nxt & 0x7fffffffu);
    return far ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
}

// new mt[k] for k in [lo, hi), reading mt[k + 397] (k < 227) or the already
// renewed mt[k - 227] (k >= 227)
__device__ __forceinline__ void mt_twist_range(uint32_t *mt, uint32_t lo,
                                               uint32_t hi)
{
    for (uint32_t k0 = lo; k0 < hi; k0 += 64u) {
        const uint32_t k = k0 + threadIdx.x;
        uint32_t v = 0;
        if (k < hi)
            v = mt_next_word(mt[k], mt[k + 1u == MT_N ? 0u : k + 1u],
                             mt[k < MT_N - MT_SHIFT ? k + MT_SHIFT : k - (MT_N - MT_SHIFT)]);
        if (k < hi)
            mt[k] = v;
    }
}

__device__ __forceinline__ void mt_twist(uint32_t *mt)
{
    mt_twist_range(mt, 0u, MT_N - MT_SHIFT);        // 0..226: old words
    mt_twist_range(mt, MT_N - MT_SHIFT, 2u * (MT_N - MT_SHIFT)); // 227..453
    mt_twist_range(mt, 2u * (MT_N - MT_SHIFT), MT_N - 1u);      // 454..622
    mt_twist_range(mt, MT_N - 1u, MT_N);            // 623: renewed mt[0], mt[396]
}

__device__ __forceinline__ uint32_t mt_temper(uint32_t y)
{
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

template <bool LDS_ARRAY>
__global__ __launch_bounds__(64) void permute_randomstate_kernel(
    const uint32_t *__restrict__ seeds, uint32_t n, int32_t *__restrict__ perm,
    int64_t stride)
{
    __shared__ uint32_t mt[MT_N];
    extern __shared__ uint16_t arr16[];
    const uint32_t lane = threadIdx.x;
    int32_t *out = perm + (int64_t)blockIdx.x * stride;
    if (lane == 0) {
        uint32_t x = seeds[blockIdx.x];
        mt[0] = x;
        for (uint32_t i = 1; i < MT_N; ++i) {
            x = 1812433253u * (x ^ (x >> 30)) + i;
            mt[i] = x;
        }
    }
    for (uint32_t i = lane; i < n; i += 64u) {
        if (LDS_ARRAY)
            arr16[i] = (uint16_t)i;
        else
            out[i] = (int32_t)i;
    }
    __syncthreads();

    // The draws do not depend on the array: the wave tempers the next 64
    // words of the state at once (lane k holds word base + k) and each draw
    // is a lane read, so the only LDS round trip per step is the swap's.
    uint32_t pos = MT_N, base = MT_N;
    uint32_t buf = 0;
    uint32_t i = n ? n - 1u : 0u;
    while (i >= 1u) {
        uint32_t mask = i;
        mask |= mask >> 1;
        mask |= mask >> 2;
        mask |= mask >> 4;
        mask |= mask >> 8;
        mask |= mask >> 16;
        uint32_t v;
        do {
            if (pos == MT_N) {
                __syncthreads();
                mt_twist(mt);
                __syncthreads();
                pos = 0;
                base = MT_N;
            }
            if (pos - base >= 64u) {
                // refill (base = MT_N after a twist makes pos - base wrap)
                base = pos;
                buf = base + lane < MT_N ? mt_temper(mt[base + lane]) : 0u;
            }
            v = read_lane(buf, pos - base) & mask;
            ++pos;
        } while (v > i);
        if (lane == 0) {
            if (LDS_ARRAY) {
                const uint16_t a = arr16[i], b = arr16[v];
                arr16[i] = b;
                arr16[v] = a;
            } else {
                const int32_t a = out[i], b = out[v];
                out[i] = b;
                out[v] = a;
            }
        }
        --i;
    }
    if (LDS_ARRAY) {
        __syncthreads();
        for (uint32_t k = lane; k < n; k += 64u)
            out[k] = (int32_t)arr16[k];
    }
}

// The same permutation in two parallel phases (runs of at most
// RS_APPLY_MAX edges).  The shuffle's swap targets depend on the MT19937
// stream alone, not on the array: phase one draws them (one wavefront per
// run) and stores the target k_i of step i in the output row; phase two
// applies all swaps at once.  Steps run i = n-1 .. 1 and step i writes only
// positions i and k_i <= i, so no later step writes position i again and
// the result at position i >= 1 is the value position k_i holds just before
// step i.  The writers of a position p before step i are the steps s > i
// with k_s = p, the last of them being the smallest; the value such a step
// s brings to p is the one position s holds just before step s.  With
//   F(p) = the smallest step s > p with k_s = p (none possible),
//   V(p) = the value position p holds just before step p
//        = V(F(p)), or p itself when F(p) does not exist,
// V is the root of p's chain p -> F(p) -> F(F(p)) ... (F only increases),
// the result at i >= 1 is V(s) for the smallest step s > i with k_s = k_i,
// or k_i itself without one, and the result at 0 is V(0).
#define RS_APPLY_MAX 32768u
#define RS_APPLY_T 1024u

// Phase one.  The wave tempers 64 state words at a time (lane l holds word
// base + l) and settles the whole block at once: for steps i, i-1, ... under
// one mask m (the smallest all-ones mask >= i), the word of lane l serves
// step i - (accepted words below l), so x = word & m is accepted for sure
// when x <= i - 63, rejected for sure when x > i, and only the few lanes in
// between are decided one by one in lane order from the count of accepted
// lanes below them.  The mask holds for c = i - m/2 more steps; when the
// block accepts c words or more, the lanes after the c-th accepted one are
// settled again under the next mask.
__global__ __launch_bounds__(64) void randomstate_targets_kernel(
    const uint32_t *__restrict__ seeds, uint32_t n, int32_t *__restrict__ perm,
    int64_t stride)
{
    __shared__ uint32_t mt[MT_N];
    const uint32_t lane = threadIdx.x;
    int32_t *out = perm + (int64_t)blockIdx.x * stride;
    if (lane == 0) {
        uint32_t x = seeds[blockIdx.x];
        mt[0] = x;
        for (uint32_t i = 1; i < MT_N; ++i) {
            x = 1812433253u * (x ^ (x >> 30)) + i;
            mt[i] = x;
        }
        if (n)
            out[0] = 0;
    }
    __syncthreads();
    uint32_t pos = MT_N;
    uint32_t i = n ? n - 1u : 0u;
    while (i >= 1u) {
        if (pos == MT_N) {
            __syncthreads();
            mt_twist(mt);
            __syncthreads();
            pos = 0;
        }
        const uint32_t nw = min(64u, MT_N - pos);
        const uint32_t w = lane < nw ? mt_temper(mt[pos + lane]) : 0u;
        uint32_t l0 = 0;
        do {
            const uint32_t m = 0xFFFFFFFFu >> __builtin_clz(i);
            const uint32_t c = i - (m >> 1);
            const uint32_t x = w & m;
            const bool valid = lane >= l0 && lane < nw;
            uint64_t acc = __ballot(valid && x + 63u <= i);
            uint64_t amb = __ballot(valid && x <= i && x + 63u > i);
            while (amb) {
                const uint32_t l = (uint32_t)__builtin_ctzll(amb);
                const uint32_t below = (uint32_t)__popcll(acc & ((1ull << l) - 1ull));
                if (read_lane(x, l) + below <= i)
                    acc |= 1ull << l;
                amb &= amb - 1ull;
            }
            const uint32_t below = __builtin_amdgcn_mbcnt_hi(
                (uint32_t)(acc >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)acc, 0u));
            const bool mine = (acc >> lane) & 1ull;
            const uint32_t cnt = (uint32_t)__popcll(acc);
            uint32_t used;
            if (cnt >= c) {
                // the c-th accepted lane takes the mask's last step
                const uint32_t last = (uint32_t)__builtin_ctzll(__ballot(mine && below == c - 1u));
                if (mine && below < c)
                    out[i - below] = (int32_t)x;
                used = c;
                l0 = last + 1u;
            } else {
                if (mine)
                    out[i - below] = (int32_t)x;
                used = cnt;
                l0 = nw;
            }
            i -= used;
        } while (l0 < nw && i >= 1u);
        pos += nw;
    }
}

// Smallest of the list entries ent[b..e) above thr, or 0xFFFF (entries are
// steps, 1 <= s < n <= 32768), and whether an entry below thr exists; four
// entries are read per round.
__device__ __forceinline__ uint32_t rs_list_scan(const uint16_t *ent, uint32_t b,
                                                 uint32_t e, uint32_t thr, bool &below)
{
    uint32_t best = 0xFFFFu;
    bool lo = false;
    for (uint32_t a = b; a < e; a += 4u) {
        uint32_t x[4];
#pragma unroll
        for (uint32_t u = 0; u < 4u; ++u)
            x[u] = ent[min(a + u, e - 1u)];
#pragma unroll
        for (uint32_t u = 0; u < 4u; ++u) {
            best = x[u] > thr ? min(best, x[u]) : best;
            lo |= x[u] < thr;
        }
    }
    below = lo;
    return best;
}

// A copy of x the compiler cannot relate to x: predicates computed from it in
// one phase are not kept alive (as 32 lane masks) for a later phase.
__device__ __forceinline__ uint32_t opaque(uint32_t x)
{
    asm volatile("" : "+s"(x));
    return x;
}
__device__ __forceinline__ uint32_t opaque_v(uint32_t x)
{
    asm volatile("" : "+v"(x));
    return x;
}

// Phase two, one workgroup of RS_APPLY_T threads per run of n <= RS_APPLY_MAX
// edges, in place over the targets.  LDS holds 16-bit list cursors of the
// positions in pairs and the lists of steps per target position (a counting
// sort: counts, exclusive scan, placement; unordered within a list); every
// thread owns indices j = tid + q * RS_APPLY_T and scans the list of k_j
// once, for the smallest step above j and for whether j is the list's
// smallest step (which then knows F(k_j)).  The list area is then reused for
// V, which pointer jumping (V(p) <- V(V(p)) until nothing changes, over
// position pairs) takes to the chain roots.  The loads of independent items
// are issued together before any is waited for.  Dynamic LDS:
// rs_apply_lds_bytes(n).
__global__ __launch_bounds__(RS_APPLY_T) void randomstate_apply_kernel(
    uint32_t n, int32_t *__restrict__ perm, int64_t stride)
{
    constexpr uint32_t T = RS_APPLY_T, PER = RS_APPLY_MAX / RS_APPLY_T;
    constexpr uint32_t PAIRS = PER / 2u, WAVES = T / 64u, LOOKUP = 0x8000u;
    constexpr uint32_t GROUP = 4u;
    extern __shared__ uint32_t rsd[];
    const uint32_t words = (n + 2u) / 2u;   // 16-bit cursors of positions 0..n
    uint32_t *cur = rsd;
    uint16_t *ent = (uint16_t *)(rsd + words);
    uint32_t *V2 = rsd + words;             // V of positions 2w, 2w + 1
    const uint16_t *V = ent;
    __shared__ uint32_t wsum[WAVES];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    int32_t *row = perm + (int64_t)blockIdx.x * stride;
    if (n == 0u)
        return;
    const uint32_t npairs = (n + 1u) / 2u;
    uint32_t k[PER];
#pragma unroll
    for (uint32_t q = 0; q < PER; ++q) {
        const uint32_t j = tid + q * T;
        const uint32_t v = (uint32_t)row[min(j, n - 1u)];
        k[q] = (j >= 1u && j < n) ? v : 0u;
    }
    for (uint32_t w = tid; w < words; w += T)
        cur[w] = 0u;
    __syncthreads();
    // steps per target position
    {
        const uint32_t nn = opaque(n), t0 = opaque_v(tid);
#pragma unroll
        for (uint32_t q = 0; q < PER; ++q) {
            const uint32_t j = t0 + q * T;
            if (j >= 1u && j < nn)
                atomicAdd(&cur[k[q] >> 1], 1u << ((k[q] & 1u) * 16u));
        }
    }
    __syncthreads();
    // exclusive prefix: cur[p] = first slot of position p's list.  Wave w
    // scans words [w * chunk, (w + 1) * chunk), 64 consecutive words per
    // round, after one pass for the wave totals.
    {
        const uint32_t chunk = ((words + WAVES - 1u) / WAVES + 63u) & ~63u;
        const uint32_t w0 = min(wave * chunk, words), w1 = min(w0 + chunk, words);
        uint32_t tot = 0;
        for (uint32_t w = w0 + lane; w < w1; w += 64u) {
            const uint32_t x = cur[w];
            tot += (x & 0xFFFFu) + (x >> 16);
        }
        tot = read_lane(wave_scan_add(tot), 63u);
        if (lane == 0u)
            wsum[wave] = tot;
        __syncthreads();
        uint32_t carry = 0;
        for (uint32_t w = 0; w < wave; ++w)
            carry += wsum[w];
        for (uint32_t b = w0; b < w1; b += 64u) {
            const uint32_t w = b + lane;
            const uint32_t x = w < w1 ? cur[w] : 0u;
            const uint32_t pair = (x & 0xFFFFu) + (x >> 16);
            const uint32_t inc = wave_scan_add(pair);
            const uint32_t a0 = carry + inc - pair, a1 = a0 + (x & 0xFFFFu);
            if (w < w1)
                cur[w] = a0 | (a1 << 16);
            carry += read_lane(inc, 63u);
        }
    }
    __syncthreads();
    // place the steps; afterwards cursor p is the end of p's list, which is
    // where p + 1's list starts
    const uint32_t np = opaque(n), tp = opaque_v(tid);
#pragma unroll
    for (uint32_t q = 0; q < PER; ++q) {
        const uint32_t j = tp + q * T;
        if (j >= 1u && j < np) {
            const uint32_t sh = (k[q] & 1u) * 16u;
            const uint32_t at = (atomicAdd(&cur[k[q] >> 1], 1u << sh) >> sh) & 0xFFFFu;
            ent[at] = (uint16_t)j;
        }
    }
    __syncthreads();
    // Each step j scans the list of its target p = k_j once: the smallest
    // step above j there (its successor), and whether j is the list's
    // smallest step, in which case j knows F(p): j itself when j > p, else
    // (j = p) its successor.  Per index, packed: low half the result, or
    // LOOKUP | s for "V(s)"; bit 31 set for a list's smallest step, with
    // F(p) in bits 16..30 (all values are below 2^15).  The list bounds of a
    // group of indices are read before the lists.
    auto half = [&](uint32_t x, uint32_t p) { return (x >> ((p & 1u) * 16u)) & 0xFFFFu; };
    uint32_t rf[PER];
    const uint32_t nr = opaque(n), tr = opaque_v(tid);
#pragma unroll
    for (uint32_t g = 0; g < PER; g += GROUP) {
        uint32_t se[GROUP], sb[GROUP];
#pragma unroll
        for (uint32_t u = 0; u < GROUP; ++u) {
            const uint32_t p = k[g + u];
            se[u] = cur[p >> 1];
            sb[u] = cur[p ? (p - 1u) >> 1 : 0u];     // unused for p = 0
        }
#pragma unroll
        for (uint32_t u = 0; u < GROUP; ++u) {
            const uint32_t q = g + u, j = tr + q * T;
            uint32_t r = LOOKUP, w = 0u;
            if (j < nr && j) {
                const uint32_t p = k[q];
                const uint32_t b = p ? half(sb[u], p - 1u) : 0u, e = half(se[u], p);
                bool below;
                const uint32_t s = rs_list_scan(ent, b, e, j, below);
                r = s == 0xFFFFu ? p : (LOOKUP | s);
                if (!below)
                    w = 0x80000000u | ((j > p ? j : (s == 0xFFFFu ? p : s)) << 16);
            }
            rf[q] = r | w;
        }
    }
    __syncthreads();
    // the results move to the cursor area; V, in the list area, starts as
    // every position itself (the positions no step targets keep that), then
    // each list's smallest step writes F of its target
    uint16_t *res = (uint16_t *)cur;
    {
        uint16_t *Vw = ent;
        const uint32_t nv = opaque(n), tv = opaque_v(tid);
#pragma unroll
        for (uint32_t q = 0; q < PER; ++q) {
            const uint32_t j = tv + q * T;
            if (j < nv) {
                res[j] = (uint16_t)rf[q];
                Vw[j] = (uint16_t)j;
            }
        }
        __syncthreads();
#pragma unroll
        for (uint32_t q = 0; q < PER; ++q)
            if (rf[q] >> 31)
                Vw[k[q]] = (uint16_t)((rf[q] >> 16) & 0x7FFFu);
    }
    __syncthreads();
    // pointer jumping over the pairs; a concurrent update only moves a
    // pointer further up its own chain, and a round without a change leaves
    // every pointer at a root.  An odd n's last pair has a padding half,
    // kept as it is.
    for (;;) {
        uint32_t v[PAIRS], y0[PAIRS], y1[PAIRS];
        const uint32_t tj = opaque_v(tid);
#pragma unroll
        for (uint32_t h = 0; h < PAIRS; ++h)
            v[h] = V2[min(tj + h * T, npairs - 1u)];
#pragma unroll
        for (uint32_t h = 0; h < PAIRS; ++h) {
            y0[h] = V[v[h] & 0xFFFFu];
            y1[h] = V[min(v[h] >> 16, n - 1u)];
        }
        bool changed = false;
#pragma unroll
        for (uint32_t h = 0; h < PAIRS; ++h) {
            const uint32_t w = tj + h * T;
            const uint32_t hi = 2u * w + 1u < n ? y1[h] : v[h] >> 16;
            const uint32_t nv = y0[h] | (hi << 16);
            const bool c = w < npairs && nv != v[h];
            if (c)
                V2[w] = nv;
            changed |= c;
        }
        if (!__syncthreads_or(changed))
            break;
    }
    const uint32_t tf = opaque_v(tid);
#pragma unroll
    for (uint32_t g = 0; g < PER; g += GROUP) {
        uint32_t r[GROUP], val[GROUP];
#pragma unroll
        for (uint32_t u = 0; u < GROUP; ++u)
            r[u] = res[min(tf + (g + u) * T, n - 1u)];
#pragma unroll
        for (uint32_t u = 0; u < GROUP; ++u)
            val[u] = (r[u] & LOOKUP) ? (uint32_t)V[min(r[u] & ~LOOKUP, n - 1u)] : r[u];
#pragma unroll
        for (uint32_t u = 0; u < GROUP; ++u) {
            const uint32_t j = tf + (g + u) * T;
            if (j < n)
                row[j] = (int32_t)val[u];
        }
    }
}

static size_t rs_apply_lds_bytes(uint32_t n)
{
    return (size_t)((n + 2u) / 2u) * 4u + (((size_t)n * 2u + 3u) & ~(size_t)3u);
}

// ---------------------------------------------------------------------------
// Union-find state.  A node's code is either its parent's index (non-root)
// or, for a root, the cluster size and which spanning side (0 none, 1 first,
// 2 second) the cluster touches.  Touching both sides is recorded once as
// the run's spanning step and not stored (spanning is monotone).
// ---------------------------------------------------------------------------

// 16-bit codes in LDS for N <= 16384, with P = 2^shift the least power of
// two >= N: c < N parent; c >= P root with c = (1 + f) * P + (size - 1).
// The largest code, 4P - 1, fits 16 bits; decoding is a shift and a mask
// (the resolver's rounds are bound by the instructions they issue).
struct LdsState {
    uint16_t *code;
    uint32_t P, shift;
    __device__ __forceinline__ uint32_t get(uint32_t i) const { return code[i]; }
    __device__ __forceinline__ void set(uint32_t i, uint32_t v) const { code[i] = (uint16_t)v; }
    __device__ __forceinline__ bool is_root(uint32_t v) const { return v >= P; }
    __device__ __forceinline__ uint32_t root(uint32_t size, uint32_t f) const { return ((1u + f) << shift) + size - 1u; }
    // LDS operations of one wavefront complete in order
    __device__ __forceinline__ void sync() const {}
    __device__ __forceinline__ void decode(uint32_t v, uint32_t &size, uint32_t &f) const
    {
        f = (v >> shift) - 1u;
        size = (v & (P - 1u)) + 1u;
    }
};

// 32-bit codes in HBM for larger N: bit 31 marks a root, bits 29-30 the
// side, bits 0-28 the size.
struct GlobalState {
    uint32_t *code;
    __device__ __forceinline__ uint32_t get(uint32_t i) const { return code[i]; }
    __device__ __forceinline__ void set(uint32_t i, uint32_t v) const { code[i] = v; }
    __device__ __forceinline__ bool is_root(uint32_t v) const { return (v >> 31) != 0u; }
    __device__ __forceinline__ uint32_t root(uint32_t size, uint32_t f) const { return 0x80000000u | (f << 29) | size; }
    // stores of some lanes must be visible to loads of others
    __device__ __forceinline__ void sync() const { __threadfence_block(); }
    __device__ __forceinline__ void decode(uint32_t v, uint32_t &size, uint32_t &f) const
    {
        size = v & 0x1FFFFFFFu;
        f = (v >> 29) & 3u;
    }
};

template <typename RecT>
__device__ __forceinline__ RecT pack_record(uint32_t a, uint32_t b);
template <>
__device__ __forceinline__ uint32_t pack_record<uint32_t>(uint32_t a, uint32_t b) { return a | (b << 16); }
template <>
__device__ __forceinline__ uint64_t pack_record<uint64_t>(uint32_t a, uint32_t b) { return (uint64_t)a | ((uint64_t)b << 32); }

__device__ __forceinline__ void unpack_record(uint32_t r, uint32_t &a, uint32_t &b)
{
    a = r & 0xFFFFu;
    b = r >> 16;
}
__device__ __forceinline__ void unpack_record(uint64_t r, uint32_t &a, uint32_t &b)
{
    a = (uint32_t)r;
    b = (uint32_t)(r >> 32);
}

// The HBM-state loop of find2, from codes cx, cy already read (rx, ry:
// whether they are root codes): path halving until both are roots.
template <class St>
__device__ __forceinline__ void follow2(const St &st, uint32_t &x, uint32_t &cx, bool rx,
                                        uint32_t &y, uint32_t &cy, bool ry)
{
    while (!(rx && ry)) {
        const uint32_t gx = rx ? cx : st.get(cx);
        const uint32_t gy = ry ? cy : st.get(cy);
        if (!rx) {
            if (st.is_root(gx)) {
                x = cx;
                cx = gx;
                rx = true;
            } else {
                st.set(x, gx);
                x = gx;
                cx = st.get(x);
                rx = st.is_root(cx);
            }
        }
        if (!ry) {
            if (st.is_root(gy)) {
                y = cy;
                cy = gy;
                ry = true;
            } else {
                st.set(y, gy);
                y = gy;
                cy = st.get(y);
                ry = st.is_root(cy);
            }
        }
    }
}

// Two interleaved finds with path halving.  Concurrent halving by other
// lanes of the wave only ever replaces a parent by an ancestor, so every
// lane still reaches its true root.  With the state in LDS each pass is
// branch-free: both grandparents are read together, then (after the
// halving stores) both new nodes' codes, so a pass is two LDS round trips
// and about half the instructions of a branch per element (which also
// waits for each read in turn).  A finished element re-reads its own node,
// which costs nothing in LDS; in HBM those reads would be traffic, so the
// global state keeps the branches.
template <class St>
__device__ __forceinline__ void find2(const St &st, uint32_t &x, uint32_t &cx,
                                      uint32_t &y, uint32_t &cy)
{
    cx = st.get(x);
    cy = st.get(y);
    bool rx = st.is_root(cx), ry = st.is_root(cy);
    if constexpr (std::is_same<St, LdsState>::value) {
        while (!(rx && ry)) {
            const uint32_t gx = st.get(rx ? x : cx);
            const uint32_t gy = st.get(ry ? y : cy);
            const bool gxr = st.is_root(gx), gyr = st.is_root(gy);
            if (!rx && !gxr)
                st.set(x, gx);
            if (!ry && !gyr)
                st.set(y, gy);
            x = rx ? x : (gxr ? cx : gx);
            y = ry ? y : (gyr ? cy : gy);
            const uint32_t nx = st.get(x), ny = st.get(y);
            cx = rx ? cx : nx;
            cy = ry ? cy : ny;
            rx = st.is_root(cx);
            ry = st.is_root(cy);
        }
        return;
    }
    follow2(st, x, cx, rx, y, cy, ry);
}

// ---------------------------------------------------------------------------
// Merge-record arithmetic.  WIDE (N > 16384) keeps S3, S4 and the moments in
// 128 bits; otherwise every Sk fits 64 bits (Sk <= N**k <= 2**56).
// ---------------------------------------------------------------------------
template <bool WIDE>
struct Wide {
    using T = typename std::conditional<WIDE, u128, uint64_t>::type;
};

template <bool WIDE>
struct Deltas {
    using MT = typename Wide<WIDE>::T;
    uint32_t dmax; // size of the merged cluster (0: no merge)
    uint32_t cnt;  // 1 for a merge
    uint64_t d2;   // (a+b)^2 - a^2 - b^2
    MT d3, d4;
    template <typename RecT>
    __device__ __forceinline__ explicit Deltas(RecT rec)
    {
        uint32_t a, b;
        unpack_record(rec, a, b);
        const bool m = rec != 0;
        dmax = m ? a + b : 0u;
        cnt = m ? 1u : 0u;
        const uint64_t ab = (uint64_t)a * b;
        const uint64_t aa = (uint64_t)a * a, bb = (uint64_t)b * b;
        d2 = 2u * ab;
        d3 = (MT)3u * (MT)ab * (MT)(a + b);
        d4 = (MT)(2u * ab) * (MT)(2u * aa + 3u * ab + 2u * bb);
    }
};

// Word w of a chunk-state entry: {max cluster size, merges, S2, S3 lo,
// S3 hi, S4 lo, S4 hi, 0}.
template <typename MT>
__device__ __forceinline__ uint64_t state_word(uint32_t w, uint64_t mx,
                                               uint64_t merges, uint64_t s2,
                                               MT s3, MT s4)
{
    switch (w) {
    case 0: return mx;
    case 1: return merges;
    case 2: return s2;
    case 3: return (uint64_t)s3;
    case 4: return (uint64_t)((u128)s3 >> 64);
    case 5: return (uint64_t)s4;
    case 6: return (uint64_t)((u128)s4 >> 64);
    default: return 0u;
    }
}

// Puts a uniform value into one lane (a compare and a select, no branch).
template <typename T>
__device__ __forceinline__ T write_lane(T v, uint32_t lane, T old)
{
    return threadIdx.x == lane ? v : old;
}

// Conflict table for the parallel merge rounds: CLAIM_SLOTS words in LDS,
// indexed by a multiplicative hash of a root's low 24 bits (a full-rate
// 24-bit multiply; a collision only blocks a merge for a round).
#define CLAIM_BITS 10
#define CLAIM_SLOTS (1u << CLAIM_BITS)

__device__ __forceinline__ uint32_t claim_slot(uint32_t root)
{
    return ((root & 0xFFFFFFu) * 0x9E3779u) >> (32 - CLAIM_BITS);
}

// Modified-root stamps of an HBM-state sweep: DIRTY_SLOTS words in LDS,
// indexed by the same kind of hash; a slot holds 1 + the last batch in
// which the resolver changed a root hashing there (0: never).  Collisions
// only make a root look changed.
#define DIRTY_BITS 11
#define DIRTY_SLOTS (1u << DIRTY_BITS)

__device__ __forceinline__ uint32_t dirty_slot(uint32_t root)
{
    return ((root & 0xFFFFFFu) * 0x9E3779u) >> (32 - DIRTY_BITS);
}

// One run on one wavefront.  Bonds are taken 64 at a time (lane = step):
//  1. every lane finds the roots of its bond's endpoints in parallel (path
//     halving); lanes whose roots differ are candidate merges;
//  2. parallel rounds: each candidate claims the table slots of its two
//     roots with an atomic max of (round tag + 63 - lane), so each slot ends
//     up held by the earliest candidate touching it.  A candidate holding
//     both of its slots shares no root with any earlier pending candidate:
//     no earlier merge can change its clusters, so it merges now, in
//     parallel with the other such candidates (their roots are distinct).
//     Star merges: the root the first blocked candidate is blocked on (the
//     hub, near and above the threshold usually the largest cluster's root)
//     takes in, in the same round, every pending candidate touching it, in
//     step order up to the first whose other root an earlier candidate also
//     touches, and up to the first other pending candidate that touches a
//     root the star absorbs; their records follow from an exclusive prefix
//     sum of the absorbed sizes and the spanning step from two flag
//     ballots.  The
//     others re-find their roots and go to the next round; the rounds stop
//     early once one frees at most a single candidate.  Hash collisions
//     only make the tests conservative;
//  3. the candidates left after `rounds` rounds are resolved in step order
//     with uniform control flow: roots re-read from the state (one address
//     for the whole wave, made uniform with readfirstlane) and followed
//     while earlier merges moved them, the union decided on scalar
//     registers, lane 0 writing it back, the record placed with a lane
//     select.
// Per-step output is one merge record; the spanning step is kept scalar.
// With chunk_out, the run's state at the start of every batch of 64 bonds
// (the chunk state of chunk_scan) is written as well: the batch's records
// are in registers at its end, so their deltas are reduced across the wave
// there and no pass over the records is needed.
// SPLIT (two wavefronts per run): wave 1, the finder, runs step 1 for batch
// b + 1 while wave 0, the resolver, runs steps 2 and 3 for batch b; one
// barrier per batch, the roots handed over through LDS.  A find that runs
// beside the merges of batch b may stop at a root linked meanwhile, which is
// still an ancestor: the resolver's find continues from it, usually with a
// single read.  With chunk_out and the state in LDS, the finder folds the
// records of batch b - 1 into the running state, off the resolver's chain;
// with the state in HBM the finder's dependent reads set the pace, so the
// resolver folds its own records in the time it would wait.  With the state
// in HBM the finder also hands over the code (size and sides) of each root
// it found, and the resolver stamps every root it changes with the batch in
// an LDS table (hashed, so a collision only costs a read): a handed root
// whose stamp is older than the batch the finder ran beside has not changed
// since the finder read it, so the resolver takes it without a read.
// Edge sources of the sweep: edge e -> its endpoints (u, v).
// EdgeList reads the graph's edge array.  LatticeEdges computes them for the
// square lattice of side L in the order square_lattice builds (rows r =
// 0..L-1; in each row but the last, node (r, c)'s edge to (r + 1, c), then
// to (r, c + 1) when c < L - 1; the last row's L - 1 edges to the right),
// with node (r, c) stored at r L + c, or, tiled (L a multiple of 8), in
// 4 x 8 tiles of one 128-byte line each: ((r / 4) (L / 8) + c / 8) 32 +
// (r mod 4) 8 + c mod 8, the order of the host's locality relabelling.  No
// memory is read for an edge.
struct EdgeList {
    const int2 *__restrict__ p;
    __device__ __forceinline__ int2 operator()(int32_t e) const { return p[e]; }
};

struct LatticeEdges {
    uint32_t L, d;   // side, edges per row but the last (2 L - 1)
    float inv_d;     // 1 / d, a first guess of the row
    uint32_t tiled;
    __device__ __forceinline__ uint32_t node(uint32_t r, uint32_t c) const
    {
        return tiled ? ((((r >> 2) * (L >> 3) + (c >> 3)) << 5) | ((r & 3u) << 3) | (c & 7u))
                     : r * L + c;
    }
    __device__ __forceinline__ int2 operator()(int32_t e) const
    {
        const uint32_t x = (uint32_t)e;
        uint32_t r = (uint32_t)((float)x * inv_d);
        // the float guess is off by at most one row either way
        if (r * d > x)
            --r;
        else if ((r + 1u) * d <= x)
            ++r;
        const uint32_t q = x - r * d;
        uint32_t r2 = r, c = q, c2 = q + 1u;   // the last row: to the right
        if (r + 1u < L) {
            c = q >> 1;
            if (q & 1u)
                c2 = c + 1u;   // to the right
            else {
                r2 = r + 1u;   // down
                c2 = c;
            }
        }
        return make_int2((int32_t)node(r, c), (int32_t)node(r2, c2));
    }
};

template <class St, typename RecT, bool CHUNKS, bool SPLIT, class ES, bool FOLDER = false>
__device__ __forceinline__ void sweep_run(const St &st, uint32_t N, uint32_t M,
                                          const ES &edges,
                                          const uint8_t *__restrict__ sides,
                                          bool span_first, uint32_t rounds,
                                          uint32_t *__restrict__ claims,
                                          uint32_t *__restrict__ handoff,
                                          uint32_t *__restrict__ dirty,
                                          const int32_t *__restrict__ perm,
                                          RecT *__restrict__ rec_out,
                                          uint32_t *__restrict__ span_out,
                                          uint64_t *__restrict__ chunk_out)
{
    // the sweep's pace is the resolver's dependent instruction chain: let
    // its waves win instruction arbitration over waves of other kernels
    // sharing the SIMD (the reductions of the previous sub-batch run beside
    // it).  The finder keeps the same priority: lowering it measured slower.
    __builtin_amdgcn_s_setprio(3);
    // HBM-state graphs (N > 16384) need 128-bit S3 and S4
    constexpr bool WIDE = std::is_same<St, GlobalState>::value;
    using MT = typename Wide<WIDE>::T;
    constexpr uint32_t THREADS = SPLIT ? (FOLDER ? 192u : 128u) : 64u;
    // who folds the records into the chunk states: with the state in LDS
    // the finder (the resolver sets the pace); with the state in HBM a third
    // wave (FOLDER), or else the resolver
    static_assert(!FOLDER || (SPLIT && CHUNKS && std::is_same<St, GlobalState>::value),
                  "a folder wave only for chunk states with the state in HBM");
    constexpr bool RES_FOLD = SPLIT && WIDE && !FOLDER;
    // HBM state with a finder: the finder hands over the codes of the roots
    // it found as well (after the two root slots), and the resolver reads a
    // root afresh only if, in the batch before (beside which the finder
    // read it), it changed a root hashing to the same stamp slot
    constexpr bool HANDED = SPLIT && WIDE;
    uint32_t *const codes_out = handoff + 256u;
    const uint32_t lane = threadIdx.x % 64u;
    // SPLIT: wave 0 resolves batch b while wave 1 finds the roots of batch
    // b + 1 (a find that runs beside the merges of batch b may stop at a
    // root linked meanwhile; that is still an ancestor, and the resolver's
    // find continues from it)
    const bool resolver = !SPLIT || threadIdx.x < 64u;
    const bool folder = FOLDER && threadIdx.x >= 128u;
    uint64_t cs_mx = 1, cs_merges = 0, cs_s2 = N;
    MT cs_s3 = N, cs_s4 = N;
    for (uint32_t i = threadIdx.x; i < N; i += THREADS)
        st.set(i, st.root(1u, sides ? (uint32_t)sides[i] : 0u));
    for (uint32_t i = threadIdx.x; i < CLAIM_SLOTS; i += THREADS)
        claims[i] = 0u;
    if (HANDED)
        for (uint32_t i = threadIdx.x; i < DIRTY_SLOTS; i += THREADS)
            dirty[i] = 0u;
    __syncthreads();

    uint32_t span_at = PC_NO_SPAN;
    uint32_t tag = 0;

    // one wavefront per run: the permutation two batches ahead and the edge
    // endpoints one batch ahead are in flight while a batch is resolved
    int32_t e_ahead = 0;
    int2 uv_next = make_int2(0, 0);
    if (!SPLIT) {
        e_ahead = (64u + lane < M) ? perm[64u + lane] : 0;
        if (lane < M)
            uv_next = edges(perm[lane]);
    }
    // codes of the next batch's endpoints, read one batch early: parent
    // links only ever move up the tree (a union links a root under a root,
    // halving replaces a parent by its grandparent), so a code read before
    // this batch's merges still names an ancestor and the next find may
    // start there, one dependent read shorter.  A root code is no hint
    // (the node may have been linked since): that find reads afresh.
    uint32_t hint_u = 0xFFFFFFFFu, hint_v = 0xFFFFFFFFu;

    // SPLIT finder: roots of the batch at b0 into handoff slot (b0 / 64) & 1
    // (u roots in words 0..63, v roots in 64..127 of the slot).  Its own
    // pipeline, issued at the start of a batch so the loads complete beside
    // the finds: the permutation three batches ahead, the endpoints two
    // batches ahead, and the codes of the next batch's endpoints (the hints
    // above; with the state in HBM the finder is the slower wave and each
    // hint saves a dependent read)
    int2 f_uv1 = make_int2(0, 0), f_uv2 = make_int2(0, 0);
    int32_t f_e3 = 0;
    uint32_t f_hu = 0xFFFFFFFFu, f_hv = 0xFFFFFFFFu;
    if (SPLIT && !resolver && !folder) {
        if (lane < M)
            f_uv1 = edges(perm[lane]);
        if (64u + lane < M)
            f_uv2 = edges(perm[64u + lane]);
        f_e3 = (128u + lane < M) ? perm[128u + lane] : 0;
    }
    auto find_batch = [&](uint32_t b0) {
        const int2 uv = f_uv1;
        const uint32_t hu = f_hu, hv = f_hv;
        f_uv1 = f_uv2;
        if (b0 + 64u + lane < M) {
            f_hu = st.get((uint32_t)f_uv1.x);
            f_hv = st.get((uint32_t)f_uv1.y);
        }
        f_uv2 = (b0 + 128u + lane < M) ? edges(f_e3) : make_int2(0, 0);
        f_e3 = (b0 + 192u + lane < M) ? perm[b0 + 192u + lane] : 0;
        uint32_t x = st.is_root(hu) ? (uint32_t)uv.x : hu;
        uint32_t y = st.is_root(hv) ? (uint32_t)uv.y : hv;
        uint32_t cx, cy;
        if (HANDED)
            cx = cy = st.root(1u, 0u);   // a root code on idle lanes
        if (b0 + lane < M)
            find2(st, x, cx, y, cy);
        const uint32_t si = ((b0 / 64u) & 1u) * 128u;
        handoff[si + lane] = x;
        handoff[si + 64u + lane] = y;
        if (HANDED) {
            codes_out[si + lane] = cx;
            codes_out[si + 64u + lane] = cy;
        }
    };
    if (SPLIT) {
        if (!resolver && !folder)
            find_batch(0u);
        __syncthreads();
    }

    // SPLIT with chunk states: the resolver hands each batch's records to
    // the wave that folds them (two slots of 64 records after the root
    // slots, after the roots' codes with the state in HBM), which folds them
    // into the running state one batch later, off the resolver's chain
    RecT *rec_slots = (RecT *)(handoff + (HANDED ? 512u : 256u));
    auto fold_batch = [&](RecT rec) {
        const Deltas<WIDE> d(rec);
        cs_mx = umax64(cs_mx, (uint64_t)read_lane(wave_scan_max(d.dmax), 63u));
        cs_merges += read_lane(wave_scan_add(d.cnt), 63u);
        cs_s2 += read_lane(wave_scan_add(d.d2), 63u);
        cs_s3 += read_lane(wave_scan_add(d.d3), 63u);
        cs_s4 += read_lane(wave_scan_add(d.d4), 63u);
    };

    for (uint32_t base = 0; base < M; base += 64u) {
        if (SPLIT && !resolver) {
            if (CHUNKS && !RES_FOLD && (folder || !FOLDER)) {
                // state at the start of this batch: fold the records the
                // resolver finished in the previous iteration
                if (base > 0u)
                    fold_batch(rec_slots[(((base / 64u) - 1u) & 1u) * 64u + lane]);
                if (lane < PC_STATE_WORDS)
                    chunk_out[(size_t)(base / 64u) * PC_STATE_WORDS + lane] =
                        state_word(lane, cs_mx, cs_merges, cs_s2, cs_s3, cs_s4);
            }
            if (!folder && base + 64u < M)
                find_batch(base + 64u);
            __syncthreads();
            continue;
        }
        const uint32_t n = base + lane;
        const bool active = n < M;
        if ((!SPLIT || RES_FOLD) && CHUNKS && lane < PC_STATE_WORDS)
            chunk_out[(size_t)(base / 64u) * PC_STATE_WORDS + lane] =
                state_word(lane, cs_mx, cs_merges, cs_s2, cs_s3, cs_s4);
        uint32_t ru, rv, cu, cv;
        if (SPLIT) {
            // the finder's roots, followed past the previous batch's merges
            const uint32_t si = ((base / 64u) & 1u) * 128u;
            ru = handoff[si + lane];
            rv = handoff[si + 64u + lane];
            if (HANDED) {
                // a root unchanged since the finder read it is still a root
                // with that code: no read
                const uint32_t b = base / 64u;
                cu = codes_out[si + lane];
                cv = codes_out[si + 64u + lane];
                if (dirty[dirty_slot(ru)] >= b)
                    cu = st.get(ru);
                if (dirty[dirty_slot(rv)] >= b)
                    cv = st.get(rv);
                follow2(st, ru, cu, st.is_root(cu), rv, cv, st.is_root(cv));
            } else {
                find2(st, ru, cu, rv, cv);
            }
        } else {
            const int2 uv = uv_next;
            uv_next = (n + 64u < M) ? edges(e_ahead) : make_int2(0, 0);
            e_ahead = (n + 128u < M) ? perm[n + 128u] : 0;
            ru = st.is_root(hint_u) ? (uint32_t)uv.x : hint_u;
            rv = st.is_root(hint_v) ? (uint32_t)uv.y : hint_v;
            find2(st, ru, cu, rv, cv);
            hint_u = st.get((uint32_t)uv_next.x);
            hint_v = st.get((uint32_t)uv_next.y);
        }
        bool cand = active && ru != rv;
        uint64_t pending = __ballot(cand);
        RecT rec = 0;

        if (tag >= 0xFFFFFFFFu - 65u * 64u) {
            // this batch's round tags (at most 64 rounds) could wrap: clear
            // the table and restart
            for (uint32_t i = lane; i < CLAIM_SLOTS; i += 64u)
                claims[i] = 0u;
            tag = 0u;
        }
        for (uint32_t k = 0; k < rounds && pending; ++k) {
            tag += 64u;
            const uint32_t mine = tag + 63u - lane;
            const uint32_t su = claim_slot(ru), sv = claim_slot(rv);
            if (cand) {
                atomicMax(&claims[su], mine);
                atomicMax(&claims[sv], mine);
            }
            const uint32_t held_u = claims[su], held_v = claims[sv];
            const bool own_u = held_u == mine, own_v = held_v == mine;
            bool go = cand && own_u && own_v;
            // star: the pending candidates that touch one hub root (the root
            // the first blocked candidate is blocked on, usually the largest
            // cluster's), in step order up to the first whose other root an
            // earlier candidate also touches, all merge into the hub now
            uint64_t star = 0;
            uint32_t hub = 0;
            const uint64_t blocked = __ballot(cand && !go);
            if (blocked) {
                const uint32_t fl = (uint32_t)__builtin_ctzll(blocked);
                hub = read_lane(own_u ? rv : ru, fl);
                const bool at_u = ru == hub;
                const bool touch = cand && (at_u || rv == hub);
                const uint64_t tm = __ballot(touch);
                const uint64_t bad = tm & ~__ballot(touch && (at_u ? own_v : own_u));
                star = bad ? (tm & ((bad & (0u - bad)) - 1u)) : tm;
                go = go && !touch;
                // a pending candidate outside the star whose root's slot is
                // held by a member touches a root the star absorbs before
                // its step: the star stops before it (the slot holder is
                // the earliest claimant, lane tag + 63 - claims[slot])
                const uint32_t hu = tag + 63u - held_u, hv = tag + 63u - held_v;
                const bool late = cand && !go && !((star >> lane) & 1u) &&
                                  (((star >> (hu & 63u)) & 1u) || ((star >> (hv & 63u)) & 1u));
                const uint64_t lm = __ballot(late);
                if (lm)
                    star &= (lm & (0u - lm)) - 1u;
            }
            const bool in_star = (star >> lane) & 1u;
            const uint64_t freed = __ballot(go) | star;
            if (star) {
                // records in step order: the hub's size grows by the sizes
                // absorbed by the earlier members (exclusive prefix sum)
                const bool at_u = ru == hub;
                const uint32_t x = at_u ? rv : ru;
                uint32_t sx, fx, sh, fh;
                st.decode(at_u ? cv : cu, sx, fx);
                st.decode(at_u ? cu : cv, sh, fh);
                const uint32_t first = (uint32_t)__builtin_ctzll(star);
                const uint32_t s_hub = read_lane(sh, first), f_hub = read_lane(fh, first);
                const uint32_t incl = wave_scan_add(in_star ? sx : 0u);
                const uint32_t total = read_lane(incl, 63u);
                const uint64_t m1 = __ballot(in_star && (fx & 1u));
                const uint64_t m2 = __ballot(in_star && (fx & 2u));
                const uint32_t f_all = f_hub | (m1 ? 1u : 0u) | (m2 ? 2u : 0u);
                if (in_star) {
                    // (size on u's side, size on v's side), as every merge
                    st.set(x, hub);
                    if (HANDED)
                        dirty[dirty_slot(x)] = base / 64u + 1u;
                    const uint32_t sh_now = s_hub + incl - sx;
                    rec = pack_record<RecT>(at_u ? sh_now : sx, at_u ? sx : sh_now);
                }
                if (lane == first) {
                    st.set(hub, st.root(s_hub + total, f_all == 3u ? 0u : f_all));
                    if (HANDED)
                        dirty[dirty_slot(hub)] = base / 64u + 1u;
                }
                if (span_first)
                    span_at = min(span_at, base + first + 1u);
                if (f_all == 3u) {
                    // first member after which both sides are touched
                    const uint32_t l1 = (f_hub & 1u) ? first : (uint32_t)__builtin_ctzll(m1);
                    const uint32_t l2 = (f_hub & 2u) ? first : (uint32_t)__builtin_ctzll(m2);
                    span_at = min(span_at, base + max(l1, l2) + 1u);
                }
            }
            bool spans_here = false;
            if (go) {
                uint32_t sa, fa, sb, fb;
                st.decode(cu, sa, fa);
                st.decode(cv, sb, fb);
                const bool a_big = sa >= sb;
                const uint32_t big = a_big ? ru : rv, small = a_big ? rv : ru;
                uint32_t f = fa | fb;
                st.set(small, big);
                st.set(big, st.root(sa + sb, f == 3u ? 0u : f));
                if (HANDED) {
                    dirty[dirty_slot(small)] = base / 64u + 1u;
                    dirty[dirty_slot(big)] = base / 64u + 1u;
                }
                rec = pack_record<RecT>(sa, sb);
                spans_here = span_first || f == 3u;
            }
            {
                const uint64_t spans = __ballot(spans_here);
                if (spans)
                    span_at = min(span_at, base + (uint32_t)__builtin_ctzll(spans) + 1u);
            }
            cand = cand && !go && !in_star;
            st.sync();
            if (__ballot(cand)) {
                // lanes still pending follow their roots past this round's merges
                if (cand)
                    find2(st, ru, cu, rv, cv);
                cand = cand && ru != rv;
            }
            pending = __ballot(cand);
            // a round that frees at most one candidate is resolving a chain
            // (typically through the largest cluster's root): the in-order
            // remainder does that more cheaply
            if (__builtin_popcountll(freed) <= 1)
                break;
        }

        while (pending) {
            const uint32_t i = (uint32_t)__builtin_ctzll(pending);
            pending &= pending - 1u;
            uint32_t a = read_lane(ru, i), b = read_lane(rv, i);
            uint32_t ca = uniform(st.get(a)), cb = uniform(st.get(b));
            while (!st.is_root(ca)) {
                a = ca;
                ca = uniform(st.get(a));
            }
            while (!st.is_root(cb)) {
                b = cb;
                cb = uniform(st.get(b));
            }
            if (a == b)
                continue; // joined by an earlier bond of this batch
            uint32_t sa, fa, sb, fb;
            st.decode(ca, sa, fa);
            st.decode(cb, sb, fb);
            const bool a_big = sa >= sb;
            const uint32_t big = a_big ? a : b, small = a_big ? b : a;
            uint32_t f = fa | fb;
            if (f == 3u || span_first)
                span_at = min(span_at, base + i + 1u);
            if (f == 3u)
                f = 0u;
            if (lane == 0) {
                st.set(small, big);
                st.set(big, st.root(sa + sb, f));
                if (HANDED) {
                    dirty[dirty_slot(small)] = base / 64u + 1u;
                    dirty[dirty_slot(big)] = base / 64u + 1u;
                }
            }
            rec = write_lane(pack_record<RecT>(sa, sb), i, rec);
        }
        if (active)
            rec_out[n] = rec;
        if (SPLIT && CHUNKS && !RES_FOLD)
            rec_slots[((base / 64u) & 1u) * 64u + lane] = rec;
        else if (CHUNKS)
            fold_batch(rec);
        if (!SPLIT)
            __syncthreads();
        else {
            st.sync();
            __syncthreads();
        }
    }
    if (threadIdx.x == 0)
        *span_out = span_at;
}

// Words of LDS before the 16-bit codes: the claim table, then (SPLIT) the
// handoff: two slots of 128 roots, then two slots of 64 records (up to 64
// bits each; with the state in HBM, two slots of the roots' 128 codes).
#define SWEEP_HANDOFF_WORDS 512u
template <bool SPLIT>
constexpr uint32_t sweep_lds_words() { return CLAIM_SLOTS + (SPLIT ? SWEEP_HANDOFF_WORDS : 0u); }

template <typename RecT, bool CHUNKS, bool SPLIT, class ES>
__global__ __launch_bounds__(SPLIT ? 128 : 64) void sweep_lds_kernel(
    uint32_t N, uint32_t M, ES edges,
    const uint8_t *__restrict__ sides, int span_first,
    const int32_t *__restrict__ perm, int64_t perm_stride,
    RecT *__restrict__ records, int64_t rec_stride,
    uint32_t *__restrict__ span_step, uint32_t rounds,
    uint64_t *__restrict__ chunk_state)
{
    // claim table first (word aligned), the handoff slots, then the codes
    extern __shared__ uint32_t lds_words[];
    const uint32_t r = blockIdx.x;
    const size_t cstride = (size_t)((M + 63u) / 64u) * PC_STATE_WORDS;
    const uint32_t shift = N > 1u ? 32u - (uint32_t)__builtin_clz(N - 1u) : 0u;
    const LdsState st{(uint16_t *)(lds_words + sweep_lds_words<SPLIT>()), 1u << shift, shift};
    sweep_run<LdsState, RecT, CHUNKS, SPLIT, ES>(
        st, N, M, edges, sides, span_first != 0, rounds, lds_words,
        lds_words + CLAIM_SLOTS, nullptr, perm + (int64_t)r * perm_stride,
        records + (int64_t)r * rec_stride, span_step + r,
        CHUNKS ? chunk_state + r * cstride : nullptr);
}

template <typename RecT, bool CHUNKS, bool SPLIT, class ES, bool FOLDER = false>
__global__ __launch_bounds__(SPLIT ? (FOLDER ? 192 : 128) : 64) void sweep_global_kernel(
    uint32_t N, uint32_t M, uint32_t R, ES edges,
    const uint8_t *__restrict__ sides, int span_first,
    const int32_t *__restrict__ perm, int64_t perm_stride,
    RecT *__restrict__ records, int64_t rec_stride,
    uint32_t *__restrict__ span_step, uint32_t *__restrict__ workspace,
    uint32_t rounds, uint64_t *__restrict__ chunk_state)
{
    // the handoff: root slots, the roots' codes, then (FOLDER) the records
    __shared__ uint32_t claims[CLAIM_SLOTS];
    __shared__ uint32_t handoff[SPLIT ? SWEEP_HANDOFF_WORDS + (FOLDER ? 256u : 0u) : 1u];
    __shared__ uint32_t dirty[SPLIT ? DIRTY_SLOTS : 1u];
    const size_t cstride = (size_t)((M + 63u) / 64u) * PC_STATE_WORDS;
    const GlobalState st{workspace + (size_t)blockIdx.x * N};
    for (uint32_t r = blockIdx.x; r < R; r += gridDim.x) {
        sweep_run<GlobalState, RecT, CHUNKS, SPLIT, ES, FOLDER>(
            st, N, M, edges, sides, span_first != 0, rounds, claims, handoff,
            dirty, perm + (int64_t)r * perm_stride, records + (int64_t)r * rec_stride,
            span_step + r, CHUNKS ? chunk_state + r * cstride : nullptr);
        __syncthreads();
    }
}

// Persistent sweep over several graphs (the lattices of a finite-size
// study) in one launch: a graph table, every graph's runs numbered in table
// order (graph g owns queue items first .. first + runs - 1), and a queue
// head in device memory.  Each workgroup takes the next item with one
// atomic, sweeps that run with the graph's state kind (LDS codes, or 32-bit
// codes in its own HBM slot), and repeats until the queue is empty, so long
// and short runs pack onto the CUs with no per-graph launch or stream.  The
// table lists the largest graphs first (their runs are the longest).  Every
// sweep folds the chunk states in the finder wave.
struct MultiGraph {
    const int2 *edges;
    LatticeEdges lattice;   // used when lattice.L != 0 (edges unused)
    const uint8_t *sides;
    const int32_t *perm;
    void *records;
    uint32_t *span_step;
    uint64_t *chunk_state;
    int64_t perm_stride, rec_stride;
    uint32_t N, M, runs, first, kind, span_first, rounds, shift;
};
#define MULTI_LDS 0u       // LDS codes, 32-bit records
#define MULTI_HBM32 1u     // HBM codes, 32-bit records
#define MULTI_HBM64 2u     // HBM codes, 64-bit records

__global__ void zero_u32_kernel(uint32_t *p) { *p = 0u; }

__global__ __launch_bounds__(128) void sweep_multi_kernel(
    const MultiGraph *__restrict__ graphs, uint32_t num_graphs, uint32_t total,
    uint32_t *__restrict__ head, uint32_t *__restrict__ workspace,
    uint64_t slot_words)
{
    extern __shared__ uint32_t lds_words[];
    __shared__ uint32_t item;
    uint32_t *claims = lds_words, *handoff = lds_words + CLAIM_SLOTS;
    // after the handoff: an LDS graph's codes, or an HBM graph's root stamps
    uint16_t *codes = (uint16_t *)(lds_words + sweep_lds_words<true>());
    uint32_t *dirty = lds_words + sweep_lds_words<true>();
    for (;;) {
        if (threadIdx.x == 0u)
            item = atomicAdd(head, 1u);
        __syncthreads();
        const uint32_t q = item;
        __syncthreads();
        if (q >= total)
            return;
        uint32_t g = 0;
        while (g + 1u < num_graphs && q >= graphs[g + 1u].first)
            ++g;
        const MultiGraph &G = graphs[g];
        const uint32_t r = q - G.first;
        const size_t cstride = (size_t)((G.M + 63u) / 64u) * PC_STATE_WORDS;
        uint64_t *cs = G.chunk_state + r * cstride;
        const int32_t *perm = G.perm + (int64_t)r * G.perm_stride;
        auto run = [&](const auto &edges) {
            if (G.kind == MULTI_LDS) {
                const LdsState st{codes, 1u << G.shift, G.shift};
                sweep_run<LdsState, uint32_t, true, true>(
                    st, G.N, G.M, edges, G.sides, G.span_first != 0u, G.rounds,
                    claims, handoff, nullptr, perm,
                    (uint32_t *)G.records + (int64_t)r * G.rec_stride,
                    G.span_step + r, cs);
            } else {
                const GlobalState st{workspace + (size_t)blockIdx.x * slot_words};
                if (G.kind == MULTI_HBM32)
                    sweep_run<GlobalState, uint32_t, true, true>(
                        st, G.N, G.M, edges, G.sides, G.span_first != 0u, G.rounds,
                        claims, handoff, dirty, perm,
                        (uint32_t *)G.records + (int64_t)r * G.rec_stride, G.span_step + r, cs);
                else
                    sweep_run<GlobalState, uint64_t, true, true>(
                        st, G.N, G.M, edges, G.sides, G.span_first != 0u, G.rounds,
                        claims, handoff, dirty, perm,
                        (uint64_t *)G.records + (int64_t)r * G.rec_stride, G.span_step + r, cs);
            }
        };
        if (G.lattice.L)
            run(G.lattice);
        else
            run(EdgeList{G.edges});
        __syncthreads();
    }
}

// Running per-run state (after some number of bonds).
template <bool WIDE>
struct RunState {
    using MT = typename Wide<WIDE>::T;
    uint64_t mx, merges, s2;
    MT s3, s4;
};

template <bool WIDE>
__device__ __forceinline__ RunState<WIDE> load_state(const uint64_t *cs)
{
    RunState<WIDE> st;
    st.mx = cs[0];
    st.merges = cs[1];
    st.s2 = cs[2];
    if constexpr (WIDE) {
        st.s3 = (u128)cs[3] | ((u128)cs[4] << 64);
        st.s4 = (u128)cs[5] | ((u128)cs[6] << 64);
    } else {
        st.s3 = cs[3];
        st.s4 = cs[5];
    }
    return st;
}

// State after each lane's step: chunk-start state plus the inclusive scan
// of the deltas of the chunk's records.
template <bool WIDE>
__device__ __forceinline__ RunState<WIDE> advance(const RunState<WIDE> &st0,
                                                  const Deltas<WIDE> &d)
{
    RunState<WIDE> st;
    st.mx = umax64(st0.mx, (uint64_t)wave_scan_max(d.dmax));
    st.merges = st0.merges + wave_scan_add(d.cnt);
    st.s2 = st0.s2 + wave_scan_add(d.d2);
    st.s3 = st0.s3 + wave_scan_add(d.d3);
    st.s4 = st0.s4 + wave_scan_add(d.d4);
    return st;
}

// The seven observables from a state: x[0] max, x[1] spanning, x[2..6]
// moments k = 0..4 of all clusters but one largest.
template <bool WIDE>
__device__ __forceinline__ void observables(const RunState<WIDE> &st,
                                            uint32_t N, bool span,
                                            typename Wide<WIDE>::T x[7])
{
    using MT = typename Wide<WIDE>::T;
    const uint64_t m = st.mx;
    x[0] = (MT)m;
    x[1] = (MT)(span ? 1u : 0u);
    x[2] = (MT)((uint64_t)N - st.merges - 1u);
    x[3] = (MT)((uint64_t)N - m);
    const MT m2 = (MT)(m * m);
    x[4] = (MT)st.s2 - m2;
    x[5] = st.s3 - m2 * (MT)m;
    x[6] = st.s4 - m2 * m2;
}

// ---------------------------------------------------------------------------
// chunk_scan in two passes.  (A) chunk_delta: every (run, chunk) in
// parallel, the totals of the chunk's deltas {max merged size, merges, dS2,
// dS3, dS4} reduced across the wave into that chunk's entry.  (B)
// chunk_prefix: one wavefront per run turns the entries in place into the
// state at the start of each chunk, 64 chunks per step (scans across lanes,
// a carry between steps).
// ---------------------------------------------------------------------------
#define CHUNK_DELTA_PER_BLOCK 8u

template <typename RecT, bool WIDE>
__global__ __launch_bounds__(64) void chunk_delta_kernel(
    uint32_t M, const RecT *__restrict__ records, int64_t rec_stride,
    uint64_t *__restrict__ chunk_state)
{
    // blocks: run-major, CHUNK_DELTA_PER_BLOCK chunks each
    const uint32_t C = (M + 63u) / 64u;
    const uint32_t per_run = (C + CHUNK_DELTA_PER_BLOCK - 1u) / CHUNK_DELTA_PER_BLOCK;
    const uint32_t r = blockIdx.x / per_run, lane = threadIdx.x;
    const RecT *rec_r = records + (int64_t)r * rec_stride;
    uint64_t *out = chunk_state + (size_t)r * C * PC_STATE_WORDS;
    const uint32_t c0 = (blockIdx.x % per_run) * CHUNK_DELTA_PER_BLOCK;
    const uint32_t c1 = min(C, c0 + CHUNK_DELTA_PER_BLOCK);
    for (uint32_t c = c0; c < c1; ++c) {
        const uint32_t q = c * 64u + lane;
        const Deltas<WIDE> d(q < M ? rec_r[q] : (RecT)0);
        const uint64_t dmax = read_lane(wave_scan_max(d.dmax), 63u);
        const uint64_t cnt = read_lane(wave_scan_add(d.cnt), 63u);
        const uint64_t d2 = read_lane(wave_scan_add(d.d2), 63u);
        const auto d3 = read_lane(wave_scan_add(d.d3), 63u);
        const auto d4 = read_lane(wave_scan_add(d.d4), 63u);
        if (lane < PC_STATE_WORDS)
            out[(size_t)c * PC_STATE_WORDS + lane] = state_word(lane, dmax, cnt, d2, d3, d4);
    }
}

template <bool WIDE>
__global__ __launch_bounds__(64) void chunk_prefix_kernel(
    uint32_t N, uint32_t M, uint64_t *__restrict__ chunk_state)
{
    using MT = typename Wide<WIDE>::T;
    const uint32_t r = blockIdx.x, lane = threadIdx.x;
    const uint32_t C = (M + 63u) / 64u;
    uint64_t *st = chunk_state + (size_t)r * C * PC_STATE_WORDS;
    // state at the start of the current group of 64 chunks
    uint64_t mx = 1, merges = 0, s2 = N;
    MT s3 = N, s4 = N;
    for (uint32_t c0 = 0; c0 < C; c0 += 64u) {
        const uint32_t c = c0 + lane;
        uint64_t w[PC_STATE_WORDS] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (c < C)
            for (uint32_t k = 0; k < PC_STATE_WORDS; ++k)
                w[k] = st[(size_t)c * PC_STATE_WORDS + k];
        const uint32_t dmax = (uint32_t)w[0];
        const uint64_t cnt = w[1], d2 = w[2];
        MT d3 = (MT)w[3], d4 = (MT)w[5];
        if constexpr (WIDE) {
            d3 |= (MT)w[4] << 64;
            d4 |= (MT)w[6] << 64;
        }
        const uint32_t imax = wave_scan_max(dmax);
        const uint64_t icnt = wave_scan_add(cnt), id2 = wave_scan_add(d2);
        const MT id3 = wave_scan_add(d3), id4 = wave_scan_add(d4);
        // exclusive prefix: the sums minus the own chunk, the maximum of the
        // chunks before (the previous lane's inclusive maximum)
        uint32_t emax = (uint32_t)__shfl_up((int)imax, 1, 64);
        if (lane == 0)
            emax = 0u;
        if (c < C)
            for (uint32_t k = 0; k < PC_STATE_WORDS; ++k)
                st[(size_t)c * PC_STATE_WORDS + k] = state_word(
                    k, umax64(mx, emax), merges + icnt - cnt, s2 + id2 - d2,
                    s3 + id3 - d3, s4 + id4 - d4);
        mx = umax64(mx, read_lane(imax, 63u));
        merges += read_lane(icnt, 63u);
        s2 += read_lane(id2, 63u);
        s3 += read_lane(id3, 63u);
        s4 += read_lane(id4, 63u);
    }
}

// ---------------------------------------------------------------------------
// Microcanonical sums: block (chunk c, run group g), lane = step.  Each lane
// accumulates its step's sums over the group's runs in registers; groups
// combine through multi-word atomics.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void square_u192(u128 x, uint64_t &w0, uint64_t &w1,
                                            uint64_t &w2)
{
    const uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
    const u128 ll = (u128)lo * lo;
    const u128 lh2 = ((u128)lo * hi) << 1; // < 2^129 only if hi >= 2^64: hi < 2^32 here
    const u128 hh = (u128)hi * hi;
    w0 = (uint64_t)ll;
    const u128 mid = (ll >> 64) + (u128)(uint64_t)lh2;
    w1 = (uint64_t)mid;
    w2 = (uint64_t)(mid >> 64) + (uint64_t)(lh2 >> 64) + (uint64_t)hh;
}

template <typename RecT, bool WIDE>
__global__ __launch_bounds__(64) void micro_sums_kernel(
    uint32_t N, uint32_t M, uint32_t R, uint32_t runs_per_group,
    const RecT *__restrict__ records, int64_t rec_stride,
    const uint32_t *__restrict__ span_step,
    const uint64_t *__restrict__ chunk_state, uint64_t *__restrict__ sum1,
    uint64_t *__restrict__ sum2)
{
    using MT = typename Wide<WIDE>::T;
    const uint32_t c = blockIdx.x, g = blockIdx.y, lane = threadIdx.x;
    const uint32_t C = (M + 63u) / 64u;
    const uint64_t S = (uint64_t)M + 1u;
    const uint32_t q = c * 64u + lane, s = q + 1u;
    const bool valid = q < M;
    const uint32_t r0 = g * runs_per_group;
    const uint32_t r1 = min(R, r0 + runs_per_group);

    u128 acc1[7];
    u128 acc2[7];
    uint64_t acc2_top[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        acc1[k] = 0;
        acc2[k] = 0;
        acc2_top[k] = 0;
    }

    if constexpr (!WIDE) {
        // N <= 2^14 and at most 2^16 runs per group bound every accumulator:
        // max, #clusters - 1 and N - max are below 2^14 (squares below 2^28,
        // group sums of squares below 2^44); the spanning flag is 0 or 1;
        // moment 2 is below N^2 / 2 (square below 2^54), moment 3 below
        // N^3 / 4 (group sum below 2^56), moment 4 below N^4 / 8.  So the
        // narrow columns use 32- and 64-bit arithmetic, exact all the same.
        uint32_t n1[4] = {0, 0, 0, 0};  // sum x: max, span, #clusters - 1, N - max
        uint64_t n2[4] = {0, 0, 0, 0};  // sum x^2 of the same
        uint64_t m1_2 = 0, m1_3 = 0;    // sum x: moments 2 and 3
        u128 m1_4 = 0, m2_2 = 0, m2_3 = 0, m2_4 = 0;
        for (uint32_t r = r0; r < r1; ++r) {
            const RunState<false> st0 = load_state<false>(
                chunk_state + ((size_t)r * C + c) * PC_STATE_WORDS);
            const RecT rec = valid ? records[(int64_t)r * rec_stride + q] : (RecT)0;
            const RunState<false> st = advance<false>(st0, Deltas<false>(rec));
            uint64_t x[7];
            observables<false>(st, N, s >= span_step[r], x);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t v = (uint32_t)x[k];
                n1[k] += v;
                n2[k] += (uint64_t)(v * v);
            }
            const uint32_t v4 = (uint32_t)x[4];
            m1_2 += v4;
            m2_2 += (u128)((uint64_t)v4 * v4);
            m1_3 += x[5];
            m2_3 += (u128)x[5] * x[5];
            m1_4 += x[6];
            m2_4 += (u128)x[6] * x[6];
        }
        // column order of the sums: max, span, moments 0..4
        const int col[4] = {0, 1, 2, 3};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            acc1[col[k]] = n1[k];
            acc2[col[k]] = n2[k];
        }
        acc1[4] = m1_2;
        acc2[4] = m2_2;
        acc1[5] = m1_3;
        acc2[5] = m2_3;
        acc1[6] = m1_4;
        acc2[6] = m2_4;
    } else {
        // N <= 2^21 (the sums' limit) and at most 2^16 runs per group: S3 is
        // below N^3 <= 2^63 (its deltas and scan stay in 64 bits), max,
        // #clusters - 1 and N - max are below 2^21 (group sums of squares
        // below 2^58), moment 2 below 2^42 (its square below 2^84), moment 3
        // below 2^63 (square below 2^126) and moment 4 below 2^84 (square
        // below 2^168).  Only moments 3 and 4 need 192-bit sums of squares.
        uint64_t n1[5] = {0, 0, 0, 0, 0}; // sum x: max, span, #clusters - 1, N - max, moment 2
        uint64_t n2[4] = {0, 0, 0, 0};    // sum x^2 of the first four
        u128 m2_2 = 0, m1_3 = 0, m1_4 = 0;
        u128 m2_3 = 0, m2_4 = 0;          // low 128 bits of sum x^2, moments 3 and 4
        uint64_t t2_3 = 0, t2_4 = 0;      // their top words
        for (uint32_t r = r0; r < r1; ++r) {
            const uint64_t *cs = chunk_state + ((size_t)r * C + c) * PC_STATE_WORDS;
            const uint64_t mx0 = cs[0], merges0 = cs[1], s2_0 = cs[2], s3_0 = cs[3];
            const u128 s4_0 = (u128)cs[5] | ((u128)cs[6] << 64);
            const RecT rec = valid ? records[(int64_t)r * rec_stride + q] : (RecT)0;
            uint32_t a, b;
            unpack_record(rec, a, b);
            const bool merge = rec != 0;
            const uint64_t ab = (uint64_t)a * b;
            const uint64_t aa = (uint64_t)a * a, bb = (uint64_t)b * b;
            const uint32_t dmax = merge ? a + b : 0u;
            const uint64_t d2 = 2u * ab;
            const uint64_t d3 = 3u * ab * (uint64_t)(a + b);
            const u128 d4 = (u128)(2u * ab) * (u128)(2u * aa + 3u * ab + 2u * bb);
            const uint64_t mx = umax64(mx0, (uint64_t)wave_scan_max(dmax));
            const uint64_t merges = merges0 + wave_scan_add(merge ? 1u : 0u);
            const uint64_t s2 = s2_0 + wave_scan_add(d2);
            const uint64_t s3 = s3_0 + wave_scan_add(d3);
            const u128 s4 = s4_0 + wave_scan_add(d4);
            const uint64_t m2 = mx * mx;
            const uint64_t x[4] = {mx, s >= span_step[r] ? 1u : 0u,
                                   (uint64_t)N - merges - 1u, (uint64_t)N - mx};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                n1[k] += x[k];
                n2[k] += x[k] * x[k];
            }
            const uint64_t x4 = s2 - m2, x5 = s3 - m2 * mx;
            const u128 x6 = s4 - (u128)m2 * m2;
            n1[4] += x4;
            m2_2 += (u128)x4 * x4;
            m1_3 += x5;
            {
                const u128 sq = (u128)x5 * x5, before = m2_3;
                m2_3 += sq;
                t2_3 += m2_3 < before ? 1u : 0u;
            }
            m1_4 += x6;
            {
                uint64_t w0, w1, w2;
                square_u192(x6, w0, w1, w2);
                const u128 before = m2_4;
                m2_4 += (u128)w0 | ((u128)w1 << 64);
                t2_4 += w2 + (m2_4 < before ? 1u : 0u);
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            acc1[k] = n1[k];
            acc2[k] = n2[k];
        }
        acc1[4] = n1[4];
        acc2[4] = m2_2;
        acc1[5] = m1_3;
        acc2[5] = m2_3;
        acc2_top[5] = t2_3;
        acc1[6] = m1_4;
        acc2[6] = m2_4;
        acc2_top[6] = t2_4;
    }

    if (valid) {
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            atomic_add_u128(sum1 + ((uint64_t)k * S + s) * 2u, acc1[k]);
            atomic_add_u192(sum2 + ((uint64_t)k * S + s) * 3u,
                            (uint64_t)acc2[k], (uint64_t)(acc2[k] >> 64),
                            acc2_top[k]);
        }
    }
    if (c == 0 && lane == 0 && r1 > r0) {
        // step 0: max 1, no spanning, every moment N - 1
        const u128 n = r1 - r0;
        const u128 x0[7] = {1, 0, N - 1u, N - 1u, N - 1u, N - 1u, N - 1u};
        for (int k = 0; k < 7; ++k) {
            atomic_add_u128(sum1 + (uint64_t)k * S * 2u, n * x0[k]);
            const u128 sq = n * x0[k] * x0[k];
            atomic_add_u192(sum2 + (uint64_t)k * S * 3u, (uint64_t)sq,
                            (uint64_t)(sq >> 64), 0u);
        }
    }
}

// Step 0 alone (graphs without edges): max 1, no spanning, moments N - 1.
__global__ void step0_sums_kernel(uint32_t N, uint32_t R, uint64_t S,
                                  uint64_t *__restrict__ sum1,
                                  uint64_t *__restrict__ sum2)
{
    const u128 n = R;
    const u128 x0[7] = {1, 0, N - 1u, N - 1u, N - 1u, N - 1u, N - 1u};
    for (int k = 0; k < 7; ++k) {
        atomic_add_u128(sum1 + (uint64_t)k * S * 2u, n * x0[k]);
        const u128 sq = n * x0[k] * x0[k];
        atomic_add_u192(sum2 + (uint64_t)k * S * 3u, (uint64_t)sq,
                        (uint64_t)(sq >> 64), 0u);
    }
}

__global__ void expand_step0_kernel(uint32_t N, uint32_t R, uint64_t S,
                                    uint32_t *__restrict__ out_max,
                                    uint8_t *__restrict__ out_span,
                                    uint64_t *__restrict__ out_mom)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R)
        return;
    const uint64_t row = (uint64_t)r * S;
    out_max[row] = 1u;
    out_span[row] = 0u;
    for (int k = 0; k < 5; ++k)
        out_mom[row * 5u + k] = (uint64_t)N - 1u;
}

__global__ void fill_zero_kernel(uint64_t *__restrict__ p, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// ---------------------------------------------------------------------------
// Per-run traces: block (chunk c, run r), lane = step.
// ---------------------------------------------------------------------------
template <typename RecT, bool WIDE>
__global__ __launch_bounds__(64) void expand_kernel(
    uint32_t N, uint32_t M, const RecT *__restrict__ records,
    int64_t rec_stride, const uint32_t *__restrict__ span_step,
    const uint64_t *__restrict__ chunk_state, uint32_t *__restrict__ out_max,
    uint8_t *__restrict__ out_span, uint64_t *__restrict__ out_mom)
{
    using MT = typename Wide<WIDE>::T;
    const uint32_t c = blockIdx.x, r = blockIdx.y, lane = threadIdx.x;
    const uint32_t C = (M + 63u) / 64u;
    const uint64_t S = (uint64_t)M + 1u;
    const uint32_t q = c * 64u + lane, s = q + 1u;
    const bool valid = q < M;
    const RunState<WIDE> st0 = load_state<WIDE>(
        chunk_state + ((size_t)r * C + c) * PC_STATE_WORDS);
    const RecT rec = valid ? records[(int64_t)r * rec_stride + q] : (RecT)0;
    const RunState<WIDE> st = advance<WIDE>(st0, Deltas<WIDE>(rec));
    MT x[7];
    observables<WIDE>(st, N, s >= span_step[r], x);
    const uint64_t row = (uint64_t)r * S + s;
    if (valid) {
        out_max[row] = (uint32_t)x[0];
        out_span[row] = (uint8_t)x[1];
#pragma unroll
        for (int k = 0; k < 5; ++k)
            out_mom[row * 5u + k] = (uint64_t)x[2 + k];
    }
    if (c == 0 && lane == 0) {
        const uint64_t row0 = (uint64_t)r * S;
        out_max[row0] = 1u;
        out_span[row0] = 0u;
        for (int k = 0; k < 5; ++k)
            out_mom[row0 * 5u + k] = (uint64_t)N - 1u;
    }
}


// ---------------------------------------------------------------------------
// Binomial weights B(n, k; p) = C(n, k) p^k (1-p)^(n-k), the convolution
// factors of percolate/percolate.py:1067-1109, evaluated per element in FP64
// with Loader's saddle-point form (C. Loader, "Fast and Accurate Computation
// of Binomial Probabilities", 2000):
//   ln B = stirlerr(n) - stirlerr(k) - stirlerr(n-k) - bd0(k, np)
//          - bd0(n-k, nq) - ln(2 pi k (n-k) / n) / 2
// with stirlerr(m) = ln m! - ln(sqrt(2 pi m) (m/e)^m) and the deviance
// bd0(x, m) = x ln(x/m) + m - x computed without cancellation.  Only the
// window |k - np| <= 12 sigma + 48 is kept: by Bernstein's inequality each
// tail outside it holds at most exp(-72) of the mass for every sigma.  The
// window is then normalised to sum 1 as the reference normalises its array.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double stirlerr(double n)
{
    const double S0 = 1.0 / 12.0, S1 = 1.0 / 360.0, S2 = 1.0 / 1260.0,
                 S3 = 1.0 / 1680.0, S4 = 1.0 / 1188.0;
    if (n <= 15.0)
        return kStirlerr[(int)n];
    const double nn = n * n;
    return (S0 - (S1 - (S2 - (S3 - S4 / nn) / nn) / nn) / nn) / n;
}

__device__ __forceinline__ double bd0(double x, double np)
{
    if (fabs(x - np) < 0.1 * (x + np)) {
        double v = (x - np) / (x + np);
        double s = (x - np) * v;
        double ej = 2.0 * x * v;
        const double v2 = v * v;
        for (int j = 1; j < 1000; ++j) {
            ej *= v2;
            const double s1 = s + ej / (double)(2 * j + 1);
            if (s1 == s)
                return s1;
            s = s1;
        }
        return s;
    }
    return x * log(x / np) + np - x;
}

__device__ __forceinline__ double binom_pmf(double x, double n, double p)
{
    const double q = 1.0 - p;
    if (p == 0.0)
        return x == 0.0 ? 1.0 : 0.0;
    if (q == 0.0)
        return x == n ? 1.0 : 0.0;
    if (x == 0.0) {
        if (n == 0.0)
            return 1.0;
        const double lc = p < 0.1 ? -bd0(n, n * q) - n * p : n * log(q);
        return exp(lc);
    }
    if (x == n) {
        const double lc = q < 0.1 ? -bd0(n, n * p) - n * q : n * log(p);
        return exp(lc);
    }
    const double lc = stirlerr(n) - stirlerr(x) - stirlerr(n - x) -
                      bd0(x, n * p) - bd0(n - x, n * q);
    const double lf = 1.8378770664093454836 /* ln 2 pi */ + log(x) +
                      log1p(-x / n);
    return exp(lc - 0.5 * lf);
}

__global__ __launch_bounds__(256) void binom_weights_kernel(
    uint32_t n, const double *__restrict__ ps, const int64_t *__restrict__ start,
    const int64_t *__restrict__ len, int64_t kmax, double *__restrict__ w)
{
    const uint32_t pi = blockIdx.y;
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= kmax)
        return;
    double v = 0.0;
    if (k < len[pi])
        v = binom_pmf((double)(start[pi] + k), (double)n, ps[pi]);
    w[(int64_t)pi * kmax + k] = v;
}

// one block per p: divide the window by its sum
__global__ __launch_bounds__(256) void binom_normalize_kernel(
    int64_t kmax, const int64_t *__restrict__ len, double *__restrict__ w)
{
    __shared__ double part[256];
    double *row = w + (int64_t)blockIdx.x * kmax;
    const int64_t L = len[blockIdx.x];
    double acc = 0.0;
    for (int64_t k = threadIdx.x; k < L; k += 256)
        acc += row[k];
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            part[threadIdx.x] += part[threadIdx.x + off];
        __syncthreads();
    }
    const double total = part[0];
    if (total > 0.0)
        for (int64_t k = threadIdx.x; k < L; k += 256)
            row[k] /= total;
}

// ---------------------------------------------------------------------------
// Canonical contraction Q[p][c] = sum_k w[p][k] X[start_p + k][c] in FP64 on
// the vector ALUs: one block per p, threads stride over the window.
// ---------------------------------------------------------------------------
#define CONTRACT_MAX_COLS 32
#define PC_CONTRACT_SLICES 16

template <int C>
__global__ __launch_bounds__(256) void contract_f64_kernel(
    const double *__restrict__ x, int64_t x_stride, const int64_t *__restrict__ start,
    const int64_t *__restrict__ len, const double *__restrict__ w, int64_t kmax,
    double *__restrict__ out, int32_t cols)
{
    __shared__ double part[256 / 64][C];
    const uint32_t pi = blockIdx.x;
    const double *wr = w + (int64_t)pi * kmax;
    const double *xr = x + start[pi] * x_stride;
    const int64_t L = len[pi];
    double acc[C];
#pragma unroll
    for (int c = 0; c < C; ++c)
        acc[c] = 0.0;
    for (int64_t k = threadIdx.x; k < L; k += 256) {
        const double wk = wr[k];
        const double *row = xr + k * x_stride;
#pragma unroll
        for (int c = 0; c < C; ++c)
            if (c < cols)
                acc[c] = fma(wk, row[c], acc[c]);
    }
    // wave reduction, then across the four waves
#pragma unroll
    for (int c = 0; c < C; ++c) {
        double v = acc[c];
        for (int off = 32; off > 0; off >>= 1)
            v += __shfl_xor(v, off, 64);
        acc[c] = v;
    }
    const uint32_t wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0)
#pragma unroll
        for (int c = 0; c < C; ++c)
            part[wave][c] = acc[c];
    __syncthreads();
    if (threadIdx.x < (uint32_t)cols) {
        const int c = threadIdx.x;
        out[(int64_t)pi * cols + c] =
            part[0][c] + part[1][c] + part[2][c] + part[3][c];
    }
}


// ---------------------------------------------------------------------------
// Canonical contraction on the matrix cores.  One wavefront computes a tile
// of 16 p-values x 16 columns over one slice of the steps (the union of the
// 16 windows is split across blockIdx.z for parallelism); slices write
// partial sums that contract_reduce_kernel adds in a fixed order, so the
// result does not depend on scheduling.
//   FP64:   v_mfma_f64_16x16x4f64, FP64 in, FP64 accumulate.
//   FP32:   v_mfma_f32_16x16x4f32 on FP32 inputs; the FP32 accumulator is
//           added into FP64 every 8 MFMAs (bounded FP32 accumulation).
//   BF16X3: FP32 inputs split into three bf16 terms hi + mid + lo, the six
//           leading products (lo*hi, hi*lo, mid*mid, mid*hi, hi*mid,
//           hi*hi) on v_mfma_f32_16x16x32_bf16, FP32 accumulator added into
//           FP64 every 8 steps.
// Lane layout of the 16x16xK MFMAs: A row i = lane % 16 holds K/4
// consecutive k starting at (lane / 16) * K/4; B column j = lane % 16 the
// same k; register r of the output holds D[4 * (lane / 16) + r][lane % 16]
// for the FP32 accumulators and D[lane / 16 + 4 r][lane % 16] for FP64.
// ---------------------------------------------------------------------------
// Row tile of workgroup b out of T: workgroups are dealt to the 8 XCDs in
// turn, so give each XCD a contiguous run of tiles; neighbouring p-values
// read the same rows of X and then share that XCD's L2.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t T)
{
    const uint32_t x = b % 8u, idx = b / 8u;
    const uint32_t base = T / 8u, extra = T % 8u;
    return x * base + min(x, extra) + idx;
}

typedef double v4f64 __attribute__((ext_vector_type(4)));
typedef float v4f32 __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf16 __attribute__((ext_vector_type(8)));

#define MFMA_FLUSH 8

struct TileCtx {
    uint32_t P, C;
    int64_t x_stride, kmax;
    const int64_t *start, *len;
};

__device__ __forceinline__ int64_t tile_lo(const TileCtx &t, uint32_t p0)
{
    int64_t lo = INT64_MAX;
    for (uint32_t i = 0; i < 16 && p0 + i < t.P; ++i)
        lo = min(lo, t.start[p0 + i]);
    return lo;
}
__device__ __forceinline__ int64_t tile_hi(const TileCtx &t, uint32_t p0)
{
    int64_t hi = 0;
    for (uint32_t i = 0; i < 16 && p0 + i < t.P; ++i)
        hi = max(hi, t.start[p0 + i] + t.len[p0 + i]);
    return hi;
}

template <typename T>
__device__ __forceinline__ T load_w(const T *__restrict__ w, const TileCtx &t,
                                    uint32_t p, int64_t g)
{
    if (p >= t.P)
        return (T)0;
    const int64_t k = g - t.start[p];
    return (k >= 0 && k < t.len[p]) ? w[(int64_t)p * t.kmax + k] : (T)0;
}

template <typename T>
__device__ __forceinline__ T load_x(const T *__restrict__ x, const TileCtx &t,
                                    int64_t g, uint32_t c, int64_t S)
{
    return (c < t.C && g < S) ? x[g * t.x_stride + c] : (T)0;
}

__device__ __forceinline__ void split3(float v, __bf16 &hi, __bf16 &mid, __bf16 &lo)
{
    hi = (__bf16)v;
    const float r1 = v - (float)hi;
    mid = (__bf16)r1;
    const float r2 = r1 - (float)mid;
    lo = (__bf16)r2;
}

template <int PREC>
__global__ __launch_bounds__(64) void contract_mfma_kernel(
    const void *__restrict__ xv, int64_t S, const void *__restrict__ wv,
    TileCtx t, uint32_t slices, double *__restrict__ partial)
{
    const uint32_t lane = threadIdx.x;
    const uint32_t p0 = blockIdx.x * 16u, c0 = blockIdx.y * 16u;
    const uint32_t slice = blockIdx.z;
    const int64_t lo = tile_lo(t, p0), hi = tile_hi(t, p0);
    constexpr int64_t KSTEP = (PREC == PC_PREC_BF16X3_MFMA) ? 32 : 4;
    const int64_t steps = hi > lo ? (hi - lo + KSTEP - 1) / KSTEP : 0;
    const int64_t per = (steps + slices - 1) / slices;
    const int64_t g_begin = lo + (int64_t)slice * per * KSTEP;
    const int64_t g_end = min(hi, g_begin + per * KSTEP);
    const uint32_t i = lane % 16u, q = lane / 16u;
    const uint32_t p = p0 + i, c = c0 + i;

    double out[4] = {0.0, 0.0, 0.0, 0.0};
    if (PREC == PC_PREC_FP64_MFMA) {
        const double *w = (const double *)wv, *x = (const double *)xv;
        v4f64 acc = {0.0, 0.0, 0.0, 0.0};
        for (int64_t g = g_begin; g < g_end; g += 4) {
            const double a = load_w(w, t, p, g + q);
            const double b = load_x(x, t, g + q, c, S);
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
        }
        for (int r = 0; r < 4; ++r)
            out[r] = acc[r];
    } else if (PREC == PC_PREC_FP32_MFMA) {
        const float *w = (const float *)wv, *x = (const float *)xv;
        v4f32 acc = {0.f, 0.f, 0.f, 0.f};
        int n = 0;
        for (int64_t g = g_begin; g < g_end; g += 4) {
            const float a = load_w(w, t, p, g + q);
            const float b = load_x(x, t, g + q, c, S);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
            if (++n == MFMA_FLUSH) {
                for (int r = 0; r < 4; ++r)
                    out[r] += (double)acc[r];
                acc = v4f32{0.f, 0.f, 0.f, 0.f};
                n = 0;
            }
        }
        for (int r = 0; r < 4; ++r)
            out[r] += (double)acc[r];
    } else {
        const float *w = (const float *)wv, *x = (const float *)xv;
        v4f32 acc = {0.f, 0.f, 0.f, 0.f};
        int n = 0;
        for (int64_t g = g_begin; g < g_end; g += 32) {
            v8bf16 ah, am, al, bh, bm, bl;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int64_t gk = g + 8 * q + e;
                __bf16 h, m, l;
                split3(load_w(w, t, p, gk), h, m, l);
                ah[e] = h; am[e] = m; al[e] = l;
                split3(load_x(x, t, gk, c, S), h, m, l);
                bh[e] = h; bm[e] = m; bl[e] = l;
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bm, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bh, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bm, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc, 0, 0, 0);
            if (++n == MFMA_FLUSH) {
                for (int r = 0; r < 4; ++r)
                    out[r] += (double)acc[r];
                acc = v4f32{0.f, 0.f, 0.f, 0.f};
                n = 0;
            }
        }
        for (int r = 0; r < 4; ++r)
            out[r] += (double)acc[r];
    }
    // output rows: D[4q + r][i] for the FP32-accumulating MFMAs, rows
    // interleaved as D[q + 4r][i] for v_mfma_f64_16x16x4f64
    for (int r = 0; r < 4; ++r) {
        const uint32_t row = (PREC == PC_PREC_FP64_MFMA) ? q + 4u * (uint32_t)r
                                                         : 4u * q + (uint32_t)r;
        const uint32_t pr = p0 + row, cc = c0 + i;
        if (pr < t.P && cc < t.C)
            partial[((int64_t)slice * t.P + pr) * t.C + cc] = out[r];
    }
}

// Wide contraction on the FP64 matrix cores, for many columns (the per-run
// canonical statistics of a whole chunk of runs at once): a GEMM of the
// banded weights with X.  Workgroup of 4 wavefronts on a tile of 64 p x TC
// columns; wave w takes the p rows 16w..16w+15 and all TC columns, TC / 16
// v_mfma_f64_16x16x4f64 accumulators sharing the weight operand.  The steps
// of the union of the tile's 64 windows advance 16 at a time: the 16 x TC
// block of X is staged through LDS (double-buffered, one barrier per step;
// rows padded by 4 doubles so the four lane groups of a read fall on
// different banks), each lane's 4 weights are consecutive steps.  Tiles are
// dealt XCD-contiguous in column-major order, so the p tiles that read the
// same columns of X share an L2.  Partial sums of K-slices (blockIdx.y) go
// to `dst` + slice * P * C and contract_reduce_kernel adds them in order.
#define WIDE_TILE 64
#ifndef WIDE_KSTEP
#define WIDE_KSTEP 16
#endif

// TC columns per tile (64 or 128): TC / 16 accumulators per wave; LDS rows
// of TC + 4 doubles; WIDE_KSTEP (16 or 32) steps per barrier
template <int TC>
__global__ __launch_bounds__(256) void contract_wide_f64_kernel(
    const double *__restrict__ x, int64_t S, const double *__restrict__ w,
    TileCtx t, uint32_t tiles_p, uint32_t slices, double *__restrict__ dst)
{
    constexpr int KS = WIDE_KSTEP, H = KS / 16;
    constexpr int STRIDE = TC + 4, NJ = TC / 16, E = TC / 16;
    __shared__ double xs[2][KS * STRIDE];
    __shared__ int64_t range[2];
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t p0 = (tile % tiles_p) * WIDE_TILE;
    const uint32_t c0 = (tile / tiles_p) * TC;
    const uint32_t slice = blockIdx.y;
    const uint32_t tid = threadIdx.x, wave = tid / 64u, lane = tid % 64u;
    const uint32_t i = lane % 16u, q = lane / 16u;
    if (wave == 0) {
        // union of the windows of the tile's p
        int64_t lo = INT64_MAX, hi = 0;
        if (p0 + lane < t.P) {
            lo = t.start[p0 + lane];
            hi = lo + t.len[p0 + lane];
        }
        for (int off = 32; off > 0; off >>= 1) {
            lo = min(lo, (int64_t)__shfl_xor((long long)lo, off, 64));
            hi = max(hi, (int64_t)__shfl_xor((long long)hi, off, 64));
        }
        if (lane == 0) {
            range[0] = lo;
            range[1] = hi;
        }
    }
    __syncthreads();
    const int64_t lo = range[0], hi = range[1];
    const int64_t steps = hi > lo ? (hi - lo + KS - 1) / KS : 0;
    const int64_t per = (steps + slices - 1) / slices;
    const int64_t g_begin = lo + (int64_t)slice * per * KS;
    const int64_t g_end = min(hi, g_begin + per * KS);

    const uint32_t p = p0 + wave * 16u + i;
    const bool prow = p < t.P;
    const int64_t wstart = prow ? t.start[p] : 0, wlen = prow ? t.len[p] : 0;
    const double *wrow = w + (int64_t)(prow ? p : 0) * t.kmax;
    // X loader: thread tid fills rows tid / 16 + 16 h, columns
    // E (tid % 16) .. + E-1
    const uint32_t lr = tid / 16u, lc = (uint32_t)E * (tid % 16u);

    v4f64 acc[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j)
        acc[j] = v4f64{0.0, 0.0, 0.0, 0.0};
    double a_cur[4 * H], a_nxt[4 * H], b_nxt[E * H];
    // weights of steps g + 16h + 4q .. + 3 for row p (zero outside its
    // window); X rows past the slice end are zero, so a partial last step
    // adds nothing that belongs to the next slice
    auto load_a = [&](int64_t g, double *a) {
#pragma unroll
        for (int h = 0; h < H; ++h)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t k = g + 16 * h + 4 * (int64_t)q + e - wstart;
                a[4 * h + e] = (k >= 0 && k < wlen) ? wrow[k] : 0.0;
            }
    };
    auto load_b = [&](int64_t g, double *b) {
#pragma unroll
        for (int h = 0; h < H; ++h) {
            const int64_t row = g + 16 * h + lr;
            const bool ok = row < g_end && row < S;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const uint32_t c = c0 + lc + (uint32_t)e;
                b[E * h + e] = (ok && c < t.C) ? x[row * t.x_stride + c] : 0.0;
            }
        }
    };
    auto store_b = [&](int buf, const double *b) {
#pragma unroll
        for (int h = 0; h < H; ++h)
#pragma unroll
            for (int e = 0; e < E; ++e)
                xs[buf][(16 * h + lr) * STRIDE + lc + e] = b[E * h + e];
    };
    if (g_begin < g_end) {
        load_b(g_begin, b_nxt);
        load_a(g_begin, a_cur);
        store_b(0, b_nxt);
        __syncthreads();
        int buf = 0;
        for (int64_t g = g_begin; g < g_end; g += KS) {
            const bool more = g + KS < g_end;
            if (more) {
                load_b(g + KS, b_nxt);
                load_a(g + KS, a_nxt);
            }
            const double *xb = xs[buf];
#pragma unroll
            for (int h = 0; h < H; ++h)
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    // step g + 16h + 4q + kk: weight a_cur[4h + kk], X row
                    // 16h + 4q + kk
                    const double *xr =
                        xb + (16u * h + 4u * q + (uint32_t)kk) * STRIDE + i;
#pragma unroll
                    for (int j = 0; j < NJ; ++j)
                        acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                            a_cur[4 * h + kk], xr[16 * j], acc[j], 0, 0, 0);
                }
            if (more) {
                store_b(buf ^ 1, b_nxt);
#pragma unroll
                for (int e = 0; e < 4 * H; ++e)
                    a_cur[e] = a_nxt[e];
            }
            __syncthreads();
            buf ^= 1;
        }
    }
    // register r of accumulator j holds D[q + 4r][i] of column block j
    double *out = dst + (int64_t)slice * t.P * t.C;
#pragma unroll
    for (int j = 0; j < NJ; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint32_t row = p0 + wave * 16u + q + 4u * (uint32_t)r;
            const uint32_t col = c0 + 16u * (uint32_t)j + i;
            if (row < t.P && col < t.C)
                out[(int64_t)row * t.C + col] = acc[j][r];
        }
}

__global__ __launch_bounds__(256) void contract_reduce_kernel(
    const double *__restrict__ partial, uint32_t slices, int64_t n,
    double *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
    double acc = 0.0;
    for (uint32_t s = 0; s < slices; ++s)
        acc += partial[(int64_t)s * n + i];
    out[i] = acc;
}

__global__ __launch_bounds__(256) void f64_to_f32_kernel(
    const double *__restrict__ in, int64_t n, float *__restrict__ out)
{
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256)
        out[i] = (float)in[i];
}

// ---------------------------------------------------------------------------
// Canonical averages with the binomial weights generated inside the
// contraction (no weight matrix in HBM):
//   out[p][c] = sum_k B(n, k; p) X[k][c] / sum_k B(n, k; p)
// over the window of p.  The normaliser is the contraction of a virtual
// all-ones column C, so it goes through the same arithmetic as the columns.
//
// Tiles: one wavefront = 16 p-values x 16 columns (column C is the ones
// column) over one of `slices` pieces of the union of the 16 windows.  The
// four lane groups q = lane / 16 each walk a contiguous quarter of the piece,
// one k per MFMA k-position, so every lane's weights form a sequence of
// consecutive k: one direct evaluation (Loader's form, binom_pmf) where the
// sequence enters the window, then the exact ratio
//   B(n, k; p) / B(n, k - 1; p) = (n - k + 1) / k * p / (1 - p),
// with a fresh direct evaluation every FUSED_REANCHOR steps, so the
// relative error stays below about 1e-13.  The MFMA's k positions only have
// to agree between the A (weights) and B (columns) operands, which they do
// because lane (i, q) supplies both for the same k.
//
// SPLIT_BF16 (Ozaki-style exact split): per window of 64 k (16 consecutive
// k in each lane group), a weight is scaled by 2^(32 - E), E per row
// (max w < 2^E over the row's 64 weights of the window), and a column value
// by 2^(32 - F), F per column (max |x| < 2^F), rounded to 32-bit integers and
// cut into four bytes, most significant first.  Bytes are exact in bf16 and
// their products (< 2^16) exact in FP32; the 13 byte-pair products with
// s + t <= 4 run on v_mfma_f32_16x16x32_bf16, one FP32 accumulator per level
// s + t.  Within a window each level sums at most 4 x 64 products of at most
// 255 x 255, below 2^24, so the FP32 sums are exact integers; they are added
// into FP64 scaled by 2^(E + F - 16 - 8 (s + t)).  Scales per window keep
// the far tails of the weights (and small column values) at full relative
// precision; the dropped levels and the rounding of the 32-bit integers
// leave a relative error near 1e-10 for non-negative columns.
// ---------------------------------------------------------------------------
#define FUSED_REANCHOR 1024
#define FUSED_F64_UNROLL 8
#ifndef FUSED_F64_AHEAD
#define FUSED_F64_AHEAD 1
#endif

struct FusedCtx {
    const double *x;
    int64_t x_stride;
    const double *ps;
    const int64_t *start, *len;
    uint32_t n, P, C;
};

// Direct evaluations in the walks are rare (one per lane and range, and one
// per FUSED_REANCHOR steps): kept out of line so the contraction loops keep
// their registers (112 instead of 160 VGPRs, 4 waves per SIMD).
__device__ __attribute__((noinline)) double binom_pmf_outline(double x, double n,
                                                               double p)
{
    return binom_pmf(x, n, p);
}

// Per-lane walk over consecutive k of one row's weights.
struct BinomWalk {
    double n, p, pq, w, anchor;
    int64_t ka, kend; // in-window k of this lane's range: [ka, kend)
};

__device__ __forceinline__ void walk_init(BinomWalk &b, const FusedCtx &f,
                                          uint32_t row, int64_t k0, int64_t k1)
{
    b.n = (double)f.n;
    b.p = 0.5;
    int64_t st = 0, ln = 0;
    if (row < f.P) {
        b.p = f.ps[row];
        st = f.start[row];
        ln = f.len[row];
        // the degenerate distributions hold all their mass at one k
        if (b.p <= 0.0) {
            st = 0;
            ln = 1;
        } else if (b.p >= 1.0) {
            st = (int64_t)f.n;
            ln = 1;
        }
    }
    b.pq = b.p / (1.0 - b.p);
    b.ka = max(k0, st);
    b.kend = min(k1, st + ln);
    b.anchor = b.ka < b.kend ? binom_pmf_outline((double)b.ka, b.n, b.p) : 0.0;
    b.w = 0.0;
}

// Weight at k (k advances by one per call) from the weight at k - 1 and
// ratio = (n - k + 1) / k; `fresh` replaces the ratio step by a direct
// evaluation (a uniform branch every FUSED_REANCHOR steps).
__device__ __forceinline__ double walk_step(BinomWalk &b, int64_t k, bool fresh,
                                            double ratio)
{
    double v;
    if (fresh && k > b.ka && k < b.kend)
        v = binom_pmf_outline((double)k, b.n, b.p);
    else
        v = b.w * ratio * b.pq;
    v = (k == b.ka) ? b.anchor : v;
    b.w = v;
    return (k >= b.ka && k < b.kend) ? v : 0.0;
}

__device__ __forceinline__ double fused_x(const FusedCtx &f, int64_t k, uint32_t c)
{
    if (c < f.C)
        return f.x[k * f.x_stride + c];
    return c == f.C ? 1.0 : 0.0;
}

// K-range of this wave: the union of the tile's windows, cut into `slices`
// pieces of a multiple of `quant` steps; returns false if the piece is empty.
__device__ __forceinline__ bool fused_range(const FusedCtx &f, uint32_t p0,
                                            uint32_t slices, uint32_t slice,
                                            int64_t quant, int64_t &g_begin,
                                            int64_t &g_end, int64_t &per)
{
    int64_t lo = INT64_MAX, hi = 0;
    for (uint32_t r = 0; r < 16 && p0 + r < f.P; ++r) {
        const double p = f.ps[p0 + r];
        int64_t st = f.start[p0 + r], ln = f.len[p0 + r];
        if (p <= 0.0) {
            st = 0;
            ln = 1;
        } else if (p >= 1.0) {
            st = (int64_t)f.n;
            ln = 1;
        }
        lo = min(lo, st);
        hi = max(hi, st + ln);
    }
    if (hi <= lo)
        return false;
    per = (hi - lo + slices - 1) / slices;
    per = (per + quant - 1) / quant * quant;
    g_begin = lo + (int64_t)slice * per;
    g_end = min(hi, g_begin + per);
    return g_begin < g_end;
}

__global__ __launch_bounds__(64) void canonical_fused_f64_kernel(
    FusedCtx f, uint32_t slices, double *__restrict__ partial)
{
    const uint32_t lane = threadIdx.x, i = lane % 16u, q = lane / 16u;
    const uint32_t p0 = xcd_tile(blockIdx.x, gridDim.x) * 16u;
    const uint32_t c0 = blockIdx.y * 16u, slice = blockIdx.z;
    const uint32_t C1 = f.C + 1u, c = c0 + i;
    int64_t g_begin = 0, g_end = 0, per = 0;
    v4f64 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
    if (fused_range(f, p0, slices, slice, 4 * FUSED_F64_UNROLL, g_begin, g_end, per)) {
        const int64_t Lq = per / 4;
        const int64_t kq0 = g_begin + (int64_t)q * Lq;
        const int64_t kq1 = min(kq0 + Lq, g_end);
        BinomWalk b;
        walk_init(b, f, p0 + i, kq0, kq1);
        // the ratio (n - k + 1) / k depends on k only, shared by the 16 rows
        // of a lane group: each lane divides once per 16 steps, for step
        // j0 + lane % 16 of group lane / 16, and the steps read it back
        // with a lane shuffle
        double rr = 0.0;
        // columns FUSED_F64_AHEAD chunks ahead (a miss to the MALL or HBM
        // outlasts one chunk of MFMAs several times over)
        double xq[FUSED_F64_AHEAD][FUSED_F64_UNROLL];
#pragma unroll
        for (int d = 0; d < FUSED_F64_AHEAD; ++d)
#pragma unroll
            for (int u = 0; u < FUSED_F64_UNROLL; ++u) {
                const int64_t k = kq0 + d * FUSED_F64_UNROLL + u;
                xq[d][u] = k < kq1 ? fused_x(f, k, c) : 0.0;
            }
        for (int64_t j0 = 0; j0 < Lq; j0 += FUSED_F64_UNROLL) {
            double a[FUSED_F64_UNROLL], x[FUSED_F64_UNROLL];
#pragma unroll
            for (int u = 0; u < FUSED_F64_UNROLL; ++u) {
                x[u] = xq[0][u];
#pragma unroll
                for (int d = 0; d + 1 < FUSED_F64_AHEAD; ++d)
                    xq[d][u] = xq[d + 1][u];
                const int64_t k = kq0 + j0 + FUSED_F64_AHEAD * FUSED_F64_UNROLL + u;
                xq[FUSED_F64_AHEAD - 1][u] = k < kq1 ? fused_x(f, k, c) : 0.0;
            }
            if ((j0 & 15) == 0) {
                const int64_t kr = g_begin + (int64_t)(lane / 16u) * Lq + j0 + (lane % 16u);
                rr = (b.n - (double)kr + 1.0) / (double)kr;
            }
            const bool fresh = j0 > 0 && j0 % FUSED_REANCHOR == 0;
#pragma unroll
            for (int u = 0; u < FUSED_F64_UNROLL; ++u) {
                const double ratio = __shfl(rr, (int)(q * 16u) + (int)(j0 & 15) + u, 64);
                a[u] = walk_step(b, kq0 + j0 + u, fresh && u == 0, ratio);
            }
#pragma unroll
            for (int u = 0; u < FUSED_F64_UNROLL; u += 2) {
                acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], x[u], acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u + 1], x[u + 1], acc1, 0, 0, 0);
            }
        }
    }
    // v_mfma_f64_16x16x4f64: register r holds D[q + 4r][i]
    for (int r = 0; r < 4; ++r) {
        const uint32_t row = p0 + q + 4u * (uint32_t)r;
        if (row < f.P && c < C1)
            partial[((int64_t)slice * f.P + row) * C1 + c] = acc0[r] + acc1[r];
    }
}

// |v| 2^e rounded to a 32-bit integer (e makes |v| 2^e < 2^32)
__device__ __forceinline__ uint32_t split_u32(double v, int e)
{
    return (uint32_t)fmin(ldexp(fabs(v), e) + 0.5, 4294967295.0);
}

// byte `s` (0 = most significant) of u as a bf16, negated when neg
__device__ __forceinline__ __bf16 byte_bf16(uint32_t u, int s, bool neg)
{
    const float v = (float)((u >> (24 - 8 * s)) & 255u);
    return (__bf16)(neg ? -v : v);
}

__global__ __launch_bounds__(64) void canonical_fused_split_kernel(
    FusedCtx f, uint32_t slices, double *__restrict__ partial)
{
    const uint32_t lane = threadIdx.x, i = lane % 16u, q = lane / 16u;
    const uint32_t p0 = xcd_tile(blockIdx.x, gridDim.x) * 16u;
    const uint32_t c0 = blockIdx.y * 16u, slice = blockIdx.z;
    const uint32_t C1 = f.C + 1u, c = c0 + i;
    int64_t g_begin = 0, g_end = 0, per = 0;
    double out[4] = {0.0, 0.0, 0.0, 0.0};
    if (fused_range(f, p0, slices, slice, 64, g_begin, g_end, per)) {
        const int64_t Lq = per / 4;
        const int64_t kq0 = g_begin + (int64_t)q * Lq;
        const int64_t kq1 = min(kq0 + Lq, g_end);
        BinomWalk b;
        walk_init(b, f, p0 + i, kq0, kq1);
        for (int64_t j0 = 0; j0 < Lq; j0 += 16) {
            double x[16], w[16];
            double m = 0.0;
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int64_t k = kq0 + j0 + e;
                x[e] = k < kq1 ? fused_x(f, k, c) : 0.0;
                m = fmax(m, fabs(x[e]));
            }
            const bool fresh = j0 > 0 && j0 % FUSED_REANCHOR == 0;
            // one division per lane per window: the ratio (n - k + 1) / k
            // for step j0 + lane % 16 of group lane / 16, read back by the
            // 16 rows of each group with a lane shuffle
            const int64_t kr = g_begin + (int64_t)(lane / 16u) * Lq + j0 + (lane % 16u);
            const double rr = (b.n - (double)kr + 1.0) / (double)kr;
            double wm = 0.0;
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const double ratio = __shfl(rr, (int)(q * 16u) + e, 64);
                w[e] = walk_step(b, kq0 + j0 + e, fresh && e == 0, ratio);
                wm = fmax(wm, w[e]);
            }
            // scales of column i and row i over the window: lanes i, i+16,
            // i+32, i+48 hold its four runs of 16 k
            m = fmax(m, __shfl_xor(m, 16, 64));
            m = fmax(m, __shfl_xor(m, 32, 64));
            wm = fmax(wm, __shfl_xor(wm, 16, 64));
            wm = fmax(wm, __shfl_xor(wm, 32, 64));
            const int F = m > 0.0 ? ilogb(m) + 1 : 0;
            const int Ew = wm > 0.0 ? ilogb(wm) + 1 : 0;
            // output register r holds row 4q + r: its scale from lane 4q + r
            int Er[4];
#pragma unroll
            for (int r = 0; r < 4; ++r)
                Er[r] = __shfl(Ew, (int)(4u * q) + r, 64);
            v4f32 acc[5];
#pragma unroll
            for (int l = 0; l < 5; ++l)
                acc[l] = v4f32{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                v8bf16 A[4], B[4];
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const uint32_t uw = split_u32(w[8 * h + e], 32 - Ew);
                    const uint32_t ux = split_u32(x[8 * h + e], 32 - F);
                    const bool neg = x[8 * h + e] < 0.0;
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        A[s][e] = byte_bf16(uw, s, false);
                        B[s][e] = byte_bf16(ux, s, neg);
                    }
                }
#pragma unroll
                for (int s = 0; s < 4; ++s)
#pragma unroll
                    for (int t = 0; t < 4; ++t)
                        if (s + t <= 4)
                            acc[s + t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                A[s], B[t], acc[s + t], 0, 0, 0);
            }
#pragma unroll
            for (int l = 0; l < 5; ++l)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    out[r] += ldexp((double)acc[l][r], Er[r] + F - 16 - 8 * l);
        }
    }
    // FP32-accumulating MFMA: register r holds D[4q + r][i]
    for (int r = 0; r < 4; ++r) {
        const uint32_t row = p0 + 4u * q + (uint32_t)r;
        if (row < f.P && c < C1)
            partial[((int64_t)slice * f.P + row) * C1 + c] = out[r];
    }
}

// out[p][c] = sum over slices of partial[.][p][c] / the same sum of the
// ones column, in a fixed order
__global__ __launch_bounds__(256) void fused_reduce_kernel(
    const double *__restrict__ partial, uint32_t slices, uint32_t P, uint32_t C,
    double *__restrict__ out)
{
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (int64_t)P * C)
        return;
    const int64_t p = idx / C, c = idx % C;
    const uint32_t C1 = C + 1u;
    double num = 0.0, den = 0.0;
    for (uint32_t s = 0; s < slices; ++s) {
        const double *row = partial + ((int64_t)s * P + p) * C1;
        num += row[c];
        den += row[C];
    }
    out[idx] = num / den;
}


// ---------------------------------------------------------------------------
// Per-run observables as FP64 columns over a window of steps, the input of
// the per-run canonical statistics (percolate/hpc.py:496-512):
// out[(s - g_lo) * (R * C) + r * C + c] for steps s in [g_lo, g_lo + g_len),
// columns [has_spanning_cluster if spanning,] max_cluster_size,
// moments[0..4] (moments as their uint64 values, like the reference's
// structured array).
// ---------------------------------------------------------------------------
#ifndef WT_RUNS
#define WT_RUNS 4
#endif

template <typename RecT, bool WIDE>
__global__ __launch_bounds__(64) void window_traces_kernel(
    uint32_t N, uint32_t M, uint32_t R, const RecT *__restrict__ records,
    int64_t rec_stride, const uint32_t *__restrict__ span_step,
    const uint64_t *__restrict__ chunk_state, int64_t g_lo, int64_t g_len,
    int spanning, double *__restrict__ out)
{
    // Block (chunk of 64 steps, group of WT_RUNS runs): lane = step while
    // the observables are computed, then the tile goes out row by row, the
    // group's WT_RUNS * cols values of a step contiguous in the output (one
    // coalesced store per step instead of 64 scattered ones per run)
    using MT = typename Wide<WIDE>::T;
    __shared__ double tile[64][WT_RUNS * 7];
    const uint32_t lane = threadIdx.x;
    const uint32_t r0 = blockIdx.y * WT_RUNS;
    const uint32_t nr = min((uint32_t)WT_RUNS, R - r0);
    const uint32_t C = (M + 63u) / 64u;
    // chunk index of this block within the steps 1..M (state after c*64)
    const int64_t first_chunk = g_lo > 0 ? (g_lo - 1) / 64 : 0;
    const int64_t c = first_chunk + blockIdx.x;
    const int cols = spanning ? 7 : 6;
    const int64_t width = (int64_t)R * cols;
    if (c >= (int64_t)C) {
        // steps past M cannot exist; only step 0 needs the no-edge case
        return;
    }
    const int64_t q = c * 64 + lane; // record index, step q + 1
    const bool valid = q < (int64_t)M;
    for (uint32_t j = 0; j < nr; ++j) {
        const uint32_t r = r0 + j;
        const RunState<WIDE> st0 = load_state<WIDE>(
            chunk_state + ((size_t)r * C + (size_t)c) * PC_STATE_WORDS);
        const RecT rec = valid ? records[(int64_t)r * rec_stride + q] : (RecT)0;
        const RunState<WIDE> st = advance<WIDE>(st0, Deltas<WIDE>(rec));
        MT x[7];
        observables<WIDE>(st, N, (uint64_t)(q + 1) >= span_step[r], x);
        double *row = tile[lane] + j * cols;
        int k = 0;
        if (spanning)
            row[k++] = (double)(uint64_t)x[1];
        row[k++] = (double)(uint64_t)x[0];
        for (int m = 0; m < 5; ++m)
            row[k++] = (double)(uint64_t)x[2 + m];
    }
    __syncthreads();
    const uint32_t n = nr * (uint32_t)cols;
    for (uint32_t i = 0; i < 64u; ++i) {
        const int64_t step = c * 64 + i + 1;
        if (step > (int64_t)M || step < g_lo || step >= g_lo + g_len)
            continue;
        double *dst = out + (step - g_lo) * width + (int64_t)r0 * cols;
        if (lane < n)
            dst[lane] = tile[i][lane];
    }
    if (c == first_chunk && g_lo == 0 && lane < nr) {
        // step 0: max 1, no spanning, moments N - 1
        double *row = out + (int64_t)(r0 + lane) * cols;
        int k = 0;
        if (spanning)
            row[k++] = 0.0;
        row[k++] = 1.0;
        for (int m = 0; m < 5; ++m)
            row[k++] = (double)((uint64_t)N - 1u);
    }
}

// Mean and sum of squared deviations over runs of q[p][r * cols + c]
// (two passes, fixed order): the (n, mean, M2) of percolate/hpc.py:561-702.
// One workgroup per p reads that p's contiguous row of R * cols values:
// thread t takes runs t, t + 256, ... (cols consecutive values each), and
// the 256 partial sums of a column meet in a fixed LDS tree, so the result
// does not depend on scheduling.
#define RUN_MOMENTS_COLS 8

__device__ __forceinline__ double block_sum_256(double v, double *red)
{
    // fixed-order tree: wave-local butterfly, then the 4 wave totals in order
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    const uint32_t wave = threadIdx.x / 64;
    __syncthreads();
    if (threadIdx.x % 64 == 0)
        red[wave] = v;
    __syncthreads();
    return (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(256) void run_moments_kernel(
    const double *__restrict__ q, uint32_t P, uint32_t R, uint32_t cols,
    double *__restrict__ mean, double *__restrict__ m2)
{
    __shared__ double red[4];
    const uint32_t p = blockIdx.x;
    const double *row = q + (uint64_t)p * R * cols;
    double s[RUN_MOMENTS_COLS];
#pragma unroll
    for (int c = 0; c < RUN_MOMENTS_COLS; ++c)
        s[c] = 0.0;
    for (uint32_t r = threadIdx.x; r < R; r += 256) {
#pragma unroll
        for (int c = 0; c < RUN_MOMENTS_COLS; ++c)
            if (c < (int)cols)
                s[c] += row[(uint64_t)r * cols + c];
    }
    double mu[RUN_MOMENTS_COLS];
#pragma unroll
    for (int c = 0; c < RUN_MOMENTS_COLS; ++c)
        mu[c] = c < (int)cols ? block_sum_256(s[c], red) / (double)R : 0.0;
#pragma unroll
    for (int c = 0; c < RUN_MOMENTS_COLS; ++c)
        s[c] = 0.0;
    for (uint32_t r = threadIdx.x; r < R; r += 256) {
#pragma unroll
        for (int c = 0; c < RUN_MOMENTS_COLS; ++c)
            if (c < (int)cols) {
                const double d = row[(uint64_t)r * cols + c] - mu[c];
                s[c] = fma(d, d, s[c]);
            }
    }
#pragma unroll
    for (int c = 0; c < RUN_MOMENTS_COLS; ++c) {
        if (c < (int)cols) {
            const double d2 = block_sum_256(s[c], red);
            if (threadIdx.x == 0) {
                mean[(uint64_t)p * cols + c] = mu[c];
                m2[(uint64_t)p * cols + c] = d2;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Microcanonical mean and sample standard deviation from the exact sums:
// mean = S1 / R, std = sqrt((R S2 - S1^2) / (R (R - 1))), with the numerator
// formed exactly in 256-bit integers before the single rounding to FP64.
// ---------------------------------------------------------------------------
struct u256 {
    uint64_t w[4];
};

__device__ __forceinline__ u256 mul_u192_u64(const uint64_t a[3], uint64_t b)
{
    u256 r;
    u128 carry = 0;
    for (int i = 0; i < 3; ++i) {
        const u128 t = (u128)a[i] * b + carry;
        r.w[i] = (uint64_t)t;
        carry = t >> 64;
    }
    r.w[3] = (uint64_t)carry;
    return r;
}

__device__ __forceinline__ u256 square_u128(uint64_t lo, uint64_t hi)
{
    u256 r;
    const u128 ll = (u128)lo * lo, lh = (u128)lo * hi, hh = (u128)hi * hi;
    r.w[0] = (uint64_t)ll;
    u128 t = (ll >> 64) + (u128)(uint64_t)lh + (u128)(uint64_t)lh;
    r.w[1] = (uint64_t)t;
    t = (t >> 64) + (lh >> 64) + (lh >> 64) + (u128)(uint64_t)hh;
    r.w[2] = (uint64_t)t;
    t = (t >> 64) + (hh >> 64);
    r.w[3] = (uint64_t)t;
    return r;
}

__device__ __forceinline__ u256 sub_u256(const u256 &a, const u256 &b)
{
    u256 r;
    uint64_t borrow = 0;
    for (int i = 0; i < 4; ++i) {
        const uint64_t d = a.w[i] - b.w[i];
        const uint64_t d2 = d - borrow;
        borrow = (a.w[i] < b.w[i]) || (d < borrow) ? 1u : 0u;
        r.w[i] = d2;
    }
    return r;
}

__device__ __forceinline__ double u256_to_double(const u256 &a)
{
    // top 64 significant bits, rounded once to double, then scaled
    int top = 3;
    while (top > 0 && a.w[top] == 0)
        --top;
    if (top == 0)
        return (double)a.w[0];
    const int lz = __clzll(a.w[top]);
    uint64_t m = lz ? (a.w[top] << lz) | (a.w[top - 1] >> (64 - lz)) : a.w[top];
    // sticky bit for the discarded words
    uint64_t rest = lz ? (a.w[top - 1] << lz) : a.w[top - 1];
    for (int i = top - 2; i >= 0; --i)
        rest |= a.w[i];
    if (rest)
        m |= 1u;
    return ldexp((double)m, 64 * top - lz);
}

__device__ __forceinline__ double u128_div_to_double(uint64_t lo, uint64_t hi,
                                                     uint64_t R)
{
    const u128 v = ((u128)hi << 64) | lo;
    const u128 q = v / R, rem = v % R;
    return (double)q + (double)(uint64_t)rem / (double)R;
}

__global__ __launch_bounds__(256) void micro_moments_kernel(
    uint64_t total, uint64_t R, const uint64_t *__restrict__ sum1,
    const uint64_t *__restrict__ sum2, double *__restrict__ mean,
    double *__restrict__ std_out)
{
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= total)
        return;
    const uint64_t lo = sum1[2 * i], hi = sum1[2 * i + 1];
    mean[i] = u128_div_to_double(lo, hi, R);
    if (R < 2) {
        std_out[i] = __longlong_as_double(0x7FF8000000000000ll); // NaN, as ddof=1
        return;
    }
    const uint64_t s2[3] = {sum2[3 * i], sum2[3 * i + 1], sum2[3 * i + 2]};
    const u256 d = sub_u256(mul_u192_u64(s2, R), square_u128(lo, hi));
    const double num = u256_to_double(d);
    std_out[i] = sqrt(num / ((double)R * (double)(R - 1)));
}

// ---------------------------------------------------------------------------
// Microcanonical statistics with their confidence bounds (the epilogue of
// percolate/percolate.py:557-705) for every step from the exact sums.
// Thread = step s; out[c * S + s] for the 21 columns
//   0..2   max cluster size: mean, lower, upper bound,
//   3..5   spanning: (k + 1) / (R + 2) and the bounds beta[2k], beta[2k + 1]
//          (k = runs with a spanning cluster at s),
//   6 + 3j, 7 + 3j, 8 + 3j   moment j = 0..4: mean, lower, upper bound,
// a bound being mean + (std / sqrt(R)) * t_lo (t_hi) where R > 1 and
// std > 0, the mean otherwise.  Products and sums are rounded one by one
// (no fused multiply-add), as the host's scale * ppf + loc.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void mean_std(const uint64_t *__restrict__ sum1,
                                         const uint64_t *__restrict__ sum2,
                                         uint64_t i, uint64_t R, double &mean,
                                         double &std)
{
    const uint64_t lo = sum1[2 * i], hi = sum1[2 * i + 1];
    mean = u128_div_to_double(lo, hi, R);
    if (R < 2) {
        std = __longlong_as_double(0x7FF8000000000000ll); // NaN, as ddof=1
        return;
    }
    const uint64_t s2[3] = {sum2[3 * i], sum2[3 * i + 1], sum2[3 * i + 2]};
    const u256 d = sub_u256(mul_u192_u64(s2, R), square_u128(lo, hi));
    std = sqrt(u256_to_double(d) / ((double)R * (double)(R - 1)));
}

__global__ __launch_bounds__(256) void micro_ci_kernel(
    uint64_t S, uint64_t R, const uint64_t *__restrict__ sum1,
    const uint64_t *__restrict__ sum2, double t_lo, double t_hi,
    const double *__restrict__ beta, double *__restrict__ out)
{
    // scale * t + mean rounded twice, never fused (HIP contracts by default)
#pragma clang fp contract(off)
    const uint64_t s = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= S)
        return;
    const double sqrt_r = sqrt((double)R);
    // columns of the sums: 0 max, 1 spanning, 2..6 moments 0..4
    const int src[6] = {0, 2, 3, 4, 5, 6};
    const int dst[6] = {0, 6, 9, 12, 15, 18};
    for (int j = 0; j < 6; ++j) {
        double m, sd;
        mean_std(sum1, sum2, (uint64_t)src[j] * S + s, R, m, sd);
        double lo = m, hi = m;
        if (R > 1 && sd != 0.0) {
            const double scale = sd / sqrt_r;
            lo = t_lo * scale + m;
            hi = t_hi * scale + m;
        }
        out[(uint64_t)dst[j] * S + s] = m;
        out[(uint64_t)(dst[j] + 1) * S + s] = lo;
        out[(uint64_t)(dst[j] + 2) * S + s] = hi;
    }
    const uint64_t k = sum1[2 * (S + s)]; // spanning runs (< 2^64)
    out[3 * S + s] = ((double)k + 1.0) / ((double)R + 2.0);
    out[4 * S + s] = beta[2 * k];
    out[5 * S + s] = beta[2 * k + 1];
}

// Finalized canonical averages (percolate/hpc.py:752-834) of P rows of C
// statistics: element i = p * C + c.  Mean and sample standard deviation
// sqrt(M2 / (n - 1)), each divided by div[c] (the number of nodes for the
// per-node statistics, 1 otherwise: x / 1 == x), and the Student-t interval
// q * (std / sqrt(n)) + mean with the caller's per-row quantiles, NaN where
// scipy's ppf refuses its arguments (n - 1 <= 0, scale not > 0, mean NaN).
// out: [4][P][C] = mean, std, lower, upper.
__global__ __launch_bounds__(256) void finalize_kernel(
    uint32_t P, uint32_t C, const uint32_t *__restrict__ runs,
    const double *__restrict__ mean, const double *__restrict__ m2,
    const double *__restrict__ div, const double *__restrict__ q_lo,
    const double *__restrict__ q_hi, double *__restrict__ out)
{
    // every product and sum rounded on its own, as numpy and scipy do
#pragma clang fp contract(off)
    const uint64_t PC = (uint64_t)P * C;
    const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= PC)
        return;
    const uint32_t p = (uint32_t)(i / C), c = (uint32_t)(i % C);
    const double n = (double)runs[p];
    const double dof = n - 1.0;
    const double d = div[c];
    const double mu = mean[i] / d;
    const double sd = sqrt(m2[i] / dof) / d;
    const double scale = sd / sqrt(n);
    const bool ok = dof > 0.0 && scale > 0.0 && mu == mu;
    const double nan = __longlong_as_double(0x7ff8000000000000ll);
    out[i] = mu;
    out[PC + i] = sd;
    out[2 * PC + i] = ok ? q_lo[p] * scale + mu : nan;
    out[3 * PC + i] = ok ? q_hi[p] * scale + mu : nan;
}

// Pairwise (n, mean, M2) merge of Chan, Golub & LeVeque (the reduction of
// percolate/hpc.py:638-702) of P rows of C statistics: element i = p * C + c,
// per-row counts n_a[p], n_b[p].  mean = mean_a + delta * (n_b / n),
// M2 = M2_a + M2_b + delta^2 * (n_a n_b / n), delta = mean_b - mean_a; the
// outputs may alias the a operands.
__global__ __launch_bounds__(256) void merge_moments_kernel(
    uint64_t PC, uint32_t C, const double *__restrict__ n_a,
    const double *mean_a, const double *m2_a, const double *__restrict__ n_b,
    const double *__restrict__ mean_b, const double *__restrict__ m2_b,
    double *mean_out, double *m2_out)
{
#pragma clang fp contract(off)
    const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= PC)
        return;
    const uint64_t p = i / C;
    const double na = n_a[p], nb = n_b[p];
    const double n = na + nb;
    const double ma = mean_a[i], delta = mean_b[i] - ma;
    const double m2 = m2_a[i] + m2_b[i] + delta * delta * (na * nb / n);
    mean_out[i] = ma + delta * (nb / n);
    m2_out[i] = m2;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
static const uint32_t kLdsMaxNodes = 16384u;
static const uint32_t kU32RecordMaxNodes = 65536u;
static const uint32_t kSweepMaxNodes = 1u << 28;
static const uint32_t kSumsMaxNodes = 1u << 21;
static const uint32_t kGlobalSlots = 4096u;

static inline hipStream_t as_stream(void *s) { return (hipStream_t)s; }

// parallel merge rounds per batch before the in-order remainder
#ifndef PC_MERGE_ROUNDS
#define PC_MERGE_ROUNDS 4
#endif
static uint32_t g_merge_rounds = PC_MERGE_ROUNDS;

// Two wavefronts per run (a finder and a resolver, see sweep_run) or one:
// unless the caller forces one (PC_SWEEP_ONE_WAVE / PC_SWEEP_TWO_WAVES in
// state_mode), two for HBM state and for LDS state from kSplitMinNodes nodes
// on.  Measured on MI355X: below that the second wave halves the runs a CU
// holds (L = 16: 4.8e10 additions/s with one wave, 3.8e10 with two), above
// it the finder wins (L = 64: 7.0e10 -> 7.4e10; L = 128: 3.3e10 -> 4.1e10).
static const uint32_t kSplitMinNodes = 4096u;

static bool sweep_split(uint32_t num_nodes, bool use_lds, uint32_t waves)
{
    if (waves == PC_SWEEP_ONE_WAVE)
        return false;
    if (waves == PC_SWEEP_TWO_WAVES)
        return true;
    return !use_lds || num_nodes >= kSplitMinNodes;
}

template <bool SPLIT>
static size_t sweep_lds_bytes(uint32_t num_nodes)
{
    return sweep_lds_words<SPLIT>() * sizeof(uint32_t) +
           (((size_t)num_nodes * sizeof(uint16_t) + 3u) & ~(size_t)3u);
}

extern "C" int pc_abi_version(void) { return 1; }

extern "C" const char *pc_last_error(void) { return g_last_error; }

extern "C" int pc_record_bytes(uint32_t num_nodes)
{
    return num_nodes <= kU32RecordMaxNodes ? 4 : 8;
}

// Bucket bits and chunks of the device permutations for num_edges edges.
static void shuffle_plan(uint32_t num_edges, uint32_t &lb1, uint32_t &chunks)
{
    lb1 = 0;
    if (num_edges > SHUF_WHOLE)
        while (((uint64_t)SHUF_BUCKET << lb1) < num_edges)
            ++lb1;
    chunks = (num_edges + SHUF_CHUNK - 1u) / SHUF_CHUNK;
}

extern "C" int pc_permute_philox_workspace_size(uint32_t num_edges,
                                                uint32_t num_runs,
                                                uint64_t *bytes)
{
    g_last_error[0] = 0;
    if (!bytes)
        return pc_fail(PC_EINVAL, "pc_permute_philox_workspace_size: bytes is NULL");
    if (num_edges > SHUF_MAX_EDGES)
        return pc_fail(PC_EOVERFLOW, "pc_permute_philox_workspace_size: num_edges > 2^23");
    uint32_t lb1 = 0, chunks = 0;
    shuffle_plan(num_edges, lb1, chunks);
    const uint64_t nb1 = 1ull << lb1;
    // a run of one bucket is shuffled in LDS from the identity: no counts
    *bytes = lb1 ? (uint64_t)num_runs * (nb1 * chunks + nb1 + 1u) * 4u : 0u;
    return PC_OK;
}

extern "C" int pc_permute_philox(uint64_t seed, uint64_t first_run,
                                 uint32_t num_runs, uint32_t num_edges,
                                 int32_t *d_perm, int64_t perm_stride,
                                 void *d_workspace, uint64_t workspace_bytes,
                                 void *stream)
{
    g_last_error[0] = 0;
    if (num_runs == 0 || num_edges == 0)
        return PC_OK;
    if (!d_perm || perm_stride < (int64_t)num_edges)
        return pc_fail(PC_EINVAL, "pc_permute_philox: bad output buffer");
    if (num_edges > SHUF_MAX_EDGES)
        return pc_fail(PC_EOVERFLOW, "pc_permute_philox: num_edges > 2^23");
    if (first_run + num_runs > (1ull << 32))
        return pc_fail(PC_EOVERFLOW, "pc_permute_philox: run indices reach 2^32");
    uint64_t need = 0;
    pc_permute_philox_workspace_size(num_edges, num_runs, &need);
    if (need && (!d_workspace || workspace_bytes < need))
        return pc_fail(PC_ENOMEM, "pc_permute_philox: workspace needs %llu bytes",
                       (unsigned long long)need);
    uint32_t lb1 = 0, chunks = 0;
    shuffle_plan(num_edges, lb1, chunks);
    const uint32_t nb1 = 1u << lb1;
    uint32_t *counts = (uint32_t *)d_workspace;
    uint32_t *starts = counts ? counts + (uint64_t)num_runs * nb1 * chunks : nullptr;
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    const uint32_t fr = (uint32_t)first_run;
    hipStream_t s = as_stream(stream);
    // runs in slabs of at most 65535 (the grid's y extent)
    for (uint32_t r0 = 0; r0 < num_runs; r0 += 65535u) {
        const uint32_t rn = min(num_runs - r0, 65535u);
        if (lb1 == 0u && num_edges <= SHUF_SMALL_CAP) {
            shuffle_group_kernel<2, true><<<dim3(1, rn), 128,
                                            shuf_group_lds_bytes(2, num_edges), s>>>(
                k0, k1, fr, 0u, r0, num_edges, nullptr, d_perm, perm_stride);
            PC_LAUNCH_CHECK("shuffle_group_kernel");
            continue;
        }
        if (lb1 == 0u) {
            shuffle_group_kernel<8, true><<<dim3(1, rn), 512,
                                            shuf_group_lds_bytes(8, num_edges), s>>>(
                k0, k1, fr, 0u, r0, num_edges, nullptr, d_perm, perm_stride);
            PC_LAUNCH_CHECK("shuffle_group_kernel");
            continue;
        }
        shuffle_count_kernel<<<shuf_blocks(chunks, rn), 64 * SHUF_W, nb1 * 4u, s>>>(
            k0, k1, fr, num_edges, lb1, chunks, r0, rn, counts);
        PC_LAUNCH_CHECK("shuffle_count_kernel");
        shuffle_scan_kernel<<<rn, 256, nb1 * 4u, s>>>(num_edges, lb1, chunks, r0,
                                                      counts, starts);
        PC_LAUNCH_CHECK("shuffle_scan_kernel");
        shuffle_scatter_kernel<<<shuf_blocks(chunks, rn), 64 * SHUF_W,
                                 (SHUF_W + 2u) * nb1 * 4u, s>>>(
            k0, k1, fr, num_edges, lb1, chunks, r0, rn, counts, d_perm, perm_stride);
        PC_LAUNCH_CHECK("shuffle_scatter_kernel");
        shuffle_group_kernel<4, false><<<dim3(nb1, rn), 256,
                                         shuf_group_lds_bytes(4, SHUF_GROUP_CAP), s>>>(
            k0, k1, fr, lb1, r0, SHUF_GROUP_CAP, starts, d_perm, perm_stride);
        PC_LAUNCH_CHECK("shuffle_group_kernel");
    }
    return PC_OK;
}

extern "C" int pc_sweep_workspace_size(uint32_t num_nodes, uint32_t num_runs,
                                       uint32_t state_mode, uint64_t *bytes)
{
    g_last_error[0] = 0;
    if (!bytes)
        return pc_fail(PC_EINVAL, "pc_sweep_workspace_size: bytes is NULL");
    state_mode &= ~PC_SWEEP_WAVES_MASK;
    if (state_mode == PC_STATE_LDS ||
        (state_mode == PC_STATE_AUTO && num_nodes <= kLdsMaxNodes)) {
        *bytes = 0;
    } else {
        const uint64_t slots = num_runs < kGlobalSlots ? num_runs : kGlobalSlots;
        *bytes = slots * (uint64_t)num_nodes * sizeof(uint32_t);
    }
    return PC_OK;
}

// The sweep launch for either edge source (pc_sweep, pc_sweep_lattice).
template <class ES>
static int sweep_launch(uint32_t num_nodes, uint32_t num_edges, ES edges,
                        bool have_edges, const uint8_t *d_node_sides,
                        int32_t span_on_first_merge, const int32_t *d_perm,
                        int64_t perm_stride, uint32_t num_runs, void *d_records,
                        int64_t rec_stride, uint32_t *d_span_step,
                        void *d_workspace, uint64_t workspace_bytes,
                        uint32_t merge_rounds, uint32_t state_mode,
                        uint64_t *d_chunk_state, void *stream)
{
    const uint32_t rounds = merge_rounds ? merge_rounds : g_merge_rounds;
    if (rounds > 64)
        return pc_fail(PC_EINVAL, "pc_sweep: merge_rounds > 64");
    const uint32_t waves = state_mode & PC_SWEEP_WAVES_MASK;
    state_mode &= ~PC_SWEEP_WAVES_MASK;
    if (state_mode > PC_STATE_HBM || waves == PC_SWEEP_WAVES_MASK)
        return pc_fail(PC_EINVAL, "pc_sweep: unknown state_mode %u", state_mode | waves);
    if (state_mode == PC_STATE_LDS && num_nodes > kLdsMaxNodes)
        return pc_fail(PC_EINVAL, "pc_sweep: LDS state needs num_nodes <= %u",
                       kLdsMaxNodes);
    const bool use_lds = state_mode == PC_STATE_LDS ||
                         (state_mode == PC_STATE_AUTO && num_nodes <= kLdsMaxNodes);
    g_last_error[0] = 0;
    if (num_runs == 0)
        return PC_OK;
    if (num_nodes == 0)
        return pc_fail(PC_EINVAL, "pc_sweep: num_nodes must be positive");
    if (num_nodes > kSweepMaxNodes)
        return pc_fail(PC_EOVERFLOW, "pc_sweep: num_nodes > 2^28");
    if (!d_span_step || (num_edges && (!have_edges || !d_perm || !d_records)))
        return pc_fail(PC_EINVAL, "pc_sweep: NULL buffer");
    if (num_edges && (perm_stride < (int64_t)num_edges ||
                      rec_stride < (int64_t)num_edges))
        return pc_fail(PC_EINVAL, "pc_sweep: stride smaller than num_edges");
    hipStream_t s = as_stream(stream);
    const bool u32rec = num_nodes <= kU32RecordMaxNodes;
    const bool split = sweep_split(num_nodes, use_lds, waves);
    // HBM state with chunk states: a third wave folds the records
    const bool folder = !use_lds && split && d_chunk_state;
    const uint32_t threads = folder ? 192u : split ? 128u : 64u;
    if (use_lds) {
        const size_t lds = split ? sweep_lds_bytes<true>(num_nodes)
                                 : sweep_lds_bytes<false>(num_nodes);
#define PC_SWEEP_LDS(CH, SP, CS)                                                    \
    sweep_lds_kernel<uint32_t, CH, SP, ES><<<num_runs, threads, lds, s>>>(               \
        num_nodes, num_edges, edges, d_node_sides, span_on_first_merge, d_perm,     \
        perm_stride, (uint32_t *)d_records, rec_stride, d_span_step, rounds, CS)
        if (d_chunk_state && split)
            PC_SWEEP_LDS(true, true, d_chunk_state);
        else if (d_chunk_state)
            PC_SWEEP_LDS(true, false, d_chunk_state);
        else if (split)
            PC_SWEEP_LDS(false, true, nullptr);
        else
            PC_SWEEP_LDS(false, false, nullptr);
#undef PC_SWEEP_LDS
        PC_LAUNCH_CHECK("sweep_lds_kernel");
        return PC_OK;
    }
    uint64_t need = 0;
    pc_sweep_workspace_size(num_nodes, num_runs, state_mode, &need);
    if (!d_workspace || workspace_bytes < (uint64_t)num_nodes * 4u)
        return pc_fail(PC_ENOMEM,
                       "pc_sweep: workspace of %llu bytes too small (need %llu)",
                       (unsigned long long)workspace_bytes,
                       (unsigned long long)need);
    uint64_t slots = workspace_bytes / ((uint64_t)num_nodes * 4u);
    if (slots > num_runs)
        slots = num_runs;
    if (slots > kGlobalSlots)
        slots = kGlobalSlots;
#define PC_SWEEP_GLOBAL(RT, CH, SP, FO)                                        \
    sweep_global_kernel<RT, CH, SP, ES, FO><<<(uint32_t)slots, threads, 0, s>>>(   \
        num_nodes, num_edges, num_runs, edges, d_node_sides,                   \
        span_on_first_merge, d_perm, perm_stride, (RT *)d_records, rec_stride, \
        d_span_step, (uint32_t *)d_workspace, rounds, d_chunk_state)
#define PC_SWEEP_GLOBAL_W(RT, CH)                                              \
    do {                                                                       \
        if (split)                                                             \
            PC_SWEEP_GLOBAL(RT, CH, true, false);                              \
        else                                                                   \
            PC_SWEEP_GLOBAL(RT, CH, false, false);                             \
    } while (0)
    if (folder && u32rec)
        PC_SWEEP_GLOBAL(uint32_t, true, true, true);
    else if (folder)
        PC_SWEEP_GLOBAL(uint64_t, true, true, true);
    else if (u32rec && d_chunk_state)
        PC_SWEEP_GLOBAL_W(uint32_t, true);
    else if (u32rec)
        PC_SWEEP_GLOBAL_W(uint32_t, false);
    else if (d_chunk_state)
        PC_SWEEP_GLOBAL_W(uint64_t, true);
    else
        PC_SWEEP_GLOBAL_W(uint64_t, false);
#undef PC_SWEEP_GLOBAL_W
#undef PC_SWEEP_GLOBAL
    PC_LAUNCH_CHECK("sweep_global_kernel");
    return PC_OK;
}

extern "C" int pc_sweep(uint32_t num_nodes, uint32_t num_edges,
                        const int32_t *d_edges, const uint8_t *d_node_sides,
                        int32_t span_on_first_merge, const int32_t *d_perm,
                        int64_t perm_stride, uint32_t num_runs, void *d_records,
                        int64_t rec_stride, uint32_t *d_span_step,
                        void *d_workspace, uint64_t workspace_bytes,
                        uint32_t merge_rounds, uint32_t state_mode,
                        uint64_t *d_chunk_state, void *stream)
{
    return sweep_launch(num_nodes, num_edges, EdgeList{(const int2 *)d_edges},
                        d_edges != nullptr, d_node_sides, span_on_first_merge,
                        d_perm, perm_stride, num_runs, d_records, rec_stride,
                        d_span_step, d_workspace, workspace_bytes, merge_rounds,
                        state_mode, d_chunk_state, stream);
}

static LatticeEdges lattice_edges(uint32_t length, bool tiled)
{
    const uint32_t d = 2u * length - 1u;
    return LatticeEdges{length, d, 1.0f / (float)d, tiled ? 1u : 0u};
}

extern "C" int pc_sweep_lattice(uint32_t length, int32_t tiled,
                                const uint8_t *d_node_sides,
                                int32_t span_on_first_merge, const int32_t *d_perm,
                                int64_t perm_stride, uint32_t num_runs,
                                void *d_records, int64_t rec_stride,
                                uint32_t *d_span_step, void *d_workspace,
                                uint64_t workspace_bytes, uint32_t merge_rounds,
                                uint32_t state_mode, uint64_t *d_chunk_state,
                                void *stream)
{
    g_last_error[0] = 0;
    if (length == 0 || length > 16384u)
        return pc_fail(PC_EINVAL, "pc_sweep_lattice: length must be in 1..16384");
    if (tiled && (length % 8u))
        return pc_fail(PC_EINVAL, "pc_sweep_lattice: the tiled order needs a multiple of 8");
    const uint32_t N = length * length, M = 2u * length * (length - 1u);
    return sweep_launch(N, M, lattice_edges(length, tiled != 0), true, d_node_sides,
                        span_on_first_merge, d_perm, perm_stride, num_runs,
                        d_records, rec_stride, d_span_step, d_workspace,
                        workspace_bytes, merge_rounds, state_mode, d_chunk_state,
                        stream);
}

extern "C" int pc_chunk_scan(uint32_t num_nodes, uint32_t num_edges,
                             uint32_t num_runs, const void *d_records,
                             int64_t rec_stride, uint64_t *d_chunk_state,
                             void *stream)
{
    g_last_error[0] = 0;
    if (num_runs == 0 || num_edges == 0)
        return PC_OK;
    if (num_nodes == 0 || num_nodes > kSweepMaxNodes)
        return pc_fail(PC_EOVERFLOW, "pc_chunk_scan: num_nodes out of range");
    if (!d_records || !d_chunk_state || rec_stride < (int64_t)num_edges)
        return pc_fail(PC_EINVAL, "pc_chunk_scan: bad buffer");
    hipStream_t s = as_stream(stream);
    const uint32_t C = (num_edges + 63u) / 64u;
    const uint64_t blocks = (uint64_t)((C + CHUNK_DELTA_PER_BLOCK - 1u) / CHUNK_DELTA_PER_BLOCK) *
                            num_runs;
    if (blocks > 0x7FFFFFFFu)
        return pc_fail(PC_EOVERFLOW, "pc_chunk_scan: too many runs x chunks in one call");
    const dim3 grid((uint32_t)blocks);
    if (num_nodes <= kLdsMaxNodes)
        chunk_delta_kernel<uint32_t, false><<<grid, 64, 0, s>>>(
            num_edges, (const uint32_t *)d_records, rec_stride, d_chunk_state);
    else if (num_nodes <= kU32RecordMaxNodes)
        chunk_delta_kernel<uint32_t, true><<<grid, 64, 0, s>>>(
            num_edges, (const uint32_t *)d_records, rec_stride, d_chunk_state);
    else
        chunk_delta_kernel<uint64_t, true><<<grid, 64, 0, s>>>(
            num_edges, (const uint64_t *)d_records, rec_stride, d_chunk_state);
    PC_LAUNCH_CHECK("chunk_delta_kernel");
    if (num_nodes <= kLdsMaxNodes)
        chunk_prefix_kernel<false><<<num_runs, 64, 0, s>>>(num_nodes, num_edges,
                                                           d_chunk_state);
    else
        chunk_prefix_kernel<true><<<num_runs, 64, 0, s>>>(num_nodes, num_edges,
                                                          d_chunk_state);
    PC_LAUNCH_CHECK("chunk_prefix_kernel");
    return PC_OK;
}

extern "C" int pc_microcanonical_sums(uint32_t num_nodes, uint32_t num_edges,
                                      uint32_t num_runs, const void *d_records,
                                      int64_t rec_stride,
                                      const uint32_t *d_span_step,
                                      const uint64_t *d_chunk_state,
                                      uint64_t *d_sum1, uint64_t *d_sum2,
                                      int32_t accumulate, void *stream)
{
    g_last_error[0] = 0;
    if (num_nodes == 0 || num_nodes > kSumsMaxNodes)
        return pc_fail(PC_EOVERFLOW,
                       "pc_microcanonical_sums: num_nodes outside 1..2^21");
    if (num_nodes <= kLdsMaxNodes) {
        // the 128-bit sums of squares hold R * (N^4 / 8)^2
        const double n4 = (double)num_nodes * num_nodes * num_nodes * num_nodes / 8.0;
        if ((double)num_runs * n4 * n4 >= 0x1p127)
            return pc_fail(PC_EOVERFLOW,
                           "pc_microcanonical_sums: too many runs for N = %u",
                           num_nodes);
    }
    if (!d_sum1 || !d_sum2 || (num_runs && num_edges &&
                               (!d_records || !d_span_step || !d_chunk_state)))
        return pc_fail(PC_EINVAL, "pc_microcanonical_sums: NULL buffer");
    hipStream_t s = as_stream(stream);
    const uint64_t S = (uint64_t)num_edges + 1u;
    if (!accumulate) {
        fill_zero_kernel<<<1024, 256, 0, s>>>(d_sum1, 7u * S * 2u);
        fill_zero_kernel<<<1024, 256, 0, s>>>(d_sum2, 7u * S * 3u);
        PC_LAUNCH_CHECK("fill_zero_kernel");
    }
    if (num_runs == 0)
        return PC_OK;
    if (num_edges == 0) {
        step0_sums_kernel<<<1, 1, 0, s>>>(num_nodes, num_runs, S, d_sum1, d_sum2);
        PC_LAUNCH_CHECK("step0_sums_kernel");
        return PC_OK;
    }
    const uint32_t C = (num_edges + 63u) / 64u;
    // enough blocks to fill 256 CUs several times over
    uint32_t groups = (8192u + C - 1u) / C;
    if (groups > num_runs)
        groups = num_runs;
    // at most 2^16 runs per group: the narrow 128-bit sum of squares of the
    // fourth moment (each below 2^112) stays exact
    if ((num_runs + groups - 1u) / groups > 65536u)
        groups = (num_runs + 65535u) / 65536u;
    const uint32_t rpg = (num_runs + groups - 1u) / groups;
    groups = (num_runs + rpg - 1u) / rpg;
    const dim3 grid(C, groups);
    if (num_nodes <= kLdsMaxNodes)
        micro_sums_kernel<uint32_t, false><<<grid, 64, 0, s>>>(
            num_nodes, num_edges, num_runs, rpg, (const uint32_t *)d_records,
            rec_stride, d_span_step, d_chunk_state, d_sum1, d_sum2);
    else if (num_nodes <= kU32RecordMaxNodes)
        micro_sums_kernel<uint32_t, true><<<grid, 64, 0, s>>>(
            num_nodes, num_edges, num_runs, rpg, (const uint32_t *)d_records,
            rec_stride, d_span_step, d_chunk_state, d_sum1, d_sum2);
    else
        micro_sums_kernel<uint64_t, true><<<grid, 64, 0, s>>>(
            num_nodes, num_edges, num_runs, rpg, (const uint64_t *)d_records,
            rec_stride, d_span_step, d_chunk_state, d_sum1, d_sum2);
    PC_LAUNCH_CHECK("micro_sums_kernel");
    return PC_OK;
}

extern "C" int pc_expand_traces(uint32_t num_nodes, uint32_t num_edges,
                                uint32_t num_runs, const void *d_records,
                                int64_t rec_stride, const uint32_t *d_span_step,
                                const uint64_t *d_chunk_state,
                                uint32_t *d_max_size, uint8_t *d_spanning,
                                uint64_t *d_moments, void *stream)
{
    g_last_error[0] = 0;
    if (num_runs == 0)
        return PC_OK;
    if (num_nodes == 0 || num_nodes > kSumsMaxNodes)
        return pc_fail(PC_EOVERFLOW, "pc_expand_traces: num_nodes outside 1..2^21");
    if (!d_max_size || !d_spanning || !d_moments ||
        (num_edges && (!d_records || !d_span_step || !d_chunk_state)))
        return pc_fail(PC_EINVAL, "pc_expand_traces: NULL buffer");
    hipStream_t s = as_stream(stream);
    const uint64_t S = (uint64_t)num_edges + 1u;
    if (num_edges == 0) {
        expand_step0_kernel<<<(num_runs + 255u) / 256u, 256, 0, s>>>(
            num_nodes, num_runs, S, d_max_size, d_spanning, d_moments);
        PC_LAUNCH_CHECK("expand_step0_kernel");
        return PC_OK;
    }
    const uint32_t C = (num_edges + 63u) / 64u;
    const dim3 grid(C, num_runs);
    if (num_nodes <= kLdsMaxNodes)
        expand_kernel<uint32_t, false><<<grid, 64, 0, s>>>(
            num_nodes, num_edges, (const uint32_t *)d_records, rec_stride,
            d_span_step, d_chunk_state, d_max_size, d_spanning, d_moments);
    else if (num_nodes <= kU32RecordMaxNodes)
        expand_kernel<uint32_t, true><<<grid, 64, 0, s>>>(
            num_nodes, num_edges, (const uint32_t *)d_records, rec_stride,
            d_span_step, d_chunk_state, d_max_size, d_spanning, d_moments);
    else
        expand_kernel<uint64_t, true><<<grid, 64, 0, s>>>(
            num_nodes, num_edges, (const uint64_t *)d_records, rec_stride,
            d_span_step, d_chunk_state, d_max_size, d_spanning, d_moments);
    PC_LAUNCH_CHECK("expand_kernel");
    return PC_OK;
}

extern "C" int pc_binomial_window(uint32_t n, double p, int64_t *start,
                                  int64_t *len)
{
    g_last_error[0] = 0;
    if (!start || !len)
        return pc_fail(PC_EINVAL, "pc_binomial_window: NULL output");
    if (!(p >= 0.0 && p <= 1.0))
        return pc_fail(PC_EINVAL, "pc_binomial_window: p outside [0, 1]");
    const double mu = (double)n * p;
    const double sigma = sqrt((double)n * p * (1.0 - p));
    const double half = 12.0 * sigma + 48.0;
    double lo = floor(mu - half), hi = ceil(mu + half);
    if (lo < 0.0)
        lo = 0.0;
    if (hi > (double)n)
        hi = (double)n;
    *start = (int64_t)lo;
    *len = (int64_t)hi - (int64_t)lo + 1;
    return PC_OK;
}

extern "C" int pc_binomial_weights(uint32_t n, uint32_t num_p,
                                   const double *d_ps, const int64_t *d_start,
                                   const int64_t *d_len, int64_t kmax,
                                   double *d_w, void *stream)
{
    g_last_error[0] = 0;
    if (num_p == 0)
        return PC_OK;
    if (!d_ps || !d_start || !d_len || !d_w || kmax <= 0)
        return pc_fail(PC_EINVAL, "pc_binomial_weights: bad buffer");
    hipStream_t s = as_stream(stream);
    const dim3 grid((uint32_t)((kmax + 255) / 256), num_p);
    binom_weights_kernel<<<grid, 256, 0, s>>>(n, d_ps, d_start, d_len, kmax, d_w);
    PC_LAUNCH_CHECK("binom_weights_kernel");
    binom_normalize_kernel<<<num_p, 256, 0, s>>>(kmax, d_len, d_w);
    PC_LAUNCH_CHECK("binom_normalize_kernel");
    return PC_OK;
}

// K-slices of the MFMA contraction: enough tiles x slices to fill the GPU
// with about 16k wavefronts, at most PC_CONTRACT_SLICES.
#ifndef PC_CONTRACT_WAVES
#define PC_CONTRACT_WAVES 16384u
#endif
static uint32_t contract_slices(uint32_t num_p, uint32_t cols)
{
    const uint64_t tiles = (uint64_t)((num_p + 15u) / 16u) * ((cols + 15u) / 16u);
    uint64_t sl = (PC_CONTRACT_WAVES + tiles - 1u) / (tiles ? tiles : 1u);
    if (sl < 1u)
        sl = 1u;
    if (sl > PC_CONTRACT_SLICES)
        sl = PC_CONTRACT_SLICES;
    return (uint32_t)sl;
}

// The FP64 matrix-core contraction takes the wide kernel from this many
// columns on, with tiles of 64 p x WIDE_TC columns.
#define WIDE_MIN_COLS 64u
#ifndef WIDE_TC
#define WIDE_TC 64
#endif

static bool contract_wide(uint32_t cols, int32_t precision)
{
    return precision == PC_PREC_FP64_MFMA && cols >= WIDE_MIN_COLS;
}

static uint32_t wide_tiles_p(uint32_t num_p) { return (num_p + WIDE_TILE - 1u) / WIDE_TILE; }
static uint32_t wide_tiles(uint32_t num_p, uint32_t cols)
{
    return wide_tiles_p(num_p) * ((cols + WIDE_TC - 1u) / WIDE_TC);
}

// K-slices of the wide kernel: none from four workgroups per CU on (1024
// tiles), otherwise enough to reach about that many, at most 16.
static uint32_t wide_slices(uint32_t num_p, uint32_t cols)
{
    const uint64_t tiles = wide_tiles(num_p, cols);
    if (tiles >= 1024u)
        return 1u;
    uint64_t sl = (1024u + tiles - 1u) / (tiles ? tiles : 1u);
    return (uint32_t)(sl > PC_CONTRACT_SLICES ? PC_CONTRACT_SLICES : sl);
}

extern "C" int pc_contract_workspace_size(uint32_t num_p, uint32_t cols,
                                          int32_t precision, uint64_t *bytes)
{
    g_last_error[0] = 0;
    if (!bytes)
        return pc_fail(PC_EINVAL, "pc_contract_workspace_size: bytes is NULL");
    if (precision == PC_PREC_FP64)
        *bytes = 0u;
    else if (contract_wide(cols, precision)) {
        const uint32_t sl = wide_slices(num_p, cols);
        *bytes = sl > 1u ? (uint64_t)sl * num_p * cols * sizeof(double) : 0u;
    } else
        *bytes = (uint64_t)contract_slices(num_p, cols) * num_p * cols *
                 sizeof(double);
    return PC_OK;
}

extern "C" int pc_canonical_contract(uint32_t num_steps, uint32_t cols,
                                     const void *d_x, int64_t x_stride,
                                     uint32_t num_p, const int64_t *d_start,
                                     const int64_t *d_len, const void *d_w,
                                     int64_t kmax, double *d_out,
                                     int32_t precision, void *d_workspace,
                                     uint64_t workspace_bytes, void *stream)
{
    g_last_error[0] = 0;
    if (num_p == 0 || cols == 0)
        return PC_OK;
    if (precision == PC_PREC_FP64 && cols > CONTRACT_MAX_COLS)
        return pc_fail(PC_EOVERFLOW,
                       "pc_canonical_contract: FP64 vector form takes at most %d columns",
                       CONTRACT_MAX_COLS);
    if (!d_x || !d_start || !d_len || !d_w || !d_out || x_stride < (int64_t)cols)
        return pc_fail(PC_EINVAL, "pc_canonical_contract: bad buffer");
    hipStream_t s = as_stream(stream);
    if (precision == PC_PREC_FP64) {
        const double *x = (const double *)d_x, *w = (const double *)d_w;
        if (cols <= 8)
            contract_f64_kernel<8><<<num_p, 256, 0, s>>>(
                x, x_stride, d_start, d_len, w, kmax, d_out, (int32_t)cols);
        else if (cols <= 16)
            contract_f64_kernel<16><<<num_p, 256, 0, s>>>(
                x, x_stride, d_start, d_len, w, kmax, d_out, (int32_t)cols);
        else
            contract_f64_kernel<32><<<num_p, 256, 0, s>>>(
                x, x_stride, d_start, d_len, w, kmax, d_out, (int32_t)cols);
        PC_LAUNCH_CHECK("contract_f64_kernel");
        return PC_OK;
    }
    uint64_t need = 0;
    pc_contract_workspace_size(num_p, cols, precision, &need);
    if (need > 0u && (!d_workspace || workspace_bytes < need))
        return pc_fail(PC_ENOMEM, "pc_canonical_contract: workspace needs %llu bytes",
                       (unsigned long long)need);
    const TileCtx t{num_p, cols, x_stride, kmax, d_start, d_len};
    if (contract_wide(cols, precision)) {
        const uint32_t sl = wide_slices(num_p, cols);
        const dim3 grid(wide_tiles(num_p, cols), sl);
        double *dst = sl > 1u ? (double *)d_workspace : d_out;
        contract_wide_f64_kernel<WIDE_TC><<<grid, 256, 0, s>>>(
            (const double *)d_x, num_steps, (const double *)d_w, t,
            wide_tiles_p(num_p), sl, dst);
        PC_LAUNCH_CHECK("contract_wide_f64_kernel");
        if (sl > 1u) {
            const int64_t n = (int64_t)num_p * cols;
            contract_reduce_kernel<<<(uint32_t)((n + 255) / 256), 256, 0, s>>>(
                dst, sl, n, d_out);
            PC_LAUNCH_CHECK("contract_reduce_kernel");
        }
        return PC_OK;
    }
    const uint32_t slices = contract_slices(num_p, cols);
    const dim3 grid((num_p + 15u) / 16u, (cols + 15u) / 16u, slices);
    double *partial = (double *)d_workspace;
    switch (precision) {
    case PC_PREC_FP64_MFMA:
        contract_mfma_kernel<PC_PREC_FP64_MFMA><<<grid, 64, 0, s>>>(
            d_x, num_steps, d_w, t, slices, partial);
        break;
    case PC_PREC_FP32_MFMA:
        contract_mfma_kernel<PC_PREC_FP32_MFMA><<<grid, 64, 0, s>>>(
            d_x, num_steps, d_w, t, slices, partial);
        break;
    case PC_PREC_BF16X3_MFMA:
        contract_mfma_kernel<PC_PREC_BF16X3_MFMA><<<grid, 64, 0, s>>>(
            d_x, num_steps, d_w, t, slices, partial);
        break;
    default:
        return pc_fail(PC_EINVAL, "pc_canonical_contract: unknown precision %d",
                       precision);
    }
    PC_LAUNCH_CHECK("contract_mfma_kernel");
    const int64_t n = (int64_t)num_p * cols;
    contract_reduce_kernel<<<(uint32_t)((n + 255) / 256), 256, 0, s>>>(
        partial, slices, n, d_out);
    PC_LAUNCH_CHECK("contract_reduce_kernel");
    return PC_OK;
}

extern "C" int pc_canonical_fused_workspace_size(uint32_t num_p, uint32_t cols,
                                                 int32_t precision, uint64_t *bytes)
{
    g_last_error[0] = 0;
    if (!bytes)
        return pc_fail(PC_EINVAL, "pc_canonical_fused_workspace_size: bytes is NULL");
    if (precision != PC_PREC_FP64_MFMA && precision != PC_PREC_SPLIT_BF16)
        return pc_fail(PC_EINVAL, "pc_canonical_fused_workspace_size: unknown precision %d",
                       precision);
    *bytes = (uint64_t)contract_slices(num_p, cols + 1u) * num_p * (cols + 1u) *
             sizeof(double);
    return PC_OK;
}

// Checks of pc_canonical_fused's arguments; returns PC_OK or a failure.
static int fused_check(const char *who, uint32_t cols, const double *d_x,
                       int64_t x_stride, const double *d_ps,
                       const int64_t *d_start, const int64_t *d_len,
                       double *d_out)
{
    if (cols > 4096u)
        return pc_fail(PC_EOVERFLOW, "%s: at most 4096 columns", who);
    if (!d_x || !d_ps || !d_start || !d_len || !d_out || x_stride < (int64_t)cols)
        return pc_fail(PC_EINVAL, "%s: bad buffer", who);
    return PC_OK;
}

// The launches of pc_canonical_fused with the partial sums in `partial`
// (pc_canonical_fused_workspace_size bytes; every entry is written before
// it is read, so it needs no clearing).
static int fused_launch(uint32_t num_edges, uint32_t cols, const double *d_x,
                        int64_t x_stride, uint32_t num_p, const double *d_ps,
                        const int64_t *d_start, const int64_t *d_len,
                        double *d_out, int32_t precision, double *partial,
                        hipStream_t s)
{
    const FusedCtx f{d_x, x_stride, d_ps, d_start, d_len, num_edges, num_p, cols};
    const uint32_t slices = contract_slices(num_p, cols + 1u);
    const dim3 grid((num_p + 15u) / 16u, (cols + 1u + 15u) / 16u, slices);
    if (precision == PC_PREC_FP64_MFMA) {
        canonical_fused_f64_kernel<<<grid, 64, 0, s>>>(f, slices, partial);
        PC_LAUNCH_CHECK("canonical_fused_f64_kernel");
    } else {
        canonical_fused_split_kernel<<<grid, 64, 0, s>>>(f, slices, partial);
        PC_LAUNCH_CHECK("canonical_fused_split_kernel");
    }
    const int64_t n = (int64_t)num_p * cols;
    fused_reduce_kernel<<<(uint32_t)((n + 255) / 256), 256, 0, s>>>(
        partial, slices, num_p, cols, d_out);
    PC_LAUNCH_CHECK("fused_reduce_kernel");
    return PC_OK;
}

extern "C" int pc_canonical_fused(uint32_t num_edges, uint32_t cols,
                                  const double *d_x, int64_t x_stride,
                                  uint32_t num_p, const double *d_ps,
                                  const int64_t *d_start, const int64_t *d_len,
                                  double *d_out, int32_t precision,
                                  void *d_workspace, uint64_t workspace_bytes,
                                  void *stream)
{
    g_last_error[0] = 0;
    if (num_p == 0 || cols == 0)
        return PC_OK;
    int rc = fused_check("pc_canonical_fused", cols, d_x, x_stride, d_ps,
                         d_start, d_len, d_out);
    if (rc != PC_OK)
        return rc;
    uint64_t need = 0;
    if (pc_canonical_fused_workspace_size(num_p, cols, precision, &need) != PC_OK)
        return PC_EINVAL;
    if (!d_workspace || workspace_bytes < need)
        return pc_fail(PC_ENOMEM, "pc_canonical_fused: workspace needs %llu bytes",
                       (unsigned long long)need);
    return fused_launch(num_edges, cols, d_x, x_stride, num_p, d_ps, d_start,
                        d_len, d_out, precision, (double *)d_workspace,
                        as_stream(stream));
}

// ---------------------------------------------------------------------------
// Scratch of pc_canonical_fused_noalloc: a file-scope device array, one copy
// per GPU, held until the process ends.  Host code never passes the array's
// name to a kernel (on the host it is only a stand-in for the symbol): the
// device address of each GPU's copy comes from hipGetSymbolAddress, looked
// up once per device with that device current, and is used everywhere.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[PC_NOALLOC_SCRATCH_BYTES];

#define PC_MAX_DEVICES 64
static std::mutex g_scratch_mutex;
static void *g_scratch_ptr[PC_MAX_DEVICES];

// Device address of the current device's scratch copy.
static int scratch_address(void **out)
{
    int dev = -1;
    PC_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= PC_MAX_DEVICES)
        return pc_fail(PC_EINVAL, "scratch: device ordinal %d outside 0..%d", dev,
                       PC_MAX_DEVICES - 1);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        PC_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p)
            return pc_fail(PC_EHIP, "scratch: hipGetSymbolAddress gave NULL");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return PC_OK;
}

extern "C" int pc_canonical_fused_noalloc(uint32_t num_edges, uint32_t cols,
                                          const double *d_x, int64_t x_stride,
                                          uint32_t num_p, const double *d_ps,
                                          const int64_t *d_start,
                                          const int64_t *d_len, double *d_out,
                                          int32_t precision, void *stream)
{
    g_last_error[0] = 0;
    if (num_p == 0 || cols == 0)
        return PC_OK;
    int rc = fused_check("pc_canonical_fused_noalloc", cols, d_x, x_stride, d_ps,
                         d_start, d_len, d_out);
    if (rc != PC_OK)
        return rc;
    uint64_t need = 0;
    if (pc_canonical_fused_workspace_size(num_p, cols, precision, &need) != PC_OK)
        return PC_EINVAL;
    if (need > (uint64_t)PC_NOALLOC_SCRATCH_BYTES)
        return pc_fail(PC_ENOMEM,
                       "pc_canonical_fused_noalloc: %u p x %u columns need %llu "
                       "bytes of scratch, above the fixed %u (use "
                       "pc_canonical_fused with a workspace)",
                       num_p, cols, (unsigned long long)need,
                       (unsigned)PC_NOALLOC_SCRATCH_BYTES);
    void *p = nullptr;
    rc = scratch_address(&p);
    if (rc != PC_OK)
        return rc;
    return fused_launch(num_edges, cols, d_x, x_stride, num_p, d_ps, d_start,
                        d_len, d_out, precision, (double *)p, as_stream(stream));
}

extern "C" int pc_convert_f64_to_f32(const double *d_in, int64_t n, float *d_out,
                                     void *stream)
{
    g_last_error[0] = 0;
    if (n <= 0)
        return PC_OK;
    if (!d_in || !d_out)
        return pc_fail(PC_EINVAL, "pc_convert_f64_to_f32: NULL buffer");
    f64_to_f32_kernel<<<2048, 256, 0, as_stream(stream)>>>(d_in, n, d_out);
    PC_LAUNCH_CHECK("f64_to_f32_kernel");
    return PC_OK;
}

extern "C" int pc_microcanonical_ci(uint32_t num_edges, uint32_t num_runs,
                                    const uint64_t *d_sum1, const uint64_t *d_sum2,
                                    double t_lo, double t_hi, const double *d_beta,
                                    double *d_out, void *stream)
{
    g_last_error[0] = 0;
    if (num_runs == 0)
        return pc_fail(PC_EINVAL, "pc_microcanonical_ci: no runs");
    if (!d_sum1 || !d_sum2 || !d_beta || !d_out)
        return pc_fail(PC_EINVAL, "pc_microcanonical_ci: NULL buffer");
    const uint64_t S = (uint64_t)num_edges + 1u;
    micro_ci_kernel<<<(uint32_t)((S + 255) / 256), 256, 0, as_stream(stream)>>>(
        S, num_runs, d_sum1, d_sum2, t_lo, t_hi, d_beta, d_out);
    PC_LAUNCH_CHECK("micro_ci_kernel");
    return PC_OK;
}

extern "C" int pc_finalize_canonical(uint32_t num_ps, uint32_t num_columns,
                                     const uint32_t *d_runs, const double *d_mean,
                                     const double *d_m2, const double *d_div,
                                     const double *d_q_lo, const double *d_q_hi,
                                     double *d_out, void *stream)
{
    g_last_error[0] = 0;
    const uint64_t total = (uint64_t)num_ps * num_columns;
    if (total == 0)
        return PC_OK;
    if (!d_runs || !d_mean || !d_m2 || !d_div || !d_q_lo || !d_q_hi || !d_out)
        return pc_fail(PC_EINVAL, "pc_finalize_canonical: NULL buffer");
    if ((total + 255u) / 256u > 0x7FFFFFFFull)
        return pc_fail(PC_EOVERFLOW, "pc_finalize_canonical: too many rows");
    finalize_kernel<<<(uint32_t)((total + 255u) / 256u), 256, 0, as_stream(stream)>>>(
        num_ps, num_columns, d_runs, d_mean, d_m2, d_div, d_q_lo, d_q_hi, d_out);
    PC_LAUNCH_CHECK("finalize_kernel");
    return PC_OK;
}

extern "C" int pc_merge_moments(uint32_t num_ps, uint32_t num_columns,
                                const double *d_n_a, const double *d_mean_a,
                                const double *d_m2_a, const double *d_n_b,
                                const double *d_mean_b, const double *d_m2_b,
                                double *d_mean_out, double *d_m2_out, void *stream)
{
    g_last_error[0] = 0;
    const uint64_t total = (uint64_t)num_ps * num_columns;
    if (total == 0)
        return PC_OK;
    if (!d_n_a || !d_mean_a || !d_m2_a || !d_n_b || !d_mean_b || !d_m2_b ||
        !d_mean_out || !d_m2_out)
        return pc_fail(PC_EINVAL, "pc_merge_moments: NULL buffer");
    if ((total + 255u) / 256u > 0x7FFFFFFFull)
        return pc_fail(PC_EOVERFLOW, "pc_merge_moments: too many rows");
    merge_moments_kernel<<<(uint32_t)((total + 255u) / 256u), 256, 0, as_stream(stream)>>>(
        total, num_columns, d_n_a, d_mean_a, d_m2_a, d_n_b, d_mean_b, d_m2_b,
        d_mean_out, d_m2_out);
    PC_LAUNCH_CHECK("merge_moments_kernel");
    return PC_OK;
}

extern "C" int pc_microcanonical_moments(uint32_t num_edges, uint32_t num_runs,
                                         const uint64_t *d_sum1,
                                         const uint64_t *d_sum2, double *d_mean,
                                         double *d_std, void *stream)
{
    g_last_error[0] = 0;
    if (num_runs == 0)
        return pc_fail(PC_EINVAL, "pc_microcanonical_moments: no runs");
    if (!d_sum1 || !d_sum2 || !d_mean || !d_std)
        return pc_fail(PC_EINVAL, "pc_microcanonical_moments: NULL buffer");
    const uint64_t total = (uint64_t)PC_COLUMNS * ((uint64_t)num_edges + 1u);
    micro_moments_kernel<<<(uint32_t)((total + 255) / 256), 256, 0,
                           as_stream(stream)>>>(total, num_runs, d_sum1, d_sum2,
                                                d_mean, d_std);
    PC_LAUNCH_CHECK("micro_moments_kernel");
    return PC_OK;
}

// Plan of a persistent multi-graph sweep: device table entries, the LDS
// bytes (the largest LDS-state graph), the HBM slot size (the largest
// HBM-state graph), the queue length and the grid (workgroups resident at
// once).
struct MultiPlan {
    std::vector<MultiGraph> table;
    size_t lds;
    uint64_t slot_words, total;
    uint32_t grid;
};

static int multi_plan(uint32_t num_graphs, const pc_sweep_graph *graphs,
                      MultiPlan &plan)
{
    if (num_graphs == 0 || !graphs)
        return pc_fail(PC_EINVAL, "pc_sweep_multi: no graphs");
    plan.table.assign(num_graphs, MultiGraph{});
    plan.lds = sweep_lds_words<true>() * sizeof(uint32_t);
    plan.slot_words = 0;
    plan.total = 0;
    for (uint32_t g = 0; g < num_graphs; ++g) {
        const pc_sweep_graph &in = graphs[g];
        MultiGraph &G = plan.table[g];
        const uint32_t rounds = in.merge_rounds ? in.merge_rounds : g_merge_rounds;
        if (in.num_nodes == 0 || in.num_nodes > kSweepMaxNodes || rounds > 64)
            return pc_fail(PC_EINVAL, "pc_sweep_multi: graph %u: bad sizes", g);
        const uint32_t L = in.lattice_length;
        if (L && ((uint64_t)L * L != in.num_nodes ||
                  in.num_edges != 2u * L * (L - 1u) ||
                  (in.num_nodes > kLdsMaxNodes && (L % 8u))))
            return pc_fail(PC_EINVAL, "pc_sweep_multi: graph %u is not the lattice of "
                           "side %u (tiled order needs a multiple of 8)", g, L);
        if (in.num_runs && in.num_edges &&
            ((!L && !in.d_edges) || !in.d_perm || !in.d_records || !in.d_span_step ||
             !in.d_chunk_state || in.perm_stride < (int64_t)in.num_edges ||
             in.rec_stride < (int64_t)in.num_edges))
            return pc_fail(PC_EINVAL, "pc_sweep_multi: graph %u: bad buffers", g);
        if (in.num_runs && !in.d_span_step)
            return pc_fail(PC_EINVAL, "pc_sweep_multi: graph %u: no span output", g);
        G.edges = (const int2 *)in.d_edges;
        G.lattice = L ? lattice_edges(L, in.num_nodes > kLdsMaxNodes)
                      : LatticeEdges{0u, 1u, 1.0f, 0u};
        G.sides = in.d_node_sides;
        G.perm = in.d_perm;
        G.records = in.d_records;
        G.span_step = in.d_span_step;
        G.chunk_state = in.d_chunk_state;
        G.perm_stride = in.perm_stride;
        G.rec_stride = in.rec_stride;
        G.N = in.num_nodes;
        G.M = in.num_edges;
        G.runs = in.num_runs;
        G.first = (uint32_t)plan.total;
        G.span_first = in.span_on_first_merge ? 1u : 0u;
        G.rounds = rounds;
        G.shift = in.num_nodes > 1u ? 32u - (uint32_t)__builtin_clz(in.num_nodes - 1u) : 0u;
        if (in.num_nodes <= kLdsMaxNodes) {
            G.kind = MULTI_LDS;
            plan.lds = std::max(plan.lds, sweep_lds_bytes<true>(in.num_nodes));
        } else {
            G.kind = in.num_nodes <= kU32RecordMaxNodes ? MULTI_HBM32 : MULTI_HBM64;
            plan.slot_words = std::max<uint64_t>(plan.slot_words, in.num_nodes);
            plan.lds = std::max(plan.lds, (sweep_lds_words<true>() + DIRTY_SLOTS) *
                                              sizeof(uint32_t));
        }
        plan.total += in.num_runs;
    }
    if (plan.total >= (1ull << 32) - 65536u)
        return pc_fail(PC_EOVERFLOW, "pc_sweep_multi: too many runs");
    int dev = 0, cus = 0, blocks = 0;
    PC_HIP_CHECK(hipGetDevice(&dev));
    PC_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    PC_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &blocks, sweep_multi_kernel, 128, plan.lds));
    if (blocks < 1)
        return pc_fail(PC_EINVAL, "pc_sweep_multi: %zu bytes of LDS do not fit", plan.lds);
    uint64_t grid = (uint64_t)cus * (uint64_t)blocks;
    if (grid > plan.total)
        grid = plan.total ? plan.total : 1u;
    plan.grid = (uint32_t)grid;
    return PC_OK;
}

// The table, then the queue head at the next 256-byte boundary.
static uint64_t multi_head_offset(uint32_t num_graphs)
{
    return ((uint64_t)num_graphs * sizeof(MultiGraph) + 255u) & ~(uint64_t)255u;
}
static uint64_t multi_table_bytes(uint32_t num_graphs)
{
    return multi_head_offset(num_graphs) + 256u;
}

extern "C" int pc_sweep_multi_workspace_size(uint32_t num_graphs,
                                             const pc_sweep_graph *graphs,
                                             uint64_t *table_bytes,
                                             uint64_t *workspace_bytes)
{
    g_last_error[0] = 0;
    if (!table_bytes || !workspace_bytes)
        return pc_fail(PC_EINVAL, "pc_sweep_multi_workspace_size: NULL output");
    MultiPlan plan;
    const int rc = multi_plan(num_graphs, graphs, plan);
    if (rc)
        return rc;
    *table_bytes = multi_table_bytes(num_graphs);
    *workspace_bytes = (uint64_t)plan.grid * plan.slot_words * 4u;
    return PC_OK;
}

extern "C" int pc_sweep_multi_prepare(uint32_t num_graphs,
                                      const pc_sweep_graph *graphs,
                                      void *d_table, uint64_t table_bytes)
{
    g_last_error[0] = 0;
    MultiPlan plan;
    const int rc = multi_plan(num_graphs, graphs, plan);
    if (rc)
        return rc;
    const size_t bytes = (size_t)num_graphs * sizeof(MultiGraph);
    if (!d_table || table_bytes < multi_table_bytes(num_graphs) ||
        ((uintptr_t)d_table & 255u))
        return pc_fail(PC_ENOMEM, "pc_sweep_multi_prepare: table needs %llu bytes, "
                       "256-byte aligned",
                       (unsigned long long)multi_table_bytes(num_graphs));
    PC_HIP_CHECK(hipMemcpy(d_table, plan.table.data(), bytes, hipMemcpyHostToDevice));
    return PC_OK;
}

extern "C" int pc_sweep_multi(uint32_t num_graphs, const pc_sweep_graph *graphs,
                              void *d_table, uint64_t table_bytes,
                              void *d_workspace, uint64_t workspace_bytes,
                              void *stream)
{
    g_last_error[0] = 0;
    MultiPlan plan;
    const int rc = multi_plan(num_graphs, graphs, plan);
    if (rc)
        return rc;
    if (!d_table || table_bytes < multi_table_bytes(num_graphs) ||
        ((uintptr_t)d_table & 255u))
        return pc_fail(PC_EINVAL, "pc_sweep_multi: table of %llu bytes too small "
                       "or not 256-byte aligned", (unsigned long long)table_bytes);
    const uint64_t need = (uint64_t)plan.grid * plan.slot_words * 4u;
    if (need && (!d_workspace || workspace_bytes < need))
        return pc_fail(PC_ENOMEM, "pc_sweep_multi: workspace needs %llu bytes",
                       (unsigned long long)need);
    if (plan.total == 0)
        return PC_OK;
    uint32_t *head = (uint32_t *)((char *)d_table + multi_head_offset(num_graphs));
    hipStream_t s = as_stream(stream);
    zero_u32_kernel<<<1, 1, 0, s>>>(head);
    PC_LAUNCH_CHECK("zero_u32_kernel");
    sweep_multi_kernel<<<plan.grid, 128, plan.lds, s>>>(
        (const MultiGraph *)d_table, num_graphs, (uint32_t)plan.total, head,
        (uint32_t *)d_workspace, plan.slot_words);
    PC_LAUNCH_CHECK("sweep_multi_kernel");
    return PC_OK;
}

extern "C" int pc_sweep_occupancy(uint32_t num_nodes, uint32_t state_mode,
                                  int32_t *blocks_per_cu)
{
    g_last_error[0] = 0;
    if (!blocks_per_cu)
        return pc_fail(PC_EINVAL, "pc_sweep_occupancy: NULL output");
    const uint32_t waves = state_mode & PC_SWEEP_WAVES_MASK;
    state_mode &= ~PC_SWEEP_WAVES_MASK;
    if (state_mode > PC_STATE_HBM || waves == PC_SWEEP_WAVES_MASK)
        return pc_fail(PC_EINVAL, "pc_sweep_occupancy: unknown state_mode %u",
                       state_mode | waves);
    int blocks = 0;
    const bool use_lds = state_mode == PC_STATE_LDS ||
                         (state_mode == PC_STATE_AUTO && num_nodes <= kLdsMaxNodes);
    const bool split = sweep_split(num_nodes, use_lds, waves);
    if (use_lds) {
        if (num_nodes > kLdsMaxNodes)
            return pc_fail(PC_EINVAL, "pc_sweep_occupancy: too many nodes for LDS");
        if (split)
            PC_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &blocks, sweep_lds_kernel<uint32_t, false, true, EdgeList>, 128,
                sweep_lds_bytes<true>(num_nodes)));
        else
            PC_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &blocks, sweep_lds_kernel<uint32_t, false, false, EdgeList>, 64,
                sweep_lds_bytes<false>(num_nodes)));
    } else if (split) {
        PC_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &blocks, sweep_global_kernel<uint32_t, false, true, EdgeList>, 128, 0));
    } else {
        PC_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &blocks, sweep_global_kernel<uint32_t, false, false, EdgeList>, 64, 0));
    }
    *blocks_per_cu = blocks;
    return PC_OK;
}

extern "C" int pc_window_traces(uint32_t num_nodes, uint32_t num_edges,
                                uint32_t num_runs, const void *d_records,
                                int64_t rec_stride, const uint32_t *d_span_step,
                                const uint64_t *d_chunk_state, int64_t g_lo,
                                int64_t g_len, int32_t spanning, double *d_out,
                                void *stream)
{
    g_last_error[0] = 0;
    if (num_runs == 0 || g_len <= 0)
        return PC_OK;
    if (num_nodes == 0 || num_nodes > kSumsMaxNodes)
        return pc_fail(PC_EOVERFLOW, "pc_window_traces: num_nodes outside 1..2^21");
    if (g_lo < 0 || g_lo + g_len > (int64_t)num_edges + 1)
        return pc_fail(PC_EINVAL, "pc_window_traces: window outside 0..num_edges");
    if (num_edges == 0)
        return pc_fail(PC_EINVAL, "pc_window_traces: graph without edges");
    if (!d_records || !d_span_step || !d_chunk_state || !d_out)
        return pc_fail(PC_EINVAL, "pc_window_traces: NULL buffer");
    const int64_t first_chunk = g_lo > 0 ? (g_lo - 1) / 64 : 0;
    const int64_t last_step = g_lo + g_len - 1;
    const int64_t last_chunk = last_step > 0 ? (last_step - 1) / 64 : 0;
    const dim3 grid((uint32_t)(last_chunk - first_chunk + 1),
                    (num_runs + WT_RUNS - 1u) / WT_RUNS);
    hipStream_t s = as_stream(stream);
    if (num_nodes <= kLdsMaxNodes)
        window_traces_kernel<uint32_t, false><<<grid, 64, 0, s>>>(
            num_nodes, num_edges, num_runs, (const uint32_t *)d_records,
            rec_stride, d_span_step, d_chunk_state, g_lo, g_len, spanning, d_out);
    else if (num_nodes <= kU32RecordMaxNodes)
        window_traces_kernel<uint32_t, true><<<grid, 64, 0, s>>>(
            num_nodes, num_edges, num_runs, (const uint32_t *)d_records,
            rec_stride, d_span_step, d_chunk_state, g_lo, g_len, spanning, d_out);
    else
        window_traces_kernel<uint64_t, true><<<grid, 64, 0, s>>>(
            num_nodes, num_edges, num_runs, (const uint64_t *)d_records,
            rec_stride, d_span_step, d_chunk_state, g_lo, g_len, spanning, d_out);
    PC_LAUNCH_CHECK("window_traces_kernel");
    return PC_OK;
}

extern "C" int pc_run_moments(const double *d_q, uint32_t num_p,
                              uint32_t num_runs, uint32_t cols, double *d_mean,
                              double *d_m2, void *stream)
{
    g_last_error[0] = 0;
    if (num_p == 0 || cols == 0)
        return PC_OK;
    if (num_runs == 0 || !d_q || !d_mean || !d_m2)
        return pc_fail(PC_EINVAL, "pc_run_moments: bad arguments");
    if (cols > RUN_MOMENTS_COLS)
        return pc_fail(PC_EINVAL, "pc_run_moments: at most 8 columns");
    run_moments_kernel<<<num_p, 256, 0, as_stream(stream)>>>(
        d_q, num_p, num_runs, cols, d_mean, d_m2);
    PC_LAUNCH_CHECK("run_moments_kernel");
    return PC_OK;
}

extern "C" int pc_permute_randomstate(const uint32_t *d_seeds, uint32_t num_runs,
                                      uint32_t num_edges, int32_t *d_perm,
                                      int64_t perm_stride, void *stream)
{
    g_last_error[0] = 0;
    if (num_runs == 0 || num_edges == 0)
        return PC_OK;
    if (!d_seeds || !d_perm || perm_stride < (int64_t)num_edges)
        return pc_fail(PC_EINVAL, "pc_permute_randomstate: bad buffer");
    if (num_edges > 0x7fffffffu)
        return pc_fail(PC_EOVERFLOW, "pc_permute_randomstate: num_edges >= 2^31");
    hipStream_t s = as_stream(stream);
    if (num_edges <= RS_APPLY_MAX) {
        // targets, then all swaps at once (in place in the output rows)
        randomstate_targets_kernel<<<num_runs, 64, 0, s>>>(d_seeds, num_edges, d_perm,
                                                             perm_stride);
        PC_LAUNCH_CHECK("randomstate_targets_kernel");
        randomstate_apply_kernel<<<num_runs, RS_APPLY_T, rs_apply_lds_bytes(num_edges), s>>>(
            num_edges, d_perm, perm_stride);
        PC_LAUNCH_CHECK("randomstate_apply_kernel");
        return PC_OK;
    }
    if (num_edges <= MT_LDS_MAX) {
        const size_t lds = ((size_t)num_edges * sizeof(uint16_t) + 3u) & ~(size_t)3u;
        permute_randomstate_kernel<true><<<num_runs, 64, lds, s>>>(
            d_seeds, num_edges, d_perm, perm_stride);
    } else {
        permute_randomstate_kernel<false><<<num_runs, 64, 0, s>>>(
            d_seeds, num_edges, d_perm, perm_stride);
    }
    PC_LAUNCH_CHECK("permute_randomstate_kernel");
    return PC_OK;
}
