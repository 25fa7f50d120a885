/*
 * TEST INFRASTRUCTURE ONLY.  Sequential restatement of this package's device
 * permutations (pc_permute_philox); it has no counterpart in the reference,
 * whose runs draw numpy RandomState(seed).permutation(M)
 * (percolate/hpc.py:195,206).  Only tests/ load it, as the checker.
 *
 * The algorithm is a Rao-Sandelius shuffle keyed by Philox4x32-10:
 *   1. Every element gets a uniform random bucket label from the Philox
 *      stream; elements are grouped by label, keeping their order inside a
 *      group (a stable counting sort, i.e. one radix pass on random keys).
 *   2. Every group is shuffled uniformly (Fisher-Yates with unbiased
 *      bounded draws), again by a level of random labels first when it is
 *      large.
 * Given independent uniform random words, the result is an exactly uniform
 * permutation: labelling iid and shuffling each label class uniformly is
 * the same as sorting by iid continuous keys (integer part = label,
 * fraction = the order within the class).
 *
 * Layout (the device kernels follow it bit for bit):
 *   - M <= whole (32768 on the device): one bucket, the whole run
 *     (lb1 = 0), never through the wide path below.  Otherwise 2**lb1
 *     buckets, lb1 the least with M <= bucket << lb1 (bucket 4096 on the
 *     device); label of element i = the top lb1 bits of key(1, 0, i).
 *     Elements grouped by label, stable.
 *   - Every bucket b of n elements, in place: n > group_cap (8192 on the
 *     device, in runs of several buckets): plain Fisher-Yates with 32-bit
 *     draws from ctr (u, run, b, 4 << 28).
 *     Otherwise 2**lb2 sub-buckets, lb2 the least with n <= sub << lb2
 *     (sub 64 on the device); label of position j = the top lb2 bits of
 *     key(2, b, j); stable grouping, then Fisher-Yates of every sub-bucket
 *     (start st, size c) with 16-bit draws: the draws take the halves of
 *     ctr (st, run, b, 3 << 28 | u), u = 0, 1, ... in turn, and a rejected
 *     word is replaced by the next half of the sub-bucket's second stream,
 *     ctr (st, run, b, 5 << 28 | v), v = 0, 1, ... (so every accepted draw
 *     takes exactly one word of the first stream, whatever was rejected).
 *   - key(tag, b, j) = 16-bit half (j >> 6) & 7 of
 *     Philox(ctr ((j >> 9) << 6 | (j & 63), run, b, tag << 28)), the halves
 *     of a word taken low then high.
 *   - A draw uniform on [0, r): Lemire's multiply-and-reject: m = x * r for
 *     a w-bit word x; reject while (m mod 2**w) < (2**w - r) mod r; the
 *     draw is m >> w.
 *   - Fisher-Yates: for j = c-1 down to 1, swap a[j] and a[draw on j+1].
 *   - key = (seed lo, seed hi); Philox rounds as in Salmon et al. (SC'11).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static const uint32_t PH_M0 = 0xD2511F53u, PH_M1 = 0xCD9E8D57u;
static const uint32_t PH_W0 = 0x9E3779B9u, PH_W1 = 0xBB67AE85u;

void nzo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int i = 0; i < 10; ++i) {
        if (i) {
            k0 += PH_W0;
            k1 += PH_W1;
        }
        const uint64_t p0 = (uint64_t)PH_M0 * c0;
        const uint64_t p1 = (uint64_t)PH_M1 * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0;
        c1 = (uint32_t)p1;
        c2 = n2;
        c3 = (uint32_t)p0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

struct shuffle_ctx {
    uint32_t key[2];
    uint32_t run;
};

static uint32_t half16(const uint32_t w[4], uint32_t h)
{
    return (w[h >> 1] >> (16u * (h & 1u))) & 0xFFFFu;
}

static uint32_t key16(const struct shuffle_ctx *x, uint32_t tag, uint32_t b,
                      uint64_t j)
{
    const uint32_t ctr[4] = {(uint32_t)(((j >> 9) << 6) | (j & 63u)), x->run, b,
                             tag << 28};
    uint32_t w[4];
    nzo_philox4x32_10(ctr, x->key, w);
    return half16(w, (uint32_t)((j >> 6) & 7u));
}

/* the two streams of 16-bit words of one sub-bucket: ctr[3] = tag << 28 | u */
struct draw16 {
    uint32_t ctr[4];
    uint32_t w[4], h, u;   /* first stream */
    uint32_t wr[4], hr, v; /* replacement words */
};

static uint32_t next16(const struct shuffle_ctx *x, uint32_t *ctr, uint32_t tag,
                       uint32_t *w, uint32_t *h, uint32_t *u)
{
    if (*h == 8u) {
        ctr[3] = (tag << 28) | *u;
        nzo_philox4x32_10(ctr, x->key, w);
        ++*u;
        *h = 0;
    }
    return half16(w, (*h)++);
}

static uint32_t bounded16(const struct shuffle_ctx *x, struct draw16 *d,
                          uint32_t r)
{
    uint32_t m = next16(x, d->ctr, 3u, d->w, &d->h, &d->u) * r;
    while ((m & 0xFFFFu) < r && (m & 0xFFFFu) < (65536u - r) % r)
        m = next16(x, d->ctr, 5u, d->wr, &d->hr, &d->v) * r;
    return m >> 16;
}

static void swap32(int32_t *a, int64_t i, int64_t j)
{
    const int32_t t = a[i];
    a[i] = a[j];
    a[j] = t;
}

/* Fisher-Yates over a whole bucket with 32-bit draws (buckets above the cap) */
static void shuffle_wide(const struct shuffle_ctx *x, int32_t *a, int64_t n,
                         uint32_t b)
{
    uint32_t ctr[4] = {0, x->run, b, 4u << 28};
    uint32_t w[4];
    uint32_t h = 4, u = 0;
    for (int64_t j = n - 1; j >= 1; --j) {
        const uint32_t r = (uint32_t)(j + 1);
        uint32_t k;
        for (;;) {
            if (h == 4u) {
                ctr[0] = u++;
                nzo_philox4x32_10(ctr, x->key, w);
                h = 0;
            }
            const uint64_t m = (uint64_t)w[h++] * r;
            const uint32_t lo = (uint32_t)m;
            if (lo < r && lo < (0u - r) % r)
                continue;
            k = (uint32_t)(m >> 32);
            break;
        }
        swap32(a, j, k);
    }
}

static void shuffle_group(const struct shuffle_ctx *x, int32_t *a, int64_t n,
                          uint32_t b, int64_t cap, int64_t sub)
{
    if (n <= 1)
        return;
    if (n > cap) {
        shuffle_wide(x, a, n, b);
        return;
    }
    uint32_t lb = 0;
    while ((sub << lb) < n)
        ++lb;
    const uint32_t B = 1u << lb;
    uint32_t *lab = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
    int64_t *start = (int64_t *)calloc((size_t)B + 1, sizeof(int64_t));
    int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    for (int64_t j = 0; j < n; ++j) {
        lab[j] = lb ? key16(x, 2u, b, (uint64_t)j) >> (16u - lb) : 0u;
        start[lab[j] + 1]++;
    }
    for (uint32_t s = 0; s < B; ++s)
        start[s + 1] += start[s];
    {
        int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)B);
        memcpy(fill, start, sizeof(int64_t) * B);
        for (int64_t j = 0; j < n; ++j)
            tmp[fill[lab[j]]++] = a[j];
        free(fill);
    }
    for (uint32_t s = 0; s < B; ++s) {
        const int64_t st = start[s], c = start[s + 1] - st;
        if (c < 2)
            continue;
        struct draw16 d = {{(uint32_t)st, x->run, b, 0}, {0, 0, 0, 0}, 8u, 0u,
                           {0, 0, 0, 0}, 8u, 0u};
        for (int64_t j = c - 1; j >= 1; --j)
            swap32(tmp + st, j, bounded16(x, &d, (uint32_t)(j + 1)));
    }
    memcpy(a, tmp, sizeof(int32_t) * (size_t)n);
    free(lab);
    free(start);
    free(tmp);
}

/*
 * The permutation of run `run` for num_edges edges under `seed`, into out.
 * bucket, group_cap, whole and sub are the layout's size parameters (4096,
 * 8192, 32768 and 64 on the device; tests shrink them to reach many
 * buckets, the wide Fisher-Yates path and sub-bucket labels on small
 * inputs).  Returns 0, or -1 for sizes outside the layout (more than 2**23
 * edges, more than 2**16 buckets or sub-buckets, or a group_cap or whole
 * above 2**16, where a sub-bucket could outgrow its 16-bit draws).
 */
int nzo_bucket_permutation(uint64_t seed, uint32_t run, int64_t num_edges,
                           int64_t bucket, int64_t group_cap, int64_t whole,
                           int64_t sub, int32_t *out)
{
    struct shuffle_ctx x = {{(uint32_t)seed, (uint32_t)(seed >> 32)}, run};
    const int64_t M = num_edges;
    if (M < 0 || M > (1ll << 23) || bucket < 1 || sub < 1 || group_cap > 65536 ||
        whole > 65536)
        return -1;
    uint32_t lb1 = 0;
    if (M > whole)
        while ((bucket << lb1) < M)
            ++lb1;
    if (lb1 > 16)
        return -1;
    const int64_t cap = lb1 ? group_cap : whole;
    const uint32_t B = 1u << lb1;
    uint32_t *lab = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(M ? M : 1));
    int64_t *start = (int64_t *)calloc((size_t)B + 1, sizeof(int64_t));
    for (int64_t i = 0; i < M; ++i) {
        lab[i] = lb1 ? key16(&x, 1u, 0u, (uint64_t)i) >> (16u - lb1) : 0u;
        start[lab[i] + 1]++;
    }
    for (uint32_t b = 0; b < B; ++b)
        start[b + 1] += start[b];
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)B);
    memcpy(fill, start, sizeof(int64_t) * B);
    for (int64_t i = 0; i < M; ++i)
        out[fill[lab[i]]++] = (int32_t)i;
    int rc = 0;
    for (uint32_t b = 0; b < B; ++b) {
        const int64_t n = start[b + 1] - start[b];
        if (n <= cap && n > (sub << 16))
            rc = -1;
        else
            shuffle_group(&x, out + start[b], n, b, cap, sub);
    }
    free(fill);
    free(lab);
    free(start);
    return rc;
}

/* Many runs (first_run + r, r < runs), out [runs][num_edges]; OpenMP over
 * runs. */
int nzo_bucket_permutations(uint64_t seed, uint32_t first_run, int64_t runs,
                            int64_t num_edges, int64_t bucket,
                            int64_t group_cap, int64_t whole, int64_t sub,
                            int32_t *out, int threads)
{
    int rc = 0;
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads) reduction(| : rc)
    for (int64_t r = 0; r < runs; ++r)
        rc |= nzo_bucket_permutation(seed, first_run + (uint32_t)r, num_edges,
                                     bucket, group_cap, whole, sub,
                                     out + r * num_edges);
    return rc;
}
