/* oracle.c -- CPU restatement of the hann hot path. TEST INFRASTRUCTURE ONLY.
 *
 * This file is the checker for the HIP library in hann_amd/, and the CPU
 * baseline timed beside it. Only tests/, __graft_entry__.smoke() and bench.py
 * (its cpu_baseline leg) load it. The product never calls it.
 *
 * Reference files are cited as path:line relative to the habedi/hann tree.
 * Parity target: the reference's AVX2+FMA tier, which every current x86 host
 * selects (core/cpu_check.go:34-44), and its AVX tier for Manhattan, which the
 * AVX2 tier aliases (core/simd_distance.c:139-149).
 *
 * What pins each function (tests/test_oracle.py):
 *   or_sqdiff/euclidean/manhattan/cosine, or_batch, or_normalize
 *       bit-for-bit against golden vectors written by the reference's own C
 *       kernels compiled from source (tests/golden/distance_golden.npz), and
 *       against the closed forms of core/distance_test.go:14-123.
 *   or_assign, or_encode
 *       against the stored Cluster/Codes of the reference-trained fixture
 *       pqivf/testdata/index_v1.gob (tests/golden/index_v1.json), and against
 *       an FP64 brute-force argmin on inputs with separated minima.
 *   or_kmeans_update
 *       against the fixture's stored coarse centroids and codebooks, which are
 *       the float32 point-order means of their members, and against numpy's
 *       sequential float32 cumsum.
 *   or_search
 *       against FP64 brute force on the reconstructed vectors, and against
 *       the reference's behavioural tests (fallback, nprobe, empty lists).
 *   or_rng_*, or_kmeans seeding, or_train's stream layout
 *       build-owned (the reference draws from Go math/rand, whose seeding
 *       table is not available here): parity unpinned; checked only for
 *       being a permutation / in range and for determinism.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ RNG --
 * splitmix64 (Steele, Lea and Flood, "Fast splittable pseudorandom number
 * generators", OOPSLA 2014) for the stream, Lemire's multiply-high map for
 * bounded draws, and the inside-out shuffle that Go's rand.Perm uses for the
 * permutation (pqivf/index.go:805 calls it). */
uint64_t or_hash64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

uint64_t or_rng_next(uint64_t *state) {
    *state += 0x9E3779B97F4A7C15ull;
    uint64_t z = *state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t or_rng_bounded(uint64_t *state, uint64_t n) {
    return (uint64_t)(((unsigned __int128)or_rng_next(state) * n) >> 64);
}

void or_rng_perm(uint64_t *state, int64_t n, int64_t *out) {
    for (int64_t i = 0; i < n; i++) {
        int64_t j = (int64_t)or_rng_bounded(state, (uint64_t)(i + 1));
        out[i] = out[j];
        out[j] = i;
    }
}

uint64_t or_stream_seed(uint64_t seed, uint64_t epoch, uint64_t role) {
    return or_hash64(or_hash64(seed ^ or_hash64(epoch)) ^ or_hash64(role + 0x632BE59BD9B4E019ull));
}

/* ------------------------------------------------------------ distances --
 * The AVX2 instantiation of core/simd_kernels.inc.h (vocabulary
 * core/simd_kernels_avx2.inc.h:4-19): four 8-lane accumulators over blocks
 * of 32 floats, a one-vector cleanup loop into accumulator 0, the pairwise
 * combine (acc0+acc1)+(acc2+acc3), the horizontal sum of core/simd_isa.h:31-41
 * and a scalar tail that GCC contracts to a fused multiply-add (cgo builds in
 * GNU C mode, where -ffp-contract=fast is the default). Lane l of
 * accumulator a at block j holds element 32j+8a+l. */
static inline float hsum8(const float s[8]) {
    /* core/simd_isa.h:31-41: lanes 0-3 plus lanes 4-7, then the movehdup /
     * movehl pairing, gives ((s0+s4)+(s1+s5)) + ((s2+s6)+(s3+s7)). */
    float t0 = s[0] + s[4], t1 = s[1] + s[5], t2 = s[2] + s[6], t3 = s[3] + s[7];
    return (t0 + t1) + (t2 + t3);
}

float or_sqdiff(const float *a, const float *b, size_t n) {
    /* core/simd_kernels.inc.h:65-102 */
    float acc[4][8];
    memset(acc, 0, sizeof(acc));
    size_t i = 0;
    size_t lim4 = n - (n % 32);
    for (; i < lim4; i += 32)
        for (int v = 0; v < 4; v++)
            for (int l = 0; l < 8; l++) {
                float d = a[i + 8 * v + l] - b[i + 8 * v + l];
                acc[v][l] = fmaf(d, d, acc[v][l]);
            }
    size_t lim = n - (n % 8);
    for (; i < lim; i += 8)
        for (int l = 0; l < 8; l++) {
            float d = a[i + l] - b[i + l];
            acc[0][l] = fmaf(d, d, acc[0][l]);
        }
    float s[8];
    for (int l = 0; l < 8; l++) s[l] = (acc[0][l] + acc[1][l]) + (acc[2][l] + acc[3][l]);
    float sum = hsum8(s);
    for (; i < n; i++) {
        float d = a[i] - b[i];
        sum = fmaf(d, d, sum);
    }
    return sum;
}

void or_normalize(float *v, size_t n) {
    /* core/simd_kernels.inc.h:267-315 on the AVX2+FMA tier (the tier
     * hann_cpu_init(2) installs, core/simd_ops.c:58-76): the squared norm
     * in four 8-lane FMA accumulators over 32-element blocks, the 8-wide
     * cleanup in accumulator 0, (acc0+acc1)+(acc2+acc3), the horizontal sum
     * of core/simd_isa.h:31-41, a fused scalar tail; sqrtf; a zero norm
     * leaves the vector as it is; every element divided by the norm. */
    float acc[4][8];
    memset(acc, 0, sizeof(acc));
    size_t i = 0;
    size_t lim4 = n - (n % 32);
    for (; i < lim4; i += 32)
        for (int a = 0; a < 4; a++)
            for (int l = 0; l < 8; l++) acc[a][l] = fmaf(v[i + 8 * a + l], v[i + 8 * a + l], acc[a][l]);
    size_t lim = n - (n % 8);
    for (; i < lim; i += 8)
        for (int l = 0; l < 8; l++) acc[0][l] = fmaf(v[i + l], v[i + l], acc[0][l]);
    float s[8];
    for (int l = 0; l < 8; l++) s[l] = (acc[0][l] + acc[1][l]) + (acc[2][l] + acc[3][l]);
    float total = hsum8(s);
    for (; i < n; i++) total = fmaf(v[i], v[i], total);
    float norm = sqrtf(total);
    if (norm == 0.0f) return;
    for (size_t j = 0; j < n; j++) v[j] = v[j] / norm;
}

float or_euclidean(const float *a, const float *b, size_t n) {
    /* core/simd_kernels.inc.h:104-107 */
    return sqrtf(or_sqdiff(a, b, n));
}

float or_manhattan(const float *a, const float *b, size_t n) {
    /* core/simd_kernels.inc.h:116-149 instantiated with HANN_UNROLL 1
     * (core/simd_kernels_avx.inc.h:6), which the AVX2 tier calls
     * (core/simd_distance.c:139-149): one 8-lane accumulator, no FMA. */
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    size_t i = 0;
    size_t lim = n - (n % 8);
    for (; i < lim; i += 8)
        for (int l = 0; l < 8; l++) acc[l] = acc[l] + fabsf(a[i + l] - b[i + l]);
    float sum = hsum8(acc);
    for (; i < n; i++) sum += fabsf(a[i] - b[i]);
    return sum;
}

float or_cosine(const float *a, const float *b, size_t n) {
    /* core/simd_kernels.inc.h:153-229, AVX2 tier */
    float dot[4][8], na[4][8], nb[4][8];
    memset(dot, 0, sizeof(dot));
    memset(na, 0, sizeof(na));
    memset(nb, 0, sizeof(nb));
    size_t i = 0;
    size_t lim4 = n - (n % 32);
    for (; i < lim4; i += 32)
        for (int v = 0; v < 4; v++)
            for (int l = 0; l < 8; l++) {
                float x = a[i + 8 * v + l], y = b[i + 8 * v + l];
                dot[v][l] = fmaf(x, y, dot[v][l]);
                na[v][l] = fmaf(x, x, na[v][l]);
                nb[v][l] = fmaf(y, y, nb[v][l]);
            }
    size_t lim = n - (n % 8);
    for (; i < lim; i += 8)
        for (int l = 0; l < 8; l++) {
            float x = a[i + l], y = b[i + l];
            dot[0][l] = fmaf(x, y, dot[0][l]);
            na[0][l] = fmaf(x, x, na[0][l]);
            nb[0][l] = fmaf(y, y, nb[0][l]);
        }
    float sd[8], sa[8], sb[8];
    for (int l = 0; l < 8; l++) {
        sd[l] = (dot[0][l] + dot[1][l]) + (dot[2][l] + dot[3][l]);
        sa[l] = (na[0][l] + na[1][l]) + (na[2][l] + na[3][l]);
        sb[l] = (nb[0][l] + nb[1][l]) + (nb[2][l] + nb[3][l]);
    }
    float d = hsum8(sd), nra = hsum8(sa), nrb = hsum8(sb);
    for (; i < n; i++) {
        d = fmaf(a[i], b[i], d);
        nra = fmaf(a[i], a[i], nra);
        nrb = fmaf(b[i], b[i], nrb);
    }
    float normA = sqrtf(nra), normB = sqrtf(nrb);
    if (normA == 0.0f || normB == 0.0f) return 1.0f;
    float denom = normA * normB;
    float sim = d / denom;
    if (sim > 1.0f) sim = 1.0f;
    if (sim < -1.0f) sim = -1.0f;
    return 1.0f - sim;
}

float or_pair(int metric, const float *a, const float *b, size_t n) {
    switch (metric) {
    case OR_SQEUCLIDEAN: return or_sqdiff(a, b, n);
    case OR_EUCLIDEAN: return or_euclidean(a, b, n);
    case OR_MANHATTAN: return or_manhattan(a, b, n);
    default: return or_cosine(a, b, n);
    }
}

void or_batch(int metric, const float *q, const float *flat, size_t dim, size_t n, double *out) {
    /* core/simd_kernels.inc.h:235-264 */
    for (size_t i = 0; i < n; i++) out[i] = (double)or_pair(metric, q, flat + i * dim, dim);
}

/* ------------------------------------------------------------- threads -- */
typedef void (*or_range_fn)(void *ctx, size_t begin, size_t end);
typedef struct { or_range_fn fn; void *ctx; size_t begin, end; } or_job;

static void *or_job_run(void *p) {
    or_job *j = (or_job *)p;
    j->fn(j->ctx, j->begin, j->end);
    return NULL;
}

/* contiguous chunks, one per worker, like pqivf/index.go:746-781 */
static void or_parallel(size_t n, int nthreads, or_range_fn fn, void *ctx) {
    if (nthreads < 1) nthreads = 1;
    if ((size_t)nthreads > n) nthreads = n > 0 ? (int)n : 1;
    if (nthreads == 1) { fn(ctx, 0, n); return; }
    size_t chunk = (n + nthreads - 1) / nthreads;
    pthread_t th[256];
    or_job jobs[256];
    int used = 0;
    for (int w = 0; w < nthreads && w < 256; w++) {
        size_t b = (size_t)w * chunk;
        if (b >= n) break;
        size_t e = b + chunk > n ? n : b + chunk;
        jobs[used] = (or_job){fn, ctx, b, e};
        pthread_create(&th[used], NULL, or_job_run, &jobs[used]);
        used++;
    }
    for (int w = 0; w < used; w++) pthread_join(th[w], NULL);
}

/* ----------------------------------------------------------- batch top-k */
typedef struct { float r; int64_t id; } or_cand;

static int or_cand_less(or_cand a, or_cand b) { return a.r < b.r || (a.r == b.r && a.id < b.id); }

/* max-heap on (rank, id) holding the k best seen so far */
static void or_heap_sift_down(or_cand *h, size_t n, size_t i) {
    for (;;) {
        size_t l = 2 * i + 1, r = l + 1, m = i;
        if (l < n && or_cand_less(h[m], h[l])) m = l;
        if (r < n && or_cand_less(h[m], h[r])) m = r;
        if (m == i) return;
        or_cand t = h[i]; h[i] = h[m]; h[m] = t;
        i = m;
    }
}

static void or_heap_sift_up(or_cand *h, size_t i) {
    while (i > 0) {
        size_t p = (i - 1) / 2;
        if (!or_cand_less(h[p], h[i])) return;
        or_cand t = h[i]; h[i] = h[p]; h[p] = t;
        i = p;
    }
}

static int or_cand_cmp(const void *x, const void *y) {
    or_cand a = *(const or_cand *)x, b = *(const or_cand *)y;
    if (or_cand_less(a, b)) return -1;
    if (or_cand_less(b, a)) return 1;
    return 0;
}

typedef struct {
    int metric; const float *queries; const float *base; size_t nb, dim, k;
    int64_t *out_ids; float *out_rank;
} or_topk_ctx;

static void or_topk_range(void *p, size_t qb, size_t qe) {
    or_topk_ctx *c = (or_topk_ctx *)p;
    size_t k = c->k < c->nb ? c->k : c->nb;
    or_cand *h = (or_cand *)malloc(sizeof(or_cand) * (k + 1));
    int rank_metric = c->metric == OR_EUCLIDEAN ? OR_SQEUCLIDEAN : c->metric;
    for (size_t q = qb; q < qe; q++) {
        const float *qv = c->queries + q * c->dim;
        size_t hn = 0;
        for (size_t i = 0; i < c->nb; i++) {
            or_cand x = {or_pair(rank_metric, qv, c->base + i * c->dim, c->dim), (int64_t)i};
            if (hn < k) { h[hn] = x; or_heap_sift_up(h, hn); hn++; }
            else if (k > 0 && or_cand_less(x, h[0])) { h[0] = x; or_heap_sift_down(h, hn, 0); }
        }
        qsort(h, hn, sizeof(or_cand), or_cand_cmp);
        for (size_t j = 0; j < c->k; j++) {
            c->out_ids[q * c->k + j] = j < hn ? h[j].id : -1;
            c->out_rank[q * c->k + j] = j < hn ? h[j].r : INFINITY;
        }
    }
    free(h);
}

void or_batch_topk(int metric, const float *queries, size_t nq, const float *base, size_t nb,
                   size_t dim, size_t k, int nthreads, int64_t *out_ids, float *out_rank) {
    or_topk_ctx c = {metric, queries, base, nb, dim, k, out_ids, out_rank};
    or_parallel(nq, nthreads, or_topk_range, &c);
}

/* --------------------------------------------------------------- k-means */
typedef struct { const float *data; size_t d; const float *cents; size_t k; int32_t *out; } or_assign_ctx;

static void or_assign_range(void *p, size_t b, size_t e) {
    or_assign_ctx *c = (or_assign_ctx *)p;
    for (size_t i = b; i < e; i++) {
        /* pqivf/index.go:768-778: one batch of rank distances, then an argmin
         * with a strict <, so the lowest centroid index wins a tie. */
        const float *x = c->data + i * c->d;
        int32_t best = 0;
        float bd = or_sqdiff(x, c->cents, c->d);
        for (size_t j = 1; j < c->k; j++) {
            float dj = or_sqdiff(x, c->cents + j * c->d, c->d);
            if (dj < bd) { bd = dj; best = (int32_t)j; }
        }
        c->out[i] = best;
    }
}

void or_assign(const float *data, size_t n, size_t d, const float *cents, size_t k,
               int nthreads, int32_t *out) {
    /* pqivf/index.go:738-789 */
    or_assign_ctx c = {data, d, cents, k, out};
    or_parallel(n, nthreads, or_assign_range, &c);
}

void or_kmeans_update(const float *data, size_t n, size_t d, const int32_t *assign, size_t k,
                      uint64_t *rng, float *cents) {
    /* pqivf/index.go:816-844: float32 sums in point order, then a float32
     * division by the count; an empty cluster takes a random data point. */
    float *sums = (float *)calloc(k * d, sizeof(float));
    int64_t *counts = (int64_t *)calloc(k, sizeof(int64_t));
    for (size_t p = 0; p < n; p++) {
        int32_t c = assign[p];
        counts[c]++;
        float *s = sums + (size_t)c * d;
        const float *x = data + p * d;
        for (size_t j = 0; j < d; j++) s[j] = s[j] + x[j];
    }
    for (size_t i = 0; i < k; i++) {
        if (counts[i] == 0) {
            uint64_t idx = or_rng_bounded(rng, n);
            memcpy(cents + i * d, data + idx * d, d * sizeof(float));
        } else {
            float cnt = (float)counts[i];
            for (size_t j = 0; j < d; j++) cents[i * d + j] = sums[i * d + j] / cnt;
        }
    }
    free(sums);
    free(counts);
}

size_t or_kmeans(const float *data, size_t n, size_t d, size_t k, int iters, uint64_t *rng,
                 int nthreads, float *out_cents) {
    /* pqivf/index.go:795-847 */
    if (n == 0) return 0;
    if (n < k) k = n;
    int64_t *perm = (int64_t *)malloc(sizeof(int64_t) * n);
    or_rng_perm(rng, (int64_t)n, perm);
    for (size_t i = 0; i < k; i++) memcpy(out_cents + i * d, data + perm[i] * d, d * sizeof(float));
    free(perm);
    int32_t *assign = (int32_t *)malloc(sizeof(int32_t) * n);
    for (int it = 0; it < iters; it++) {
        or_assign(data, n, d, out_cents, k, nthreads, assign);
        or_kmeans_update(data, n, d, assign, k, rng, out_cents);
    }
    free(assign);
    return k;
}

/* k-means++ seeding (a build extension: the reference seeds with a
 * permutation). A restatement of the device's order of operations, so that
 * seeds compare bit for bit: D^2 in the reference FP32 order (NaN as +inf),
 * per-256-row FP64 tree sums, 1024 sequential chunk sums, a Hillis-Steele
 * inclusive scan over the chunks, and a walk of the chunk, blocks and rows
 * that hold u * total. */
static float or_sane(float v) { return v != v ? INFINITY : v; }

void or_kmeanspp_seed(const float *data, size_t n, size_t d, size_t k, uint64_t *rng, float *out) {
    if (n == 0 || k == 0) return;
    const size_t first = (size_t)or_rng_bounded(rng, (uint64_t)n);
    double *u = (double *)malloc(sizeof(double) * (k > 1 ? k - 1 : 1));
    for (size_t j = 0; j + 1 < k; j++) u[j] = (double)(or_rng_next(rng) >> 11) * (1.0 / 9007199254740992.0);
    memcpy(out, data + first * d, d * sizeof(float));
    const size_t B = 256, nb = (n + B - 1) / B, per = (nb + 1023) / 1024;
    float *D = (float *)malloc(sizeof(float) * n);
    double *bsum = (double *)malloc(sizeof(double) * nb);
    double s[1024], blk[256];
    for (size_t j = 1; j < k; j++) {
        const float *c = out + (j - 1) * d;
        for (size_t i = 0; i < n; i++) {
            const float e = or_sane(or_sqdiff(data + i * d, c, d));
            D[i] = j == 1 ? e : (e < D[i] ? e : D[i]);
        }
        for (size_t b = 0; b < nb; b++) {
            for (size_t t = 0; t < B; t++) blk[t] = b * B + t < n ? (double)D[b * B + t] : 0.0;
            for (size_t o = B / 2; o > 0; o >>= 1)
                for (size_t t = 0; t < o; t++) blk[t] += blk[t + o];
            bsum[b] = blk[0];
        }
        for (size_t t = 0; t < 1024; t++) {
            const size_t b0 = t * per, b1 = b0 + per < nb ? b0 + per : nb;
            double m = 0.0;
            for (size_t b = b0; b < b1; b++) m += bsum[b];
            s[t] = m;
        }
        for (size_t o = 1; o < 1024; o <<= 1) {
            double v[1024];
            for (size_t t = 0; t < 1024; t++) v[t] = t >= o ? s[t - o] : 0.0;
            for (size_t t = 0; t < 1024; t++) s[t] += v[t];
        }
        const double total = s[1023];
        int64_t pick = -1;
        if (!(total > 0.0)) {
            int64_t r = (int64_t)(u[j - 1] * (double)n);
            pick = r < (int64_t)n ? r : (int64_t)n - 1;
        } else {
            const double target = u[j - 1] * total;
            for (size_t t = 0; t < 1024 && pick < 0; t++) {
                const double excl = t > 0 ? s[t - 1] : 0.0;
                if (!(excl <= target && target < s[t])) continue;
                const size_t b0 = t * per, b1 = b0 + per < nb ? b0 + per : nb;
                double run = excl;
                int64_t found = -1, last_pos = -1;
                for (size_t b = b0; b < b1 && found < 0; b++) {
                    if (run + bsum[b] <= target) {
                        run += bsum[b];
                        continue;
                    }
                    const size_t r1 = (b + 1) * B < n ? (b + 1) * B : n;
                    for (size_t r = b * B; r < r1; r++) {
                        if (D[r] > 0.f) last_pos = (int64_t)r;
                        run += (double)D[r];
                        if (run > target && D[r] > 0.f) {
                            found = (int64_t)r;
                            break;
                        }
                    }
                }
                pick = found >= 0 ? found : last_pos;
                if (pick < 0) pick = -2;   /* the claiming chunk found no row */
            }
            if (pick < 0) {
                pick = -1;
                for (int64_t r = (int64_t)n - 1; r >= 0; r--)
                    if (D[r] > 0.f) {
                        pick = r;
                        break;
                    }
                if (pick < 0) pick = 0;
            }
        }
        memcpy(out + j * d, data + (size_t)pick * d, d * sizeof(float));
    }
    free(D);
    free(bsum);
    free(u);
}

size_t or_kmeans_init(const float *data, size_t n, size_t d, size_t k, int iters, uint64_t *rng, int init,
                      int nthreads, float *out_cents) {
    if (init == 0) return or_kmeans(data, n, d, k, iters, rng, nthreads, out_cents);
    if (n == 0) return 0;
    if (n < k) k = n;
    or_kmeanspp_seed(data, n, d, k, rng, out_cents);
    int32_t *assign = (int32_t *)malloc(sizeof(int32_t) * n);
    for (int it = 0; it < iters; it++) {
        or_assign(data, n, d, out_cents, k, nthreads, assign);
        or_kmeans_update(data, n, d, assign, k, rng, out_cents);
    }
    free(assign);
    return k;
}

/* -------------------------------------------------------------------- PQ */
int or_encode(const float *vec, const float *centroid, size_t d, size_t m,
              const float *codebooks, const size_t *cb_len, size_t pmax, int32_t *codes) {
    /* pqivf/index.go:641-666: residual in float32, contiguous sub-vectors
     * (splitVector, 716-728), per-subquantizer argmin with a strict <. */
    size_t s = d / m;
    float res[4096];
    for (size_t j = 0; j < d; j++) res[j] = vec[j] - centroid[j];
    for (size_t i = 0; i < m; i++) {
        size_t P = cb_len[i];
        if (P == 0) return -1;
        const float *cb = codebooks + i * pmax * s;
        int32_t best = 0;
        float bd = or_sqdiff(res + i * s, cb, s);
        for (size_t j = 1; j < P; j++) {
            float dj = or_sqdiff(res + i * s, cb + j * s, s);
            if (dj < bd) { bd = dj; best = (int32_t)j; }
        }
        codes[i] = best;
    }
    return 0;
}

int or_train(const float *vecs, size_t n, size_t d, size_t coarse_k, size_t m, size_t pq_k,
             int iters, uint64_t seed, uint64_t epoch, int nthreads,
             float *coarse, size_t *kc_out, int32_t *assign,
             float *codebooks, size_t *cb_len, int32_t *codes) {
    /* pqivf/index.go:451-605, over vectors already in ascending id order */
    if (n < coarse_k) return -1;
    uint64_t rng = or_stream_seed(seed, epoch, 0);
    size_t kc = or_kmeans(vecs, n, d, coarse_k, iters, &rng, nthreads, coarse);
    *kc_out = kc;
    or_assign(vecs, n, d, coarse, kc, nthreads, assign);
    /* lists in ascending id order; clusters visited in ascending order */
    int64_t *off = (int64_t *)calloc(kc + 1, sizeof(int64_t));
    for (size_t i = 0; i < n; i++) off[assign[i] + 1]++;
    for (size_t c = 0; c < kc; c++) off[c + 1] += off[c];
    int64_t *order = (int64_t *)malloc(sizeof(int64_t) * n);
    int64_t *cur = (int64_t *)malloc(sizeof(int64_t) * kc);
    memcpy(cur, off, sizeof(int64_t) * kc);
    for (size_t i = 0; i < n; i++) order[cur[assign[i]]++] = (int64_t)i;
    size_t s = d / m;
    float *sub = (float *)malloc(sizeof(float) * n * s);
    for (size_t i = 0; i < m; i++) {
        /* dataPerSub[i] in cluster-ascending, list order (pqivf/index.go:520-541) */
        for (size_t r = 0; r < n; r++) {
            const float *x = vecs + order[r] * d;
            const float *c = coarse + (size_t)assign[order[r]] * d;
            for (size_t j = 0; j < s; j++) sub[r * s + j] = x[i * s + j] - c[i * s + j];
        }
        uint64_t srng = or_stream_seed(seed, epoch, 1 + i);
        cb_len[i] = or_kmeans(sub, n, s, pq_k, iters, &srng, nthreads, codebooks + i * pq_k * s);
    }
    free(sub);
    for (size_t i = 0; i < n; i++)
        if (or_encode(vecs + i * d, coarse + (size_t)assign[i] * d, d, m, codebooks, cb_len, pq_k,
                      codes + i * m) != 0) {
            free(off); free(order); free(cur);
            return -2;
        }
    free(off);
    free(order);
    free(cur);
    return 0;
}

/* ---------------------------------------------------------------- search */
typedef struct { float r; int64_t c; } or_cent;

static int or_cent_cmp(const void *x, const void *y) {
    const or_cent *a = (const or_cent *)x, *b = (const or_cent *)y;
    if (a->r < b->r) return -1;
    if (a->r > b->r) return 1;
    return a->c < b->c ? -1 : (a->c > b->c ? 1 : 0);
}

/* The reference's rank of entry e of list c for query q: decodePQCode
 * (678-690) then vectorAdd (705-714) and the squared distance in the
 * reference's order, or the distance to the stored vector when the entry has
 * no valid codes (pqivf/index.go:905-936). `approx` holds d floats. */
static float or_entry_rank(const float *q, size_t d, const float *coarse, int64_t c, int64_t e,
                           const float *vecs, const int32_t *codes, const uint8_t *has_codes,
                           const float *codebooks, const size_t *cb_len, size_t pmax, size_t m,
                           int have_codebooks, float *approx) {
    size_t s = m ? d / m : 0;
    int use_codes = have_codebooks && has_codes[e];
    if (use_codes) {
        const int32_t *ce = codes + e * m;
        for (size_t i = 0; i < m && use_codes; i++)
            if (ce[i] < 0 || (size_t)ce[i] >= cb_len[i]) use_codes = 0;
        if (use_codes) {
            const float *cent = coarse + c * d;
            for (size_t i = 0; i < m; i++) {
                const float *cb = codebooks + (i * pmax + (size_t)ce[i]) * s;
                for (size_t j = 0; j < s; j++) approx[i * s + j] = cent[i * s + j] + cb[j];
            }
        }
    }
    return use_codes ? or_sqdiff(q, approx, d) : or_sqdiff(q, vecs + e * d, d);
}

/* the probed lists of q: nearestCentroids (pqivf/index.go:184-206) sorts by
 * rank with an unstable sort; ties here go to the lower cluster index */
static or_cent *or_probe_order(const float *q, size_t d, const float *coarse, size_t coarse_k) {
    or_cent *cs = (or_cent *)malloc(sizeof(or_cent) * (coarse_k ? coarse_k : 1));
    for (size_t c = 0; c < coarse_k; c++) cs[c] = (or_cent){or_sqdiff(q, coarse + c * d, d), (int64_t)c};
    qsort(cs, coarse_k, sizeof(or_cent), or_cent_cmp);
    return cs;
}

int or_search(const float *q, size_t d, const float *coarse, size_t coarse_k,
              const int64_t *list_off, const int64_t *ids, const float *vecs,
              const int32_t *codes, const uint8_t *has_codes,
              const float *codebooks, const size_t *cb_len, size_t pmax, size_t m,
              int have_codebooks, int nprobe, int allow_fallback, int k,
              int64_t *out_ids, double *out_dist, int *fell_back) {
    /* pqivf/index.go:850-949 */
    *fell_back = 0;
    if (k <= 0) return 0;
    or_cent *cs = or_probe_order(q, d, coarse, coarse_k);
    size_t np = (size_t)nprobe < coarse_k ? (size_t)nprobe : coarse_k;
    size_t ncand = 0;
    for (size_t i = 0; i < np; i++) ncand += list_off[cs[i].c + 1] - list_off[cs[i].c];
    int all = 0;
    if (ncand < (size_t)k && allow_fallback) { all = 1; *fell_back = 1; }
    size_t total = all ? (size_t)list_off[coarse_k] : ncand;
    or_cand *cand = (or_cand *)malloc(sizeof(or_cand) * (total ? total : 1));
    float approx[4096];
    size_t nc = 0;
    size_t nl = all ? coarse_k : np;
    for (size_t li = 0; li < nl; li++) {
        int64_t c = all ? (int64_t)li : cs[li].c;
        for (int64_t e = list_off[c]; e < list_off[c + 1]; e++) {
            float r = or_entry_rank(q, d, coarse, c, e, vecs, codes, has_codes, codebooks, cb_len, pmax, m,
                                    have_codebooks, approx);
            cand[nc++] = (or_cand){r, ids[e]};
        }
    }
    qsort(cand, nc, sizeof(or_cand), or_cand_cmp);
    int kk = (size_t)k < nc ? k : (int)nc;
    for (int i = 0; i < kk; i++) {
        out_ids[i] = cand[i].id;
        out_dist[i] = sqrt((double)cand[i].r); /* FromRank, core/metric.go:156 */
    }
    free(cand);
    free(cs);
    return kk;
}

/* candidates of the rerank: (PQ rank, storage position) */
typedef struct { float r; int64_t e; } or_pos;

static int or_pos_cmp(const void *x, const void *y) {
    const or_pos *a = (const or_pos *)x, *b = (const or_pos *)y;
    if (a->r < b->r) return -1;
    if (a->r > b->r) return 1;
    return a->e < b->e ? -1 : (a->e > b->e ? 1 : 0);
}

int or_search_rerank(const float *q, size_t d, const float *coarse, size_t coarse_k,
                     const int64_t *list_off, const int64_t *ids, const float *vecs,
                     const int32_t *codes, const uint8_t *has_codes,
                     const float *codebooks, const size_t *cb_len, size_t pmax, size_t m,
                     int have_codebooks, int nprobe, int allow_fallback, int k, int rerank,
                     int64_t *out_ids, double *out_dist, int *fell_back, float *sel_bound) {
    /* Search (pqivf/index.go:850-949) up to its candidate ranks: the probed
     * lists, the fallback rule (fewer than k probed entries: every list) and
     * the reference's PQ rank of every candidate. Then the build's rerank
     * (no counterpart in the reference): the rerank*k best candidates by
     * (PQ rank, storage position) -- the position is the entry's index in
     * the CSR lists, lists in ascending cluster order, each in storage
     * order -- are ranked by the reference's distance to the stored vector
     * (the rank of an entry without codes, pqivf/index.go:928), and the best
     * k by (rank, id) are returned with FromRank applied. *sel_bound is the
     * PQ rank of the last candidate kept (+inf when every candidate is
     * kept). */
    *fell_back = 0;
    *sel_bound = INFINITY;
    if (k <= 0 || rerank <= 0) return 0;
    or_cent *cs = or_probe_order(q, d, coarse, coarse_k);
    size_t np = (size_t)nprobe < coarse_k ? (size_t)nprobe : coarse_k;
    size_t ncand = 0;
    for (size_t i = 0; i < np; i++) ncand += list_off[cs[i].c + 1] - list_off[cs[i].c];
    int all = 0;
    if (ncand < (size_t)k && allow_fallback) { all = 1; *fell_back = 1; }
    size_t total = all ? (size_t)list_off[coarse_k] : ncand;
    or_pos *cand = (or_pos *)malloc(sizeof(or_pos) * (total ? total : 1));
    float approx[4096];
    size_t nc = 0;
    size_t nl = all ? coarse_k : np;
    for (size_t li = 0; li < nl; li++) {
        int64_t c = all ? (int64_t)li : cs[li].c;
        for (int64_t e = list_off[c]; e < list_off[c + 1]; e++) {
            float r = or_entry_rank(q, d, coarse, c, e, vecs, codes, has_codes, codebooks, cb_len, pmax, m,
                                    have_codebooks, approx);
            cand[nc++] = (or_pos){r, e};
        }
    }
    qsort(cand, nc, sizeof(or_pos), or_pos_cmp);
    size_t sel = (size_t)k * (size_t)rerank;
    if (sel < nc) *sel_bound = cand[sel - 1].r;
    else sel = nc;
    or_cand *rr = (or_cand *)malloc(sizeof(or_cand) * (sel ? sel : 1));
    for (size_t i = 0; i < sel; i++) rr[i] = (or_cand){or_sqdiff(q, vecs + cand[i].e * d, d), ids[cand[i].e]};
    qsort(rr, sel, sizeof(or_cand), or_cand_cmp);
    int kk = (size_t)k < sel ? k : (int)sel;
    for (int i = 0; i < kk; i++) {
        out_ids[i] = rr[i].id;
        out_dist[i] = sqrt((double)rr[i].r);
    }
    free(rr);
    free(cand);
    free(cs);
    return kk;
}

typedef struct {
    const float *qs; size_t d; const float *coarse; size_t coarse_k; const int64_t *list_off;
    const int64_t *ids; const float *vecs; const int32_t *codes; const uint8_t *has_codes;
    const float *codebooks; const size_t *cb_len; size_t pmax, m; int have_cb, nprobe, fb, k;
    int64_t *out_ids; double *out_dist; int32_t *out_counts;
} or_search_ctx;

static void or_search_range(void *p, size_t b, size_t e) {
    or_search_ctx *c = (or_search_ctx *)p;
    for (size_t q = b; q < e; q++) {
        int fell;
        c->out_counts[q] = or_search(c->qs + q * c->d, c->d, c->coarse, c->coarse_k, c->list_off,
                                     c->ids, c->vecs, c->codes, c->has_codes, c->codebooks,
                                     c->cb_len, c->pmax, c->m, c->have_cb, c->nprobe, c->fb, c->k,
                                     c->out_ids + q * c->k, c->out_dist + q * c->k, &fell);
    }
}

typedef struct {
    or_search_ctx base; int rerank; const int32_t *k_q; const int32_t *np_q; int kstride; float *bound;
} or_rr_ctx;

static void or_rr_range(void *p, size_t b, size_t e) {
    or_rr_ctx *r = (or_rr_ctx *)p;
    or_search_ctx *c = &r->base;
    for (size_t q = b; q < e; q++) {
        int fell;
        int kq = r->k_q ? r->k_q[q] : c->k;
        int npq = r->np_q ? r->np_q[q] : c->nprobe;
        c->out_counts[q] = or_search_rerank(c->qs + q * c->d, c->d, c->coarse, c->coarse_k, c->list_off,
                                            c->ids, c->vecs, c->codes, c->has_codes, c->codebooks, c->cb_len,
                                            c->pmax, c->m, c->have_cb, npq, c->fb, kq, r->rerank,
                                            c->out_ids + q * r->kstride, c->out_dist + q * r->kstride, &fell,
                                            r->bound + q);
    }
}

void or_search_rerank_many(const float *qs, size_t nq, size_t d, const float *coarse, size_t coarse_k,
                           const int64_t *list_off, const int64_t *ids, const float *vecs,
                           const int32_t *codes, const uint8_t *has_codes,
                           const float *codebooks, const size_t *cb_len, size_t pmax, size_t m,
                           int have_codebooks, const int32_t *nprobe_q, int allow_fallback, const int32_t *k_q,
                           int kstride, int rerank, int nthreads, int64_t *out_ids, double *out_dist,
                           int32_t *out_counts, float *out_bound) {
    or_rr_ctx c = {{qs, d, coarse, coarse_k, list_off, ids, vecs, codes, has_codes, codebooks, cb_len, pmax, m,
                    have_codebooks, 0, allow_fallback, kstride, out_ids, out_dist, out_counts},
                   rerank, k_q, nprobe_q, kstride, out_bound};
    or_parallel(nq, nthreads, or_rr_range, &c);
}

void or_search_many(const float *qs, size_t nq, size_t d, const float *coarse, size_t coarse_k,
                    const int64_t *list_off, const int64_t *ids, const float *vecs,
                    const int32_t *codes, const uint8_t *has_codes,
                    const float *codebooks, const size_t *cb_len, size_t pmax, size_t m,
                    int have_codebooks, int nprobe, int allow_fallback, int k, int nthreads,
                    int64_t *out_ids, double *out_dist, int32_t *out_counts) {
    or_search_ctx c = {qs, d, coarse, coarse_k, list_off, ids, vecs, codes, has_codes, codebooks,
                       cb_len, pmax, m, have_codebooks, nprobe, allow_fallback, k,
                       out_ids, out_dist, out_counts};
    or_parallel(nq, nthreads, or_search_range, &c);
}
