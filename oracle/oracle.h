/* oracle.h -- CPU oracle for the hann hot path. TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load liboracle.so, and only as the checker or the CPU baseline. The product
 * library (hann_amd) never links, loads or calls anything here.
 *
 * Parity status of each function (what pins it) is listed in oracle.c.
 */
#ifndef HANN_ORACLE_H
#define HANN_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* metric ids, shared with include/hann_amd.h */
enum { OR_SQEUCLIDEAN = 0, OR_EUCLIDEAN = 1, OR_MANHATTAN = 2, OR_COSINE = 3 };

/* build-owned random stream (restated by hann_amd/csrc/rng.h) */
uint64_t or_hash64(uint64_t x);
uint64_t or_rng_next(uint64_t *state);
uint64_t or_rng_bounded(uint64_t *state, uint64_t n);
void or_rng_perm(uint64_t *state, int64_t n, int64_t *out);
uint64_t or_stream_seed(uint64_t seed, uint64_t epoch, uint64_t role);

/* per-pair kernels, reference AVX2+FMA tier order (AVX tier for manhattan) */
float or_sqdiff(const float *a, const float *b, size_t n);
/* core/simd_kernels.inc.h:267-315 (AVX2 tier), in place */
void or_normalize(float *v, size_t n);
float or_euclidean(const float *a, const float *b, size_t n);
float or_manhattan(const float *a, const float *b, size_t n);
float or_cosine(const float *a, const float *b, size_t n);
float or_pair(int metric, const float *a, const float *b, size_t n);
void or_batch(int metric, const float *q, const float *flat, size_t dim, size_t n, double *out);

/* exact top-k by (rank, id) over nb rows for each of nq queries */
void or_batch_topk(int metric, const float *queries, size_t nq, const float *base, size_t nb,
                   size_t dim, size_t k, int nthreads, int64_t *out_ids, float *out_rank);

/* k-means pieces (pqivf/index.go:738-847) */
void or_assign(const float *data, size_t n, size_t d, const float *cents, size_t k,
               int nthreads, int32_t *out);
void or_kmeans_update(const float *data, size_t n, size_t d, const int32_t *assign, size_t k,
                      uint64_t *rng, float *cents /* in: old, out: new */);
size_t or_kmeans(const float *data, size_t n, size_t d, size_t k, int iters, uint64_t *rng,
                 int nthreads, float *out_cents /* [min(k,n)][d] */);
/* k-means++ seeding (build extension; device order of operations) and
 * runKMeans with init 0 (the reference's permutation) or 1 (k-means++) */
void or_kmeanspp_seed(const float *data, size_t n, size_t d, size_t k, uint64_t *rng, float *out);
size_t or_kmeans_init(const float *data, size_t n, size_t d, size_t k, int iters, uint64_t *rng, int init,
                      int nthreads, float *out_cents);

/* PQ encode of one vector (pqivf/index.go:641-666). codebooks: [m][pmax][s],
 * cb_len[i] rows valid in subquantizer i. returns 0, or -1 on an empty codebook */
int or_encode(const float *vec, const float *centroid, size_t d, size_t m,
              const float *codebooks, const size_t *cb_len, size_t pmax, int32_t *codes);

/* Train (pqivf/index.go:451-605) over vectors already in ascending id order.
 * Outputs: coarse [kc_out][d], assign [n], codebooks [m][pmax][s] with cb_len,
 * codes [n][m]. Returns 0 or a negative error. */
int or_train(const float *vecs, size_t n, size_t d, size_t coarse_k, size_t m, size_t pq_k,
             int iters, uint64_t seed, uint64_t epoch, int nthreads,
             float *coarse, size_t *kc_out, int32_t *assign,
             float *codebooks, size_t *cb_len, int32_t *codes);

/* Search (pqivf/index.go:850-949) of one query over an index held as CSR
 * lists. Entry e of the index: ids[e], vecs[e*d], codes[e*m], has_codes[e].
 * list_off has coarse_k+1 entries. Returns the result count; *fell_back is
 * set when the brute-force fallback ran. Results ordered by (distance, id). */
int or_search(const float *q, size_t d, const float *coarse, size_t coarse_k,
              const int64_t *list_off, const int64_t *ids, const float *vecs,
              const int32_t *codes, const uint8_t *has_codes,
              const float *codebooks, const size_t *cb_len, size_t pmax, size_t m,
              int have_codebooks, int nprobe, int allow_fallback, int k,
              int64_t *out_ids, double *out_dist, int *fell_back);

/* Search with the build's exact rerank: the rerank*k best candidates by
 * (the reference's PQ rank, storage position), then the best k of those by
 * (the reference's distance to the stored vector, id). *sel_bound: the PQ
 * rank of the last candidate kept (+inf when all are kept). */
int or_search_rerank(const float *q, size_t d, const float *coarse, size_t coarse_k,
                     const int64_t *list_off, const int64_t *ids, const float *vecs,
                     const int32_t *codes, const uint8_t *has_codes,
                     const float *codebooks, const size_t *cb_len, size_t pmax, size_t m,
                     int have_codebooks, int nprobe, int allow_fallback, int k, int rerank,
                     int64_t *out_ids, double *out_dist, int *fell_back, float *sel_bound);
/* the same for nq queries with per-query k and nprobe; outputs [nq][kstride] */
void or_search_rerank_many(const float *qs, size_t nq, size_t d, const float *coarse, size_t coarse_k,
                           const int64_t *list_off, const int64_t *ids, const float *vecs,
                           const int32_t *codes, const uint8_t *has_codes,
                           const float *codebooks, const size_t *cb_len, size_t pmax, size_t m,
                           int have_codebooks, const int32_t *nprobe_q, int allow_fallback, const int32_t *k_q,
                           int kstride, int rerank, int nthreads, int64_t *out_ids, double *out_dist,
                           int32_t *out_counts, float *out_bound);

/* many queries, T threads, one query per thread at a time (CPU baseline) */
void or_search_many(const float *qs, size_t nq, size_t d, const float *coarse, size_t coarse_k,
                    const int64_t *list_off, const int64_t *ids, const float *vecs,
                    const int32_t *codes, const uint8_t *has_codes,
                    const float *codebooks, const size_t *cb_len, size_t pmax, size_t m,
                    int have_codebooks, int nprobe, int allow_fallback, int k, int nthreads,
                    int64_t *out_ids, double *out_dist, int32_t *out_counts);

#ifdef __cplusplus
}
#endif
#endif
