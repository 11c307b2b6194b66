/* hann_amd.h -- C-ABI of the MI355X (gfx950) library behind hann's
 * batch-distance and PQIVF hot path.
 *
 * Every entry point takes plain pointers and sizes. Pointers are host
 * pointers unless the name ends in _dev, in which case they are device
 * pointers and `stream` is a hipStream_t passed as void* (NULL = the
 * library's own stream for that call). Host buffers are borrowed only for the
 * duration of the call; the library never keeps a caller pointer (the rule of
 * core/distance.go:69-72). Calls that return int return a hann_status; the
 * message of the last failure on the calling thread is hann_last_error().
 *
 * Reference interfaces replaced (habedi/hann):
 *   simd_* / hann_cpu_init          core/simd_distance.h:26-35, core/simd_ops.h:11-14
 *   Metric.DistanceBatch/RankBatch  core/metric.go:64-100
 *   pqivf.New + With* options       pqivf/index.go:56-146
 *   Index / BulkIndex / Trainer     core/index.go:14-69, pqivf/index.go:212-1101
 *   IndexStats                      core/index.go:131-139
 *   test hooks of export_test.go    pqivf/export_test.go:5-31
 */
#ifndef HANN_AMD_H
#define HANN_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status -- */
typedef enum {
    HANN_OK = 0,
    HANN_ERR_INVALID_ARG = 1,  /* validation failure (dimension, k, ids ...) */
    HANN_ERR_NOT_TRAINED = 2,  /* search on an untrained index */
    HANN_ERR_EMPTY = 3,        /* "index is empty", too few vectors to train */
    HANN_ERR_DUPLICATE_ID = 4, /* "id %d already exists ..." */
    HANN_ERR_NOT_FOUND = 5,    /* "id %d not found" */
    HANN_ERR_OOM = 6,          /* device or host allocation failed, or over budget */
    HANN_ERR_DEVICE = 7,       /* HIP runtime error, or no gfx950 device */
    HANN_ERR_FORMAT = 8,       /* bad or newer serialized index */
    HANN_ERR_UNSUPPORTED = 9   /* a parameter outside what the kernels support */
} hann_status;

/* Message of the last failing call on this thread ("" if none). */
const char *hann_last_error(void);

/* Version string of the library. */
const char *hann_version(void);

/* Select and initialise a device (replaces hann_cpu_init's role as the
 * one-time set-up, core/simd_ops.c:58-76). Idempotent. */
int hann_init(int device);

/* Drop-in for hann_cpu_init (core/simd_ops.h:11, called from
 * core/cpu_check.go:34-44). The SIMD level argument is accepted and ignored:
 * there is one code path, on the GPU. Initialises device 0. */
void hann_cpu_init(int support_level);

/* ------------------------------------------- reference distance symbols --
 * Same names and signatures as core/simd_distance.h:26-34, host pointers,
 * synchronous. Results are bit-identical to the reference's AVX2+FMA tier
 * (AVX tier for Manhattan, which that tier aliases). */
float simd_euclidean(const float *a, const float *b, size_t n);
float simd_squared_euclidean(const float *a, const float *b, size_t n);
float simd_manhattan(const float *a, const float *b, size_t n);
float simd_cosine_distance(const float *a, const float *b, size_t n);
void simd_euclidean_batch(const float *q, const float *flat, size_t dim, size_t n, double *out);
void simd_squared_euclidean_batch(const float *q, const float *flat, size_t dim, size_t n, double *out);
void simd_manhattan_batch(const float *q, const float *flat, size_t dim, size_t n, double *out);
void simd_cosine_distance_batch(const float *q, const float *flat, size_t dim, size_t n, double *out);

/* Drop-in for simd_normalize (core/simd_ops.h:14, bound by NormalizeVector,
 * core/vector_ops.go:16-22): scales vec to unit Euclidean norm in place,
 * bit-identical to the reference's AVX2+FMA tier
 * (core/simd_kernels.inc.h:267-315); a zero vector is left as it is. */
void simd_normalize(float *vec, size_t len);

/* NormalizeBatch (core/vector_ops.go:24-54) for n vectors of dim floats in
 * one flat array, in place, in one launch (build extension of the batch
 * form: the reference loops simd_normalize over worker goroutines). The
 * _dev form takes a device pointer and is asynchronous on `stream`. */
int hann_normalize_batch(float *flat, size_t n, size_t dim);
int hann_normalize_batch_dev(float *flat, size_t n, size_t dim, void *stream);

/* ------------------------------------------------------ batch distance -- */
typedef enum {
    HANN_METRIC_SQEUCLIDEAN = 0,
    HANN_METRIC_EUCLIDEAN = 1,
    HANN_METRIC_MANHATTAN = 2,
    HANN_METRIC_COSINE = 3
} hann_metric;

typedef enum {
    HANN_PREC_FP32 = 0, /* exact reference-order FP32 (VALU) */
    HANN_PREC_BF16 = 1, /* BF16 MFMA candidate pass + exact FP32 rerank */
} hann_precision;

/* Metric.DistanceBatch / RankBatch (core/metric.go:64-100) for nq queries at
 * once: out[q*nb + i] = metric(queries[q], base[i]) as the reference computes
 * it and widens to double. rank != 0 gives the rank value (squared distance
 * for Euclidean, core/metric.go:155-158). Validation as batchSize
 * (core/metric.go:104-117): dim == 0 is an error; nb == 0 is a no-op. */
int hann_distance_batch(int metric, int rank, const float *queries, size_t nq, const float *base,
                        size_t nb, size_t dim, double *out);
int hann_distance_batch_dev(int metric, int rank, const float *queries, size_t nq,
                            const float *base, size_t nb, size_t dim, double *out, void *stream);

/* Exact top-k: for each query the k base rows with the smallest rank value,
 * ordered by (rank, row index). out_ids[q*k + j] is a row index (-1 when
 * nb < k), out_dist is the true distance in double (FromRank applied).
 * precision HANN_PREC_BF16 runs an MFMA pass over BF16 copies of the inputs
 * that keeps `overfetch` candidates per query (0 = 4k), then reranks those
 * in exact FP32; it supports the (squared) Euclidean metrics. */
int hann_batch_topk(int metric, const float *queries, size_t nq, const float *base, size_t nb,
                    size_t dim, size_t k, int precision, size_t overfetch, int64_t *out_ids,
                    double *out_dist);
int hann_batch_topk_dev(int metric, const float *queries, size_t nq, const float *base, size_t nb,
                        size_t dim, size_t k, int precision, size_t overfetch, int64_t *out_ids,
                        double *out_dist, void *stream);

/* The FP32 exact form of hann_batch_topk_dev without any allocation: no
 * counterpart in the reference. It is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Its kernels keep their per-(query, row
 * range) partial lists in a static device array of 32 MiB per GPU, held
 * until the process ends, so nq * k * 12 bytes must be at most 32 MiB
 * (HANN_ERR_INVALID_ARG otherwise). Results are identical to
 * hann_batch_topk_dev with HANN_PREC_FP32. Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU (on different streams or from
 * different threads) race and must be ordered by the caller. stream NULL:
 * the device's null stream, and the call waits. */
int hann_batch_topk_noalloc_dev(int metric, const float *queries, size_t nq, const float *base, size_t nb,
                                size_t dim, size_t k, int64_t *out_ids, double *out_dist, void *stream);

/* ------------------------------------------------------------- k-means --
 * assignPoints (pqivf/index.go:738-789): out[i] = nearest centroid of row i,
 * lowest index on a tie. precision HANN_PREC_BF16 uses an MFMA pass and
 * rechecks the two best candidates in exact FP32, falling back to the exact
 * scan for any point whose margin does not certify the winner; results are
 * identical to HANN_PREC_FP32. */
int hann_assign_points(const float *data, size_t n, size_t d, const float *centroids, size_t k,
                       int precision, int32_t *out);
int hann_assign_points_dev(const float *data, size_t n, size_t d, const float *centroids, size_t k,
                           int precision, int32_t *out, void *stream);

/* runKMeans (pqivf/index.go:795-847) with the library's own random stream:
 * `seed` selects the stream, init 0 = random permutation (as the reference),
 * 1 = k-means++ (extension). Writes min(k, n) centroids and returns that
 * count in *k_out. */
int hann_kmeans(const float *data, size_t n, size_t d, size_t k, int iters, uint64_t seed,
                int init, int precision, float *out_centroids, size_t *k_out);

/* --------------------------------------------------------------- PQIVF -- */
typedef struct hann_pqivf hann_pqivf;

/* Options of pqivf.New (pqivf/index.go:56-146). A zero field takes the
 * reference default: coarse_k 16, pq_k 16, kmeans_iters 10,
 * candidate_clusters 3, num_subquantizers = largest of 8,4,2,1 dividing dim.
 * To pass an explicit non-positive value (which New rejects), set the
 * matching bit in `explicit_mask` (bit 0 coarse_k, 1 num_subquantizers,
 * 2 pq_k, 3 kmeans_iters, 4 candidate_clusters). The fallback is on unless
 * bit 5 is set, in which case bruteforce_fallback decides
 * (WithBruteForceFallback, pqivf/index.go:92-94). */
typedef struct {
    int64_t coarse_k;
    int64_t num_subquantizers;
    int64_t pq_k;
    int64_t kmeans_iters;
    int64_t candidate_clusters;
    int32_t bruteforce_fallback; /* read when explicit_mask bit 5 is set */
    uint32_t explicit_mask;
    uint64_t seed;               /* random stream seed (the HANN_SEED role, core/utils.go:12-20) */
    /* build-side extensions */
    int32_t train_precision;     /* hann_precision for k-means assignment */
    int32_t kmeans_init;         /* 0 perm (reference), 1 k-means++ */
    int64_t train_sample;        /* coarse k-means on a seeded sample of this size (0 = all) */
    int64_t pq_train_sample;     /* PQ k-means on a seeded residual sample (0 = all) */
} hann_pqivf_options;

void hann_pqivf_default_options(hann_pqivf_options *opts);

/* pqivf.New: validation order and messages as pqivf/index.go:106-146. */
int hann_pqivf_new(int64_t dim, const hann_pqivf_options *opts, hann_pqivf **out);
void hann_pqivf_free(hann_pqivf *idx);

/* Add (pqivf/index.go:212-260). `len` is the vector's length, which the
 * call validates against the index dimension. */
int hann_pqivf_add(hann_pqivf *idx, int64_t id, const float *vec, size_t len);
/* BulkAdd (266-297): vectors packed back to back in `flat`, vector i has
 * lens[i] floats (lens == NULL: every vector has dim floats). Applied in
 * ascending id order and not atomic, exactly as the reference. */
int hann_pqivf_bulk_add(hann_pqivf *idx, const int64_t *ids, const float *flat, const size_t *lens,
                        size_t n);
int hann_pqivf_delete(hann_pqivf *idx, int64_t id);                        /* 300-345 */
int hann_pqivf_bulk_delete(hann_pqivf *idx, const int64_t *ids, size_t n); /* 350-384 */
int hann_pqivf_update(hann_pqivf *idx, int64_t id, const float *vec, size_t len); /* 402-413 */
int hann_pqivf_bulk_update(hann_pqivf *idx, const int64_t *ids, const float *flat,
                           const size_t *lens, size_t n);                 /* 419-444 */
int hann_pqivf_train(hann_pqivf *idx);                                     /* 451-605 */

/* Scan modes of the inverted-list scan. */
typedef enum {
    HANN_SCAN_EXACT = 0,   /* decode + reference-order FP32 distance: bit-identical to
                              pqivf/index.go:905-936 on the AVX2 tier */
    HANN_SCAN_ADC_F32 = 1, /* FP32 ADC lookup tables */
    HANN_SCAN_ADC_F16 = 2, /* FP16 ADC lookup tables, FP32 accumulation */
    HANN_SCAN_LIST_BF16 = 3,   /* list-major scan: each probed list decoded once per batch
                                  and scored against all its queries on the matrix cores
                                  (BF16 operands, FP32 accumulation); needs rerank > 0,
                                  otherwise (and for indexes with entries lacking codes)
                                  the persistent FP32 table scan runs */
    HANN_SCAN_LIST_BF16X3 = 4, /* the same with split BF16 operands (hi + lo, three
                                  products): inner products to about FP32 accuracy */
    HANN_SCAN_LIST_FP8 = 5     /* the same with FP8 (e4m3) operands: the codebook and
                                  each query scaled by powers of two into FP8's range */
} hann_scan_mode;

typedef struct {
    int32_t scan_mode;        /* hann_scan_mode */
    int32_t coarse_precision; /* hann_precision of the coarse step: HANN_PREC_BF16 is the
                                 certified MFMA pass, whose probes and ranks equal FP32's */
    int32_t rerank;           /* 0: PQ distances (reference); R>0: rerank R*k candidates
                                 with exact FP32 distances on the stored vectors */
    int32_t fused;            /* 1 (ADC modes): the persistent scan with one FP32 table per
                                 query and a per-entry list term; scan_mode ADC_F32 and
                                 ADC_F16 both take it */
    int32_t list_config;      /* the list-major scans' work items and partitions: 0 the
                                 autotuner's choice for the batch shape, i > 0 the tuner's
                                 i-th candidate (results are identical for every choice) */
} hann_search_params;

void hann_search_default_params(hann_search_params *p);

/* Search (pqivf/index.go:850-949) of one query. Writes up to k results to
 * out_ids/out_dist in non-decreasing distance order, *out_count results.
 * Validation order and messages as the reference. */
int hann_pqivf_search(hann_pqivf *idx, const float *query, size_t len, int64_t k, int64_t *out_ids,
                      double *out_dist, int64_t *out_count);

/* Batched search: nq queries of dim floats, per-query k[q] and nprobe[q]
 * (nprobe == NULL: the index's candidate_clusters for every query).
 * Outputs are [nq][kmax] with kmax = max k; counts[q] results for query q. */
int hann_pqivf_search_batch(hann_pqivf *idx, const float *queries, size_t nq, const int32_t *k,
                            const int32_t *nprobe, const hann_search_params *params,
                            int64_t *out_ids, double *out_dist, int32_t *out_counts);
/* Same with device-resident queries, k, nprobe and outputs (out_dist in FP32
 * true distance here); asynchronous on `stream`. Each distinct stream gets
 * its own scratch workspace (so batches on different streams run on the GPU
 * at once), sized by the largest batch it has run and held until the index
 * is freed: callers should search from a fixed set of streams. NULL: the
 * index's own stream and workspace, and the call waits. */
int hann_pqivf_search_batch_dev(hann_pqivf *idx, const float *queries, size_t nq, const int32_t *k,
                                const int32_t *nprobe, int32_t kmax, int32_t max_nprobe,
                                const hann_search_params *params, int64_t *out_ids,
                                float *out_dist, int32_t *out_counts, void *stream);

/* The device search above bound to fixed buffers and replayed as one HIP
 * graph (build extension: the reference has no counterpart; it serves
 * loops that search batch after batch of the same shape, writing new
 * queries, k and nprobe into the same device buffers before each launch).
 * create records the arguments and runs nothing. Each launch searches
 * asynchronously on `stream` (NULL: the index's stream, and the launch
 * waits); the first launch, and the first after a change of the index,
 * runs the search directly and captures it, later launches replay the
 * graph. Results equal hann_pqivf_search_batch_dev's. Profiled searches,
 * per-query nprobe at or above the number of lists, and a BF16 coarse step
 * over more than 65,536 centroids run directly on every launch. A graph
 * holds its own workspace; launches of one graph are ordered on its stream,
 * and the index must outlive it. */
typedef struct hann_search_graph hann_search_graph;
int hann_pqivf_search_graph_create(hann_pqivf *idx, const float *queries, size_t nq, const int32_t *k,
                                   const int32_t *nprobe, int32_t kmax, int32_t max_nprobe,
                                   const hann_search_params *params, int64_t *out_ids, float *out_dist,
                                   int32_t *out_counts, void *stream, hann_search_graph **out);
int hann_search_graph_launch(hann_search_graph *g);
/* How often the graph was captured, replayed and run directly. */
int hann_search_graph_counts(hann_search_graph *g, int64_t *captures, int64_t *replays, int64_t *eager);
/* Waits for the graph's stream, then releases the graph and its workspace. */
void hann_search_graph_destroy(hann_search_graph *g);

/* IndexStats (core/index.go:131-139). */
typedef struct {
    int64_t count;
    int64_t dimension;
    int64_t fallback_searches;
    char distance[32]; /* "euclidean" */
} hann_index_stats;
int hann_pqivf_stats(hann_pqivf *idx, hann_index_stats *out);

/* Save / Load (pqivf/index.go:968-1101) in the library's binary format.
 * save: *buf is malloc'ed, release with hann_free_buffer. */
int hann_pqivf_save(hann_pqivf *idx, void **buf, size_t *len);
int hann_pqivf_load(hann_pqivf *idx, const void *buf, size_t len);

/* The reference's own on-disk format, Go encoding/gob: files written by
 * Go hann's Index.Save load here and files saved here load in Go hann.
 * gob_encode / gob_decode replace GobEncode / GobDecode
 * (pqivf/index.go:991-1084): one gob stream of serializedPQIVF.
 * save_gob / load_gob replace Save / Load (pqivf/index.go:1086-1100): that
 * stream wrapped as one value of the *Index GobEncoder type.
 * format_version 1 writes the current fields; 0 writes the layout from
 * before FormatVersion and NumCandidateClusters existed. Maps are written in
 * ascending key order. Decoding checks the format version and parameters
 * with the reference's messages (HANN_ERR_FORMAT) and also rejects what the
 * device index cannot hold: centroids, entries or pending vectors whose
 * length is not the dimension, codebooks not of numSubquantizers x (dim/m),
 * codes outside their codebook, lists without a centroid. Buffers from the
 * encoders are malloc'ed, release with hann_free_buffer. */
int hann_pqivf_gob_encode(hann_pqivf *idx, int format_version, void **buf, size_t *len);
int hann_pqivf_gob_decode(hann_pqivf *idx, const void *buf, size_t len);
int hann_pqivf_save_gob(hann_pqivf *idx, int format_version, void **buf, size_t *len);
int hann_pqivf_load_gob(hann_pqivf *idx, const void *buf, size_t len);
void hann_free_buffer(void *buf);

/* Stage timers of the search path (measurement; no counterpart in the
 * reference). When enabled, every search records HIP events on its stream
 * around the coarse step, the list scan and the rerank.
 * hann_pqivf_stage_times waits for them, returns the accumulated
 * milliseconds and the number of searches, and resets the totals. */
int hann_pqivf_profile(hann_pqivf *idx, int enable);
int hann_pqivf_stage_times(hann_pqivf *idx, double *coarse_ms, double *scan_ms, double *rerank_ms,
                           int64_t *calls);
/* Milliseconds spent in the main list-scan kernel of the profiled searches
 * (one launch per search: k_scan_list, k_scan_block or k_scan), from HIP
 * events on the search stream; resets the total. Call after
 * hann_pqivf_stage_times, which resolves the events (build extension). */
int hann_pqivf_scan_kernel_time(hann_pqivf *idx, double *kernel_ms);
/* The same main-kernel total (the list-major scan's filter pass) and the
 * list-major scan's sample pass, from HIP events on the search stream;
 * resets both totals (build extension). */
int hann_pqivf_scan_kernel_times(hann_pqivf *idx, double *filter_ms, double *sample_ms);
/* The scan the last search on this index ran (build extension): a
 * HANN_SCAN_* mode (the list modes only when the list-major scan ran; it
 * falls back to the FP32 table scan for shapes it does not cover) and
 * fused: 1 when the persistent table scan ran, 2 when the exact mode ran
 * the list-major exact scan (a rerank, d in {64, 96, 128}, d / m in {4, 8}),
 * 0 otherwise. */
int hann_pqivf_last_scan(hann_pqivf *idx, int *mode, int *fused);
/* The autotuner's last choice on this index as text (build extension): the
 * persistent table scan's waves per query, or a list-major scan's work items
 * and partitions, with the time that won; empty before any tuning. The
 * tuners time each candidate once on the first batch of a shape (selection
 * size k * R, nprobe bucket) and keep the fastest; results do not depend on
 * the choice. */
int hann_pqivf_last_tune(hann_pqivf *idx, char *buf, size_t len);

/* Bulk export of the clustered entries in list order (CSR): list_off
 * [k+1], ids [N], codes [N][m] (rows of -1: entry without codes), vecs
 * [N][dim], N = list_off[k]. Any output may be NULL. Build extension for
 * tools that need the whole trained state at once (what Save holds). */
int hann_pqivf_export_lists(hann_pqivf *idx, int64_t *list_off, int64_t *ids, int32_t *codes,
                            float *vecs);

/* ------------------------------------------------------------ serving --
 * A request loop over a trained, device-resident index (build extension;
 * the reference has no server). Requests of a few queries each, with their
 * own k and nprobe, are batched continuously: up to `slots` batches run on
 * the device at once, and a batch is admitted when a slot is free and its
 * device buffers fit in the HBM budget. Mutations submitted to the server
 * run in arrival order after every earlier batch and before every later one
 * (the reader/writer rule of pqivf/index.go:37). */
typedef struct hann_server hann_server;

typedef struct {
    int32_t max_batch;            /* queries per launched batch (default 4096) */
    int32_t slots;                /* batches in flight (default 2, at most 8) */
    int64_t hbm_budget;           /* bytes of batch buffers in flight (default 1 GiB) */
    hann_search_params params;    /* scan of every batch (default: the persistent ADC scan) */
} hann_server_options;

typedef struct {
    int64_t completed_queries, completed_requests, batches, mutations;
    int64_t queue_depth, in_flight, budget_in_use, budget_peak;
    /* device bytes the slots' workspaces and buffers held at most (they are
     * grow-only), and the most any batch grew its slot's workspace past the
     * scratch bound it was charged (0: every charge bounded the device
     * memory its search added) */
    int64_t held_peak, scratch_over_bound;
} hann_server_counters;

void hann_server_default_options(hann_server_options *o);
int hann_server_create(hann_pqivf *idx, const hann_server_options *o, hann_server **out);
/* drains every queued and in-flight request, then frees */
void hann_server_free(hann_server *srv);
int hann_server_submit(hann_server *srv, const float *queries, size_t nq, const int32_t *k,
                       const int32_t *nprobe, uint64_t *ticket);
int hann_server_submit_bulk_add(hann_server *srv, const int64_t *ids, const float *flat, size_t n,
                                uint64_t *ticket);
int hann_server_submit_bulk_delete(hann_server *srv, const int64_t *ids, size_t n, uint64_t *ticket);
/* Results of a finished request ([nq][kmax], kmax = the request's largest
 * k; distances are true distances). wait == 0 returns HANN_ERR_EMPTY while
 * the request runs. A ticket can be collected once. latency_ms (may be NULL)
 * is submit-to-completion time. */
int hann_server_poll(hann_server *srv, uint64_t ticket, int wait, int64_t *out_ids, double *out_dist,
                     int32_t *out_counts, double *latency_ms);
int hann_server_stats(hann_server *srv, hann_server_counters *out);
/* Open-loop load driver: request r is submitted at at_us[r] after the call
 * starts (kind 0 search of q_cnt[r] queries from q_off[r], 1 bulk add of the
 * next q_cnt[r] rows of add_ids/add_vecs, 2 bulk delete of the next q_cnt[r]
 * del_ids). Writes latencies, search results ([query][kmax_out]) and the
 * wall time of the run. */
int hann_server_replay(hann_server *srv, size_t nreq, const double *at_us, const int32_t *kind,
                       const int64_t *q_off, const int32_t *q_cnt, const float *queries, const int32_t *k,
                       const int32_t *nprobe, const int64_t *add_ids, const float *add_vecs,
                       const int64_t *del_ids, int32_t kmax_out, int64_t *out_ids, double *out_dist,
                       double *latency_ms, double *wall_s);

/* ------------------------------------------ test hooks (export_test.go) -- */
int hann_pqivf_is_trained(hann_pqivf *idx);
/* sizes: *k = coarse centroids, *m subquantizers, *p = max codebook rows */
int hann_pqivf_shape(hann_pqivf *idx, int64_t *k, int64_t *m, int64_t *p, int64_t *dim);
int hann_pqivf_coarse_centroids(hann_pqivf *idx, float *out /* [k][dim] */);
/* codebooks [m][p][dim/m]; rows beyond cb_len[i] are zero */
int hann_pqivf_codebooks(hann_pqivf *idx, float *out, int64_t *cb_len);
/* list l's ids in storage order: call with out == NULL to get the length */
int hann_pqivf_list_ids(hann_pqivf *idx, int64_t list, int64_t *out, int64_t *len);
/* list l's PQ codes [len][m] in storage order (-1 rows: entry has no codes) */
int hann_pqivf_list_codes(hann_pqivf *idx, int64_t list, int32_t *out, int64_t *len);
/* encodeVector (pqivf/index.go:669-675) */
int hann_pqivf_encode_vector(hann_pqivf *idx, const float *vec, size_t len, int64_t cluster,
                             int32_t *codes);

#ifdef __cplusplus
}
#endif
#endif
