// batch.hip -- batch top-k dispatch: exact FP32, or the certified BF16
// MFMA candidate pass followed by an exact FP32 ranking.
#include <cmath>
#include <algorithm>
#include <vector>

#include "batch.h"
#include "mfma.h"
#include "scan_list.h"

namespace hann {

// bases up to this many rows: overflowed queries are ranked over every row
// by the candidate-ranking kernel
constexpr int64_t kRankEveryMax = 65536;

int64_t certified_no_sync_max_rows() { return kRankEveryMax; }

// The row sample of the certified pass: every stride-th row. A query's
// threshold bounds about its (k stride)-th distance, so the candidates it
// ranks exactly number about k nb / ns; the sample costs nq ns distances
// (an MFMA pass that writes them, then a per-query threshold pass that reads
// them), a candidate about eight times one sample distance (a random row
// read and its exact rank). The total is least near ns = sqrt(8 k nb): 4,096
// rows for 16,384 centroids and k = 128, about 28,000 for a million rows and
// k = 100. At least 4 k sample rows.
static int64_t sample_stride(int64_t nb, int k) {
    const double ns = std::sqrt(8.0 * (double)k * (double)nb);
    const int64_t want = std::max<int64_t>(4 * (int64_t)k, (int64_t)ns);
    return std::max<int64_t>(1, nb / std::max<int64_t>(1, want));
}

// The dense form of the certified pass, for bases of at most kRankEveryMax
// rows (a multiple of 8) and widths up to 128 (a padded width the
// register-resident MFMA kernel takes): every (query, row) approximate
// distance is written as a 16-bit key, and each query's threshold comes from
// the k-th smallest of all its keys, not of a sample, so no second threshold
// and no candidate-slot sizing is needed; queries run in chunks of at most
// 256 MB of keys.
static bool dense_path(int64_t nb, int dp, bool hinted) {
    return !hinted && nb <= kRankEveryMax && nb <= dense_select_max_rows() && nb % 8 == 0 && dp <= 128;
}
static int64_t dense_qchunk(int64_t nq, int64_t nb) {
    return std::max<int64_t>(1, std::min<int64_t>(nq, (int64_t)(256ll << 20) / (2 * nb)));
}
static int dense_cap(int64_t nb, int k) { return (int)std::min<int64_t>(nb, std::max<int64_t>(256, 8ll * k)); }

__global__ void k_copy_u32(unsigned *__restrict__ dst, const unsigned *__restrict__ src, int n) {
    const int i = threadIdx.x;
    if (i < n) dst[i] = src[i];
}

size_t certified_topk_scratch_bytes(int64_t nq, int64_t nb, int dim, int k, bool cached) {
    // the buffers certified_topk takes, by the same sizing rules (256 bytes
    // of slack each, as the workspace adds)
    const int64_t dp = ceil_div(dim, 32) * 32;
    size_t b = 0;
    auto add = [&](int64_t bytes) { b += (size_t)bytes + 256; };
    add(16);
    if (!cached) {
        add(nb * dp * 2);
        add(nb * 4);
        add(nb * 4);
    }
    add(nq * dp * 2);
    add(nq * 4);
    add(nq * 4);
    add(nq * 4);
    const int64_t stride = sample_stride(nb, k);
    const int64_t ns = ceil_div(nb, stride);
    add(ns * dp * 2);
    add(ns * 4);
    const int64_t qchunk = std::max<int64_t>(1, std::min<int64_t>(nq, (int64_t)(256ll << 20) / (4 * ns)));
    add(std::max<int64_t>(qchunk * ns * 4, nq * 8));
    const int64_t expect = std::max<int64_t>(k, ceil_div((int64_t)k * nb, ns));
    const int64_t cap = std::min<int64_t>(nb, std::max<int64_t>(256, 8 * expect));
    const int64_t chunk = std::max<int64_t>(128, std::min<int64_t>(nq, (int64_t)(512ll << 20) / (4ll * cap)));
    add(nq * 4);
    add(chunk * cap * 4);
    if (mfma_filter_writes_approx((int)dp)) {
        add(chunk * cap * 4);   // the candidates' approximate distances
        add(nq * 4);            // and the second thresholds
    }
    if (dense_path(nb, (int)dp, false)) {
        // the dense form: its keys in slot 10 and its candidates in slot 13
        // can be larger than the sampled form's buffers there
        const int64_t qc = dense_qchunk(nq, nb);
        b += (size_t)std::max<int64_t>(0, qc * nb * 2 - std::max<int64_t>(qchunk * ns * 4, nq * 8));
        b += (size_t)std::max<int64_t>(0, qc * (int64_t)dense_cap(nb, k) * 4 - chunk * cap * 4);
        b += (size_t)std::max<int64_t>(0, qc * 8 - nq * 4);   // slot 11: the key ranges
    }
    return b;
}

// Certified BF16 path for the (squared) Euclidean metrics. Results equal the
// exact path's: the filter keeps every row whose exact FP32 rank can be at
// most the query's sample k-th rank, and those rows are ranked exactly.
// A query whose candidate list overflows is ranked over every row (small
// bases, on the device) or rerun on the tiled exact path.
void certified_topk(Context &c, const float *queries, int64_t nq, const float *base, int64_t nb, int dim, int k,
                    const int32_t *hint, int sqrt_out, int64_t *out_ids, double *out_dist, hipStream_t st,
                    const int32_t *k_q, BF16Base *cache, Workspace *ws) {
    Workspace &w = ws ? *ws : c.ws_bf16;
    const int dp = (int)ceil_div(dim, 32) * 32;
    const int metric = sqrt_out ? HANN_METRIC_EUCLIDEAN : HANN_METRIC_SQEUCLIDEAN;
    // BF16 copies, norms, max norm and exactness flag of the base (cached)
    // and of the queries
    __bf16 *xb;
    float *xn, *xf;
    unsigned *stat = w.get<unsigned>(3, 4);   // [0] max norm bits, [1] inexact flag, [2] queries' max norm
    if (cache) {
        if (!cache->ready) {
            cache->xb.resize((size_t)nb * dp);
            cache->xn.resize((size_t)nb);
            cache->xf.resize((size_t)nb);
            cache->stat.resize(4);
            HANN_HIP(hipMemsetAsync(cache->stat.get(), 0, 4 * sizeof(unsigned), st));
            launch_to_bf16(base, nb, dim, dp, cache->xb.get(), cache->xn.get(), cache->xf.get(), cache->stat.get(),
                           reinterpret_cast<int *>(cache->stat.get() + 1), st);
            // other streams may read the cache once this call returns
            HANN_HIP(hipStreamSynchronize(st));
            cache->ready = true;
        }
        xb = cache->xb.get();
        xn = cache->xn.get();
        xf = cache->xf.get();
        // a copy kernel, not a memcpy: the search may be captured in a graph
        hipLaunchKernelGGL(k_copy_u32, dim3(1), dim3(64), 0, st, stat, cache->stat.get(), 4);
        HANN_LAUNCH_CHECK();
    } else {
        xb = w.get<__bf16>(0, (size_t)nb * dp);
        xn = w.get<float>(1, (size_t)nb);
        xf = w.get<float>(2, (size_t)nb);
        HANN_HIP(hipMemsetAsync(stat, 0, 4 * sizeof(unsigned), st));
        launch_to_bf16(base, nb, dim, dp, xb, xn, xf, stat, reinterpret_cast<int *>(stat + 1), st);
    }
    (void)xf;
    __bf16 *qb = w.get<__bf16>(4, (size_t)nq * dp);
    float *qn = w.get<float>(5, (size_t)nq);
    float *qf = w.get<float>(6, (size_t)nq);
    launch_to_bf16(queries, nq, dim, dp, qb, qn, qf, stat + 2, reinterpret_cast<int *>(stat + 1), st);
    if (dense_path(nb, dp, hint && k == 1)) {
        const int64_t qchunk = dense_qchunk(nq, nb);
        const int cap = dense_cap(nb, k);
        uint16_t *keys = w.get<uint16_t>(10, (size_t)(qchunk * nb));
        int32_t *cnt = w.get<int32_t>(12, (size_t)nq);
        int32_t *cand = w.get<int32_t>(13, (size_t)qchunk * cap);
        unsigned *krange = w.get<unsigned>(11, (size_t)(2 * qchunk));   // each query's least and largest key
        for (int64_t q0 = 0; q0 < nq; q0 += qchunk) {
            const int64_t qn_ = std::min(qchunk, nq - q0);
            launch_mfma_keys(qb + q0 * dp, qn_, xb, nb, dp, qn + q0, xn, keys, krange, krange + qchunk, st);
            launch_dense_select(keys, krange, krange + qchunk, qn_, nb, k, k_q ? k_q + q0 : nullptr, qf + q0, stat,
                                reinterpret_cast<int *>(stat + 1), dim, cap, cnt + q0, cand, st);
            // a query with more than cap candidates is ranked over every row
            launch_rank_candidates(queries + q0 * dim, qn_, dim, base, nb, cnt + q0, cand, nullptr, nullptr, cap, k,
                                   k_q ? k_q + q0 : nullptr, sqrt_out, out_ids + q0 * k, out_dist + q0 * k, st);
        }
        // the shared buffers are free again only once the stream has run
        // this far
        if (!ws) HANN_HIP(hipStreamSynchronize(st));
        return;
    }
    // filter thresholds: from the hint row's exact rank (k == 1), or from
    // the k-th smallest approximate distance over a row sample (every
    // stride-th row), which bounds the query's k-th exact rank
    float *thr = w.get<float>(11, (size_t)nq);
    int64_t ns = nb;
    if (hint && k == 1) {
        double *srank = w.get<double>(10, (size_t)nq);
        launch_hint_rank(queries, nq, dim, base, hint, srank, st);
        launch_thresholds(srank, 1, (int)nq, qf, stat, reinterpret_cast<int *>(stat + 1), dim, thr, st);
        ns = std::max<int64_t>(1, nb / 64);   // sizes the candidate room only
    } else {
        const int64_t stride = sample_stride(nb, k);
        ns = ceil_div(nb, stride);
        __bf16 *sb = w.get<__bf16>(8, (size_t)ns * dp);
        float *sn = w.get<float>(9, (size_t)ns);
        launch_gather_bf16(xb, dp, xn, stride, ns, sb, sn, st);
        const int64_t qchunk = std::max<int64_t>(1, std::min<int64_t>(nq, (int64_t)(256ll << 20) / (4 * ns)));
        float *dense = w.get<float>(10, (size_t)qchunk * ns);
        for (int64_t q0 = 0; q0 < nq; q0 += qchunk) {
            const int64_t qn_ = std::min(qchunk, nq - q0);
            launch_mfma_dense(qb + q0 * dp, qn_, sb, ns, dp, qn + q0, sn, dense, st);
            launch_sample_thresholds(dense, qn_, ns, nullptr, k, k_q ? k_q + q0 : nullptr, qf + q0, stat,
                                     reinterpret_cast<int *>(stat + 1), dim, thr + q0, st);
        }
    }
    // candidates: expected about k * nb / ns per query; room for 8x that
    const int64_t expect = std::max<int64_t>(k, ceil_div((int64_t)k * nb, ns));
    const int cap = (int)std::min<int64_t>(nb, std::max<int64_t>(256, 8 * expect));
    // a query whose candidates overflow the room: with a small base the
    // ranking kernel ranks every row itself (and the host does not wait for
    // the counts); otherwise it is redone below by the tiled exact path
    const bool rank_every = nb <= kRankEveryMax;
    // rows per workgroup: expected candidates per query and workgroup about
    // half its LDS slots (the rest, and the workgroup's shared overflow
    // list, absorb the spread)
    const int slots = mfma_filter_slots(dp);
    const int rt = (int)std::max<int64_t>(1, std::min<int64_t>(64, nb * (slots / 2) / (128 * expect)));
    const int64_t chunk = std::max<int64_t>(128, std::min<int64_t>(nq, (int64_t)(512ll << 20) / (4ll * cap)));
    int32_t *cnt = w.get<int32_t>(12, (size_t)nq);
    int32_t *cand = w.get<int32_t>(13, (size_t)chunk * cap);
    // with the candidates' approximate distances, the k-th smallest of them
    // gives a second, tighter threshold (by the same bound), and only the
    // candidates within it are ranked; not after a hint row's threshold,
    // which is already an exact distance's
    const bool refine = mfma_filter_writes_approx(dp) && !(hint && k == 1);
    float *cand_ap = mfma_filter_writes_approx(dp) ? w.get<float>(16, (size_t)chunk * cap) : nullptr;
    float *thr2 = refine ? w.get<float>(17, (size_t)nq) : nullptr;
    launch_fill_i32(cnt, nq, 0, st);   // a fill kernel: the search may be captured in a graph
    for (int64_t q0 = 0; q0 < nq; q0 += chunk) {
        const int64_t qn_ = std::min(chunk, nq - q0);
        launch_mfma_filter(qb + q0 * dp, qn_, xb, nb, dp, qn + q0, xn, thr + q0, rt, cnt + q0, cand, cand_ap, cap, st);
        if (refine)
            launch_sample_thresholds(cand_ap, qn_, cap, cnt + q0, k, k_q ? k_q + q0 : nullptr, qf + q0, stat,
                                     reinterpret_cast<int *>(stat + 1), dim, thr2 + q0, st);
        launch_rank_candidates(queries + q0 * dim, qn_, dim, base, rank_every ? nb : 0, cnt + q0, cand,
                               refine ? cand_ap : nullptr, refine ? thr2 + q0 : nullptr, cap, k,
                               k_q ? k_q + q0 : nullptr, sqrt_out,
                               out_ids + q0 * k, out_dist + q0 * k, st);
    }
    if (rank_every) {
        // the shared buffers are free again only once the stream has run
        // this far
        if (!ws) HANN_HIP(hipStreamSynchronize(st));
        return;
    }
    // overflowed queries take the exact path
    std::vector<int32_t> hc((size_t)nq);
    HANN_HIP(hipMemcpyAsync(hc.data(), cnt, (size_t)nq * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HANN_HIP(hipStreamSynchronize(st));
    std::vector<int64_t> redo;
    for (int64_t q = 0; q < nq; ++q)
        if (hc[q] > cap) redo.push_back(q);
    if (redo.empty()) return;
    const int64_t nr = (int64_t)redo.size();
    int64_t *dredo = w.get<int64_t>(7, (size_t)nr);
    HANN_HIP(hipMemcpyAsync(dredo, redo.data(), (size_t)nr * sizeof(int64_t), hipMemcpyHostToDevice, st));
    float *rq = w.get<float>(8, (size_t)nr * dim);
    launch_gather_rows(queries, dim, dredo, nr, rq, st);
    int64_t *rid = w.get<int64_t>(9, (size_t)nr * k);
    double *rdist = w.get<double>(10, (size_t)nr * k);
    launch_topk_exact(metric, rq, nr, base, nb, dim, k, rid, rdist, c.ws, st);
    for (int64_t i = 0; i < nr; ++i) {
        HANN_HIP(hipMemcpyAsync(out_ids + redo[i] * k, rid + i * k, (size_t)k * sizeof(int64_t),
                                hipMemcpyDeviceToDevice, st));
        HANN_HIP(hipMemcpyAsync(out_dist + redo[i] * k, rdist + i * k, (size_t)k * sizeof(double),
                                hipMemcpyDeviceToDevice, st));
    }
    HANN_HIP(hipStreamSynchronize(st));
}

__global__ void k_ids_to_assign(const int64_t *__restrict__ ids, int64_t n, int32_t *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (int32_t)ids[i];
}

void assign_bf16(Context &c, const float *data, int64_t n, const float *cents, int kc, int d, const int32_t *hint,
                 int32_t *out, hipStream_t st) {
    if (n == 0) return;
    int64_t *ids = c.ws_bf16.get<int64_t>(14, (size_t)n);
    double *dist = c.ws_bf16.get<double>(15, (size_t)n);
    certified_topk(c, data, n, cents, kc, d, 1, hint, 0, ids, dist, st);
    hipLaunchKernelGGL(k_ids_to_assign, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, ids, n, out);
    HANN_LAUNCH_CHECK();
}

void batch_topk(Context &c, int metric, const float *queries, int64_t nq, const float *base, int64_t nb, int dim,
                int k, int precision, int overfetch, int64_t *out_ids, double *out_dist, hipStream_t st) {
    (void)overfetch;
    if (precision == HANN_PREC_FP32 || nb == 0) {
        launch_topk_exact(metric, queries, nq, base, nb, dim, k, out_ids, out_dist, c.ws, st);
        return;
    }
    if (precision == HANN_PREC_BF16) {
        if (metric != HANN_METRIC_SQEUCLIDEAN && metric != HANN_METRIC_EUCLIDEAN)
            fail(HANN_ERR_UNSUPPORTED, "BF16 batch top-k supports the (squared) Euclidean metrics");
        if (nb >= (1ll << 31)) fail(HANN_ERR_UNSUPPORTED, "BF16 batch top-k needs fewer than 2^31 rows");
        certified_topk(c, queries, nq, base, nb, dim, k, nullptr, metric == HANN_METRIC_EUCLIDEAN ? 1 : 0, out_ids,
                       out_dist, st);
        return;
    }
    fail(HANN_ERR_UNSUPPORTED, strf("precision %d is not supported for batch top-k", precision));
}

}  // namespace hann
