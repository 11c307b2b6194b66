// batch.h -- batch top-k dispatch (exact FP32, or BF16 MFMA + FP32 rerank).
#pragma once

#include "kmeans.h"
#include "runtime.h"

namespace hann {

// out_dist is the true distance (FromRank applied) in double.
void batch_topk(Context &c, int metric, const float *queries, int64_t nq, const float *base,
                int64_t nb, int dim, int k, int precision, int overfetch, int64_t *out_ids,
                double *out_dist, hipStream_t st);

// A base matrix's BF16 copy, norms, max norm and exactness flag, kept across
// calls while the base does not change (the owner clears `ready` when it does).
struct BF16Base {
    DevBuf<__bf16> xb;
    DevBuf<float> xn, xf;
    DevBuf<unsigned> stat;   // [0] max norm bits, [1] inexact flag
    bool ready = false;
};

// Certified BF16 top-k over (squared) Euclidean ranks: identical to the
// exact path. hint (k == 1 only, may be null): a row per query whose exact
// rank bounds the query's nearest, e.g. the previous k-means assignment.
// k_q (may be null): each query's own k, clamped to [1, k]; its row (stride
// k) holds its k_q results, then (-1, inf). cache (may be null): the base's
// BF16 copy, reused while ready. ws (may be null): the pass's buffers, owned
// by the caller's stream, and then the call does not wait for the stream when
// it can avoid it; null: the context's shared buffers, and the call returns
// with its work complete.
void certified_topk(Context &c, const float *queries, int64_t nq, const float *base, int64_t nb, int dim, int k,
                    const int32_t *hint, int sqrt_out, int64_t *out_ids, double *out_dist, hipStream_t st,
                    const int32_t *k_q = nullptr, BF16Base *cache = nullptr, Workspace *ws = nullptr);

// an upper bound of the workspace bytes certified_topk takes for this shape
// (cached: the base's BF16 copy lives in a cache, not in the workspace),
// before an overflowed query's exact rerun
size_t certified_topk_scratch_bytes(int64_t nq, int64_t nb, int dim, int k, bool cached);

// bases of at most this many rows: certified_topk with a workspace runs
// without host synchronisation (an overflowed query is ranked over every row
// on the device)
int64_t certified_no_sync_max_rows();

// assignPoints on the BF16 path: out[i] = nearest centroid, lowest index on ties
void assign_bf16(Context &c, const float *data, int64_t n, const float *cents, int kc, int d, const int32_t *hint,
                 int32_t *out, hipStream_t st);

}  // namespace hann
