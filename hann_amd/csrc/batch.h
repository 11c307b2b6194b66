// batch.h -- batch top-k dispatch (exact FP32, or BF16 MFMA + FP32 rerank).
#pragma once

#include "runtime.h"

namespace hann {

// out_dist is the true distance (FromRank applied) in double.
void batch_topk(Context &c, int metric, const float *queries, int64_t nq, const float *base,
                int64_t nb, int dim, int k, int precision, int overfetch, int64_t *out_ids,
                double *out_dist, hipStream_t st);

}  // namespace hann
