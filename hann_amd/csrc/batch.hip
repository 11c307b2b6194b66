// batch.hip -- batch top-k dispatch.
#include "batch.h"

namespace hann {

void batch_topk(Context &c, int metric, const float *queries, int64_t nq, const float *base,
                int64_t nb, int dim, int k, int precision, int overfetch, int64_t *out_ids,
                double *out_dist, hipStream_t st) {
    (void)overfetch;
    if (precision == HANN_PREC_FP32) {
        launch_topk_exact(metric, queries, nq, base, nb, dim, k, out_ids, out_dist, c.ws, st);
        return;
    }
    fail(HANN_ERR_UNSUPPORTED, strf("precision %d is not supported for batch top-k", precision));
}

}  // namespace hann
