// kmeans_api.cpp -- C-ABI of assignPoints and runKMeans.
#include <cstring>

#include "kmeans.h"
#include "batch.h"
#include "pqivf.h"
#include "rng.h"
#include "runtime.h"

using namespace hann;

extern "C" {

int hann_assign_points_dev(const float *data, size_t n, size_t d, const float *centroids, size_t k, int precision,
                           int32_t *out, void *stream) {
    HANN_API_BEGIN
    if (d == 0) fail(HANN_ERR_INVALID_ARG, "query must not be empty");
    if (k == 0) fail(HANN_ERR_INVALID_ARG, "no centroids");
    if (precision != HANN_PREC_FP32 && precision != HANN_PREC_BF16)
        fail(HANN_ERR_UNSUPPORTED, strf("precision %d is not supported for assignment", precision));
    if (n == 0) return HANN_OK;
    Context &c = global_ctx();
    hipStream_t st = stream ? (hipStream_t)stream : c.stream;
    assign_points(data, (int64_t)n, centroids, (int)k, (int)d, precision, nullptr, out, st);
    if (!stream) HANN_HIP(hipStreamSynchronize(st));
    HANN_API_END
}

int hann_assign_points(const float *data, size_t n, size_t d, const float *centroids, size_t k, int precision,
                       int32_t *out) {
    HANN_API_BEGIN
    if (d == 0) fail(HANN_ERR_INVALID_ARG, "query must not be empty");
    if (k == 0) fail(HANN_ERR_INVALID_ARG, "no centroids");
    if (precision != HANN_PREC_FP32 && precision != HANN_PREC_BF16)
        fail(HANN_ERR_UNSUPPORTED, strf("precision %d is not supported for assignment", precision));
    if (n == 0) return HANN_OK;
    Context &c = global_ctx();
    DevBuf<float> dd(n * d), dc(k * d);
    DevBuf<int32_t> da(n);
    h2d(dd.get(), data, n * d * sizeof(float), c.stream);
    h2d(dc.get(), centroids, k * d * sizeof(float), c.stream);
    assign_points(dd.get(), (int64_t)n, dc.get(), (int)k, (int)d, precision, nullptr, da.get(), c.stream);
    d2h(out, da.get(), n * sizeof(int32_t), c.stream);
    sync(c.stream);
    HANN_API_END
}

int hann_kmeans(const float *data, size_t n, size_t d, size_t k, int iters, uint64_t seed, int init, int precision,
                float *out_centroids, size_t *k_out) {
    HANN_API_BEGIN
    *k_out = 0;
    if (n == 0) fail(HANN_ERR_EMPTY, "no data for k-means training");
    if (d == 0) fail(HANN_ERR_INVALID_ARG, "query must not be empty");
    if (k == 0) fail(HANN_ERR_INVALID_ARG, "k must be positive");
    if (init != 0 && init != 1) fail(HANN_ERR_UNSUPPORTED, strf("k-means init %d is not supported", init));
    if (precision != HANN_PREC_FP32 && precision != HANN_PREC_BF16)
        fail(HANN_ERR_UNSUPPORTED, strf("precision %d is not supported for k-means", precision));
    Context &c = global_ctx();
    DevBuf<float> dd(n * d), dc(k * d);
    h2d(dd.get(), data, n * d * sizeof(float), c.stream);
    std::vector<uint64_t> st{stream_seed(seed, 0, 0)};
    KMeansScratch sc;
    const int64_t kk =
        run_kmeans(dd.get(), (int64_t)n, (int64_t)d, (int)d, (int)k, iters, 1, st, dc.get(), sc, c.stream, precision,
                   init);
    d2h(out_centroids, dc.get(), (size_t)kk * d * sizeof(float), c.stream);
    sync(c.stream);
    *k_out = (size_t)kk;
    HANN_API_END
}

}  // extern "C"
