// runtime.cpp -- device context, error plumbing and the C-ABI of the
// distance and top-k entry points.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <vector>

#include "batch.h"
#include "common.h"
#include "kernels.h"
#include "runtime.h"

namespace hann {

static thread_local std::string g_last_error;

void set_last_error(const std::string &msg) { g_last_error = msg; }

std::string strf(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    return std::string(buf);
}

static std::mutex g_init_mu;
static int g_device = -1;

void ensure_device() {
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (g_device >= 0) {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur == g_device) return;
        HANN_HIP(hipSetDevice(g_device));
        return;
    }
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
        (void)hipGetLastError();
        fail(HANN_ERR_DEVICE, "no HIP device is available (hann_amd runs only on a gfx950 GPU)");
    }
    HANN_HIP(hipSetDevice(0));
    hipDeviceProp_t prop;
    HANN_HIP(hipGetDeviceProperties(&prop, 0));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        fail(HANN_ERR_DEVICE, strf("device 0 is %s; hann_amd is built for gfx950 only", prop.gcnArchName));
    g_device = 0;
}

static void select_device(int device) {
    std::lock_guard<std::mutex> lk(g_init_mu);
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
        (void)hipGetLastError();
        fail(HANN_ERR_DEVICE, "no HIP device is available (hann_amd runs only on a gfx950 GPU)");
    }
    if (device < 0 || device >= count)
        fail(HANN_ERR_INVALID_ARG, strf("device %d out of range (%d devices)", device, count));
    HANN_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    HANN_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        fail(HANN_ERR_DEVICE, strf("device %d is %s; hann_amd is built for gfx950 only", device,
                                   prop.gcnArchName));
    g_device = device;
}

Context &global_ctx() {
    ensure_device();
    static Context *c = [] {
        auto *x = new Context();
        HANN_HIP(hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking));
        return x;
    }();
    return *c;
}

// Pinned host staging is not used for the synchronous host entry points:
// hipMemcpy from pageable memory keeps the caller's buffers borrowed for the
// duration of the call only.
void h2d(void *dst, const void *src, size_t bytes, hipStream_t st) {
    if (bytes == 0) return;
    HANN_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
}

void d2h(void *dst, const void *src, size_t bytes, hipStream_t st) {
    if (bytes == 0) return;
    HANN_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
}

void sync(hipStream_t st) { HANN_HIP(hipStreamSynchronize(st)); }

}  // namespace hann

using namespace hann;

extern "C" {

const char *hann_last_error(void) { return g_last_error.c_str(); }

const char *hann_version(void) { return "hann_amd 0.1 (gfx950)"; }

int hann_init(int device) {
    HANN_API_BEGIN
    select_device(device);
    (void)global_ctx();
    // the allocation-free top-k's scratch address on this device
    (void)noalloc_scratch_current();
    HANN_API_END
}

void hann_cpu_init(int support_level) {
    (void)support_level;
    if (hann_init(0) != HANN_OK) {
        std::fprintf(stderr, "hann_amd: hann_cpu_init failed: %s\n", hann_last_error());
    }
}

// ------------------------------------------------ reference distance symbols
static float pair_call(int metric, const float *a, const float *b, size_t n) {
    try {
        Context &c = global_ctx();
        std::lock_guard<std::mutex> lk(c.mu);
        float *d = c.ws.get<float>(2, 2 * n + 1);
        double *o = c.ws.get<double>(3, 1);
        h2d(d, a, n * sizeof(float), c.stream);
        h2d(d + n, b, n * sizeof(float), c.stream);
        launch_distance_batch(metric, d, 1, d + n, 1, (int)n, o, c.stream);
        double r = 0;
        d2h(&r, o, sizeof(double), c.stream);
        sync(c.stream);
        return (float)r;
    } catch (const hann::Error &e) {
        set_last_error(e.what());
        std::fprintf(stderr, "hann_amd: distance call failed: %s\n", e.what());
        return std::nanf("");
    }
}

static void batch_call(int metric, const float *q, const float *flat, size_t dim, size_t n, double *out) {
    try {
        if (n == 0) return;
        Context &c = global_ctx();
        std::lock_guard<std::mutex> lk(c.mu);
        float *dq = c.ws.get<float>(2, dim + n * dim);
        double *o = c.ws.get<double>(3, n);
        h2d(dq, q, dim * sizeof(float), c.stream);
        h2d(dq + dim, flat, n * dim * sizeof(float), c.stream);
        launch_distance_batch(metric, dq, 1, dq + dim, (int64_t)n, (int)dim, o, c.stream);
        d2h(out, o, n * sizeof(double), c.stream);
        sync(c.stream);
    } catch (const hann::Error &e) {
        set_last_error(e.what());
        std::fprintf(stderr, "hann_amd: batch distance call failed: %s\n", e.what());
        for (size_t i = 0; i < n; ++i) out[i] = std::nan("");
    }
}

float simd_euclidean(const float *a, const float *b, size_t n) { return pair_call(HANN_METRIC_EUCLIDEAN, a, b, n); }
float simd_squared_euclidean(const float *a, const float *b, size_t n) {
    return pair_call(HANN_METRIC_SQEUCLIDEAN, a, b, n);
}
float simd_manhattan(const float *a, const float *b, size_t n) { return pair_call(HANN_METRIC_MANHATTAN, a, b, n); }
float simd_cosine_distance(const float *a, const float *b, size_t n) { return pair_call(HANN_METRIC_COSINE, a, b, n); }
void simd_euclidean_batch(const float *q, const float *flat, size_t dim, size_t n, double *out) {
    batch_call(HANN_METRIC_EUCLIDEAN, q, flat, dim, n, out);
}
void simd_squared_euclidean_batch(const float *q, const float *flat, size_t dim, size_t n, double *out) {
    batch_call(HANN_METRIC_SQEUCLIDEAN, q, flat, dim, n, out);
}
void simd_manhattan_batch(const float *q, const float *flat, size_t dim, size_t n, double *out) {
    batch_call(HANN_METRIC_MANHATTAN, q, flat, dim, n, out);
}
void simd_cosine_distance_batch(const float *q, const float *flat, size_t dim, size_t n, double *out) {
    batch_call(HANN_METRIC_COSINE, q, flat, dim, n, out);
}

// simd_normalize (core/simd_ops.h:14): in place on a host vector; a failed
// call leaves the vector as it was and prints the reason (the reference's
// void signature has no error channel)
void simd_normalize(float *vec, size_t len) {
    try {
        if (len == 0) return;
        Context &c = global_ctx();
        std::lock_guard<std::mutex> lk(c.mu);
        float *d = c.ws.get<float>(2, len);
        h2d(d, vec, len * sizeof(float), c.stream);
        launch_normalize(d, 1, (int)len, c.stream);
        d2h(vec, d, len * sizeof(float), c.stream);
        sync(c.stream);
    } catch (const hann::Error &e) {
        set_last_error(e.what());
        std::fprintf(stderr, "hann_amd: normalize failed: %s\n", e.what());
    }
}

// ------------------------------------------------------------ batch API --
int hann_normalize_batch_dev(float *flat, size_t n, size_t dim, void *stream) {
    HANN_API_BEGIN
    if (n == 0 || dim == 0) return HANN_OK;
    if (dim > (size_t)INT32_MAX) fail(HANN_ERR_INVALID_ARG, "vector length above 2^31 - 1");
    Context &c = global_ctx();
    hipStream_t st = stream ? (hipStream_t)stream : c.stream;
    launch_normalize(flat, (int64_t)n, (int)dim, st);
    if (!stream) sync(st);
    HANN_API_END
}

int hann_normalize_batch(float *flat, size_t n, size_t dim) {
    HANN_API_BEGIN
    if (n == 0 || dim == 0) return HANN_OK;
    if (dim > (size_t)INT32_MAX) fail(HANN_ERR_INVALID_ARG, "vector length above 2^31 - 1");
    Context &c = global_ctx();
    std::lock_guard<std::mutex> lk(c.mu);
    float *d = c.ws.get<float>(2, n * dim);
    h2d(d, flat, n * dim * sizeof(float), c.stream);
    launch_normalize(d, (int64_t)n, (int)dim, c.stream);
    d2h(flat, d, n * dim * sizeof(float), c.stream);
    sync(c.stream);
    HANN_API_END
}

static void check_metric(int metric) {
    if (metric < 0 || metric > 3) fail(HANN_ERR_INVALID_ARG, strf("unknown metric %d", metric));
}

int hann_distance_batch_dev(int metric, int rank, const float *queries, size_t nq, const float *base,
                            size_t nb, size_t dim, double *out, void *stream) {
    HANN_API_BEGIN
    check_metric(metric);
    if (dim == 0) fail(HANN_ERR_INVALID_ARG, "query must not be empty");
    if (nb == 0 || nq == 0) return HANN_OK;
    Context &c = global_ctx();
    hipStream_t st = stream ? (hipStream_t)stream : c.stream;
    int m = (rank && metric == HANN_METRIC_EUCLIDEAN) ? HANN_METRIC_SQEUCLIDEAN : metric;
    launch_distance_batch(m, queries, (int64_t)nq, base, (int64_t)nb, (int)dim, out, st);
    if (!stream) sync(st);
    HANN_API_END
}

int hann_distance_batch(int metric, int rank, const float *queries, size_t nq, const float *base,
                        size_t nb, size_t dim, double *out) {
    HANN_API_BEGIN
    check_metric(metric);
    if (dim == 0) fail(HANN_ERR_INVALID_ARG, "query must not be empty");
    if (nb == 0 || nq == 0) return HANN_OK;
    Context &c = global_ctx();
    std::lock_guard<std::mutex> lk(c.mu);
    float *dq = c.ws.get<float>(2, nq * dim);
    float *db = c.ws.get<float>(4, nb * dim);
    double *o = c.ws.get<double>(3, nq * nb);
    h2d(dq, queries, nq * dim * sizeof(float), c.stream);
    h2d(db, base, nb * dim * sizeof(float), c.stream);
    int m = (rank && metric == HANN_METRIC_EUCLIDEAN) ? HANN_METRIC_SQEUCLIDEAN : metric;
    launch_distance_batch(m, dq, (int64_t)nq, db, (int64_t)nb, (int)dim, o, c.stream);
    d2h(out, o, nq * nb * sizeof(double), c.stream);
    sync(c.stream);
    HANN_API_END
}

int hann_batch_topk_dev(int metric, const float *queries, size_t nq, const float *base, size_t nb,
                        size_t dim, size_t k, int precision, size_t overfetch, int64_t *out_ids,
                        double *out_dist, void *stream) {
    HANN_API_BEGIN
    check_metric(metric);
    if (dim == 0) fail(HANN_ERR_INVALID_ARG, "query must not be empty");
    if (k == 0) fail(HANN_ERR_INVALID_ARG, "k must be positive");
    if (nq == 0) return HANN_OK;
    Context &c = global_ctx();
    hipStream_t st = stream ? (hipStream_t)stream : c.stream;
    std::unique_lock<std::mutex> lk(c.mu);
    batch_topk(c, metric, queries, (int64_t)nq, base, (int64_t)nb, (int)dim, (int)k, precision,
               (int)overfetch, out_ids, out_dist, st);
    if (!stream) sync(st);
    HANN_API_END
}

int hann_batch_topk_noalloc_dev(int metric, const float *queries, size_t nq, const float *base, size_t nb,
                                size_t dim, size_t k, int64_t *out_ids, double *out_dist, void *stream) {
    HANN_API_BEGIN
    check_metric(metric);
    if (dim == 0) fail(HANN_ERR_INVALID_ARG, "query must not be empty");
    if (k == 0) fail(HANN_ERR_INVALID_ARG, "k must be positive");
    if (nq == 0) return HANN_OK;
    ensure_device();
    // NULL: the device's null stream, and the call waits
    hipStream_t st = (hipStream_t)stream;
    launch_topk_exact_noalloc(metric, queries, (int64_t)nq, base, (int64_t)nb, (int)dim, (int)k, out_ids, out_dist,
                              st);
    if (!stream) sync(st);
    HANN_API_END
}

int hann_batch_topk(int metric, const float *queries, size_t nq, const float *base, size_t nb,
                    size_t dim, size_t k, int precision, size_t overfetch, int64_t *out_ids,
                    double *out_dist) {
    HANN_API_BEGIN
    check_metric(metric);
    if (dim == 0) fail(HANN_ERR_INVALID_ARG, "query must not be empty");
    if (k == 0) fail(HANN_ERR_INVALID_ARG, "k must be positive");
    if (nq == 0) return HANN_OK;
    Context &c = global_ctx();
    std::lock_guard<std::mutex> lk(c.mu);
    DevBuf<float> dq(nq * dim), db(nb * dim + 1);
    DevBuf<int64_t> oi(nq * k);
    DevBuf<double> od(nq * k);
    h2d(dq.get(), queries, nq * dim * sizeof(float), c.stream);
    h2d(db.get(), base, nb * dim * sizeof(float), c.stream);
    batch_topk(c, metric, dq.get(), (int64_t)nq, db.get(), (int64_t)nb, (int)dim, (int)k, precision,
               (int)overfetch, oi.get(), od.get(), c.stream);
    d2h(out_ids, oi.get(), nq * k * sizeof(int64_t), c.stream);
    d2h(out_dist, od.get(), nq * k * sizeof(double), c.stream);
    sync(c.stream);
    HANN_API_END
}

void hann_free_buffer(void *buf) { std::free(buf); }

}  // extern "C"
