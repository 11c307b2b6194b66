// common.h -- shared host/device helpers of the hann_amd library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../../include/hann_amd.h"

namespace hann {

// ---------------------------------------------------------------- errors --
// Library calls catch hann::Error at the C-ABI and turn it into a status
// code plus the thread-local message returned by hann_last_error().
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &msg);

[[noreturn]] inline void fail(int code, const std::string &msg) { throw Error(code, msg); }

#define HANN_HIP(call)                                                                           \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            ::hann::fail(HANN_ERR_DEVICE, std::string(#call " failed: ") + hipGetErrorString(e_)); \
    } while (0)

#define HANN_LAUNCH_CHECK() HANN_HIP(hipGetLastError())

std::string strf(const char *fmt, ...);

// Make sure a gfx950 device is current (throws HANN_ERR_DEVICE otherwise).
void ensure_device();

// -------------------------------------------------------- device buffers --
template <typename T>
struct DevBuf {
    T *ptr = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { resize(count); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : ptr(o.ptr), n(o.n) { o.ptr = nullptr; o.n = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) { release(); ptr = o.ptr; n = o.n; o.ptr = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        n = 0;
    }
    // grow-only: keeps the allocation when it is big enough
    void resize(size_t count) {
        if (count <= n && ptr) return;
        release();
        if (count == 0) return;
        hipError_t e = hipMalloc(&ptr, count * sizeof(T));
        if (e != hipSuccess) {
            ptr = nullptr;
            (void)hipGetLastError();
            fail(HANN_ERR_OOM, strf("device allocation of %zu bytes failed: %s", count * sizeof(T),
                                    hipGetErrorString(e)));
        }
        n = count;
    }
    T *get() const { return ptr; }
};

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace hann

// ------------------------------------------------------------ device side --
#ifdef __HIPCC__

namespace hann {
namespace dev {

constexpr int kWave = 64;

// Exact-order squared difference over n floats: the AVX2+FMA instantiation
// of core/simd_kernels.inc.h:65-102 with the horizontal sum of
// core/simd_isa.h:31-41. Slot c of the 32-wide block holds accumulator c/8,
// lane c%8; a block of 32 elements updates every slot once, the 8-wide
// cleanup updates slots 0-7, the slots combine as
// (acc0+acc1)+(acc2+acc3) per lane, the lanes reduce as
// ((s0+s4)+(s1+s5))+((s2+s6)+(s3+s7)) and a fused scalar tail follows.
__device__ __forceinline__ float hsum8(const float s[8]) {
    float t0 = s[0] + s[4], t1 = s[1] + s[5], t2 = s[2] + s[6], t3 = s[3] + s[7];
    return (t0 + t1) + (t2 + t3);
}

template <typename LoadA, typename LoadB>
__device__ __forceinline__ float sqdiff_exact(LoadA la, LoadB lb, int n) {
    float acc[32];
#pragma unroll
    for (int c = 0; c < 32; ++c) acc[c] = 0.f;
    int i = 0;
    const int lim4 = n - (n % 32);
    for (; i < lim4; i += 32) {
#pragma unroll
        for (int c = 0; c < 32; ++c) {
            float d = la(i + c) - lb(i + c);
            acc[c] = __builtin_fmaf(d, d, acc[c]);
        }
    }
    const int lim = n - (n % 8);
    for (; i < lim; i += 8) {
#pragma unroll
        for (int l = 0; l < 8; ++l) {
            float d = la(i + l) - lb(i + l);
            acc[l] = __builtin_fmaf(d, d, acc[l]);
        }
    }
    float s[8];
#pragma unroll
    for (int l = 0; l < 8; ++l) s[l] = (acc[l] + acc[8 + l]) + (acc[16 + l] + acc[24 + l]);
    float sum = hsum8(s);
    for (; i < n; ++i) {
        float d = la(i) - lb(i);
        sum = __builtin_fmaf(d, d, sum);
    }
    return sum;
}

// Manhattan in the AVX tier's order (core/simd_kernels.inc.h:116-149 with
// HANN_UNROLL 1; the AVX2 tier aliases it, core/simd_distance.c:139-149).
template <typename LoadA, typename LoadB>
__device__ __forceinline__ float manhattan_exact(LoadA la, LoadB lb, int n) {
    float acc[8];
#pragma unroll
    for (int l = 0; l < 8; ++l) acc[l] = 0.f;
    int i = 0;
    const int lim = n - (n % 8);
    for (; i < lim; i += 8) {
#pragma unroll
        for (int l = 0; l < 8; ++l) acc[l] = acc[l] + fabsf(la(i + l) - lb(i + l));
    }
    float sum = hsum8(acc);
    for (; i < n; ++i) sum = sum + fabsf(la(i) - lb(i));
    return sum;
}

// Cosine distance in the AVX2 order (core/simd_kernels.inc.h:153-229).
template <typename LoadA, typename LoadB>
__device__ __forceinline__ float cosine_exact(LoadA la, LoadB lb, int n) {
    float dt[32], na[32], nb[32];
#pragma unroll
    for (int c = 0; c < 32; ++c) { dt[c] = 0.f; na[c] = 0.f; nb[c] = 0.f; }
    int i = 0;
    const int lim4 = n - (n % 32);
    for (; i < lim4; i += 32) {
#pragma unroll
        for (int c = 0; c < 32; ++c) {
            float x = la(i + c), y = lb(i + c);
            dt[c] = __builtin_fmaf(x, y, dt[c]);
            na[c] = __builtin_fmaf(x, x, na[c]);
            nb[c] = __builtin_fmaf(y, y, nb[c]);
        }
    }
    const int lim = n - (n % 8);
    for (; i < lim; i += 8) {
#pragma unroll
        for (int l = 0; l < 8; ++l) {
            float x = la(i + l), y = lb(i + l);
            dt[l] = __builtin_fmaf(x, y, dt[l]);
            na[l] = __builtin_fmaf(x, x, na[l]);
            nb[l] = __builtin_fmaf(y, y, nb[l]);
        }
    }
    float sd[8], sa[8], sb[8];
#pragma unroll
    for (int l = 0; l < 8; ++l) {
        sd[l] = (dt[l] + dt[8 + l]) + (dt[16 + l] + dt[24 + l]);
        sa[l] = (na[l] + na[8 + l]) + (na[16 + l] + na[24 + l]);
        sb[l] = (nb[l] + nb[8 + l]) + (nb[16 + l] + nb[24 + l]);
    }
    float d = hsum8(sd), ra = hsum8(sa), rb = hsum8(sb);
    for (; i < n; ++i) {
        float x = la(i), y = lb(i);
        d = __builtin_fmaf(x, y, d);
        ra = __builtin_fmaf(x, x, ra);
        rb = __builtin_fmaf(y, y, rb);
    }
    float normA = __builtin_sqrtf(ra), normB = __builtin_sqrtf(rb);
    if (normA == 0.0f || normB == 0.0f) return 1.0f;
    float denom = normA * normB;
    float sim = d / denom;
    if (sim > 1.0f) sim = 1.0f;
    if (sim < -1.0f) sim = -1.0f;
    return 1.0f - sim;
}

// Metric value in the reference's arithmetic. rank_only: Euclidean yields
// its rank (the squared distance, core/metric.go:155-158).
template <typename LoadA, typename LoadB>
__device__ __forceinline__ float metric_exact(int metric, LoadA la, LoadB lb, int n) {
    switch (metric) {
    case HANN_METRIC_SQEUCLIDEAN: return sqdiff_exact(la, lb, n);
    case HANN_METRIC_EUCLIDEAN: return __builtin_sqrtf(sqdiff_exact(la, lb, n));
    case HANN_METRIC_MANHATTAN: return manhattan_exact(la, lb, n);
    default: return cosine_exact(la, lb, n);
    }
}

struct PtrLoad {
    const float *p;
    __device__ __forceinline__ float operator()(int i) const { return p[i]; }
};

// ------------------------------------------------------ wave-scope sync --
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// ordering of candidates: (key, id) lexicographic; NaN keys sort last
__device__ __forceinline__ float sane_key(float k) { return k != k ? __builtin_inff() : k; }

__device__ __forceinline__ bool cand_less(float ka, int64_t ia, float kb, int64_t ib) {
    return ka < kb || (ka == kb && ia < ib);
}

// ------------------------------------------------------ wave top-k select --
// One wavefront keeps the best k (key, id) pairs of a candidate stream in an
// LDS buffer of CAP = 2*KP slots (KP = power of two >= max(k, 64)).
// Candidates strictly better than the current k-th best are appended with a
// ballot/mbcnt compaction; a full buffer is bitonic-sorted and cut to k.
// Every lane of the wave must call every method (uniform control flow).
struct WaveTopK {
    float *keys;
    int64_t *ids;
    int cap;   // 2*KP
    int k;
    int n;     // wave-uniform fill
    float tau_k;
    int64_t tau_i;

    __device__ void init(float *kbuf, int64_t *ibuf, int k_, int cap_) {
        keys = kbuf;
        ids = ibuf;
        k = k_;
        cap = cap_;
        n = 0;
        tau_k = __builtin_inff();
        tau_i = INT64_MAX;
    }

    // bitonic sort of the filled span: the first power of two >= n slots
    // (unused ones padded with +inf); slots past it are never read (every
    // reader stops at n)
    __device__ void sort_all() {
        const int lane = lane_id();
        int span = 2;
        while (span < n && span < cap) span <<= 1;
        for (int s = n + lane; s < span; s += kWave) {
            keys[s] = __builtin_inff();
            ids[s] = INT64_MAX;
        }
        wave_sync();
        for (int size = 2; size <= span; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int t = lane; t < (span >> 1); t += kWave) {
                    int lo = 2 * t - (t & (stride - 1));
                    int hi = lo + stride;
                    bool up = (lo & size) == 0;
                    float ka = keys[lo], kb = keys[hi];
                    int64_t ia = ids[lo], ib = ids[hi];
                    bool swap = up ? cand_less(kb, ib, ka, ia) : cand_less(ka, ia, kb, ib);
                    if (swap) {
                        keys[lo] = kb; keys[hi] = ka;
                        ids[lo] = ib; ids[hi] = ia;
                    }
                }
                wave_sync();
            }
        }
    }

    __device__ void compact() {
        sort_all();
        if (n > k) n = k;
        if (n == k) {
            tau_k = keys[k - 1];
            tau_i = ids[k - 1];
        }
        wave_sync();
    }

    __device__ void push(float key, int64_t id, bool valid) {
        key = sane_key(key);
        bool pass = valid && cand_less(key, id, tau_k, tau_i);
        uint64_t mask = __ballot(pass);
        if (mask == 0) return;
        int cnt = __popcll(mask);
        if (n + cnt > cap) compact();
        // re-test against the possibly tightened threshold
        pass = pass && cand_less(key, id, tau_k, tau_i);
        mask = __ballot(pass);
        cnt = __popcll(mask);
        if (pass) {
            int pos = n + __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                    __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
            keys[pos] = key;
            ids[pos] = id;
        }
        n += cnt;
        wave_sync();
    }

    // final sort; the first min(n, k) slots hold the result in order
    __device__ int finish() {
        compact();
        return n;
    }
};

}  // namespace dev
}  // namespace hann

#endif  // __HIPCC__
