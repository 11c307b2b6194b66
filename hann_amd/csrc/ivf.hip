// ivf.hip -- inverted-list scan, ADC lookup tables and rerank (gfx950).
//
// The scan follows Search (pqivf/index.go:850-949): the entries of the
// first min(nprobe, K) lists, or of every list when they number fewer than k
// and the fallback is on; the PQ distance of each coded entry, the exact
// distance of each entry without codes; the k best by (distance, id).
#include "common.h"
#include "ivf.h"

namespace hann {
namespace dev {

// Exact-order squared distance between a query held in LDS and a row held
// in LDS (both plain float arrays).
struct LdsLoad {
    const float *p;
    __device__ __forceinline__ float operator()(int i) const { return p[i]; }
};

// One wave per query. Dynamic LDS layout (floats unless noted):
//   q[d] | c[d] | MODE_EXACT: xh[64][d+1]  /  MODE_ADC: lut[m][P] (f32 or f16)
//   | keys[cap] | ids[cap] (int64)
template <int MODE, typename CodeT>
__global__ void __launch_bounds__(64) k_scan(ScanArgs a) {
    extern __shared__ __align__(16) unsigned char s_raw[];
    const int lane = threadIdx.x;
    int64_t q = blockIdx.x;
    if (a.q_order) {
        const int b = a.order_base + (int)blockIdx.x;
        if (b >= *a.q_order_n) return;
        q = a.q_order[b];
    }
    const int d = a.d, m = a.m, s = a.s, P = a.pmax_cb;
    float *s_q = reinterpret_cast<float *>(s_raw);
    float *s_c = s_q + d;
    float *s_x = s_c + d;                       // exact: xhat rows; adc: lut
    size_t body = MODE == SCAN_MODE_EXACT ? (size_t)64 * (d + 1) * 4
                                          : (MODE == SCAN_MODE_ADC_F16 ? (size_t)m * P * 2 : (size_t)m * P * 4);
    body = (body + 15) & ~(size_t)15;
    float *s_keys = reinterpret_cast<float *>(reinterpret_cast<unsigned char *>(s_x) + body);
    int64_t *s_ids = reinterpret_cast<int64_t *>(s_keys + a.cap);

    int kq = a.k_q ? a.k_q[q] : a.k_default;
    kq = kq < 1 ? 1 : (kq > a.k_default ? a.k_default : kq);
    const int selk = a.rerank > 0 ? kq * a.rerank : kq;
    int npq = a.nprobe_q ? a.nprobe_q[q] : a.nprobe_default;
    if (npq > a.K) npq = a.K;
    if (npq > a.probe_cap) npq = a.probe_cap;
    if (npq < 0) npq = 0;
    const int64_t *probes = a.probes + q * a.probe_stride;

    for (int e = lane; e < d; e += kWave) s_q[e] = a.queries[q * d + e];

    // candidates in the probed lists
    int64_t total = 0;
    for (int j = lane; j < npq; j += kWave) {
        total += a.list_len[probes[j]];
    }
    for (int off = 32; off > 0; off >>= 1) total += __shfl_xor(total, off);
    const bool fb = total < kq && a.allow_fallback;
    if (lane == 0 && a.fell_back) a.fell_back[q] = fb ? 1 : 0;
    if (lane == 0 && fb && a.fb_counter) atomicAdd(a.fb_counter, 1ull);
    const int nl = fb ? a.K : npq;

    WaveTopK sel;
    sel.init(s_keys, s_ids, selk, a.cap);
    wave_sync();

    for (int li = 0; li < nl; ++li) {
        const int64_t l = fb ? li : probes[li];
        const int64_t e0 = a.list_start[l], e1 = e0 + a.list_len[l];
        if (e1 == e0) continue;
        const float *cent = a.coarse + l * d;
        for (int e = lane; e < d; e += kWave) s_c[e] = cent[e];
        wave_sync();
        if (MODE != SCAN_MODE_EXACT) {
            // residual query r = q - c (float32), then LUT[i][j] = |r_i - cb_ij|^2
            // in the reference's order for s-dim vectors
            for (int e = lane; e < d; e += kWave) s_c[e] = s_q[e] - s_c[e];
            wave_sync();
            const int tot = m * P;
            for (int t = lane; t < tot; t += kWave) {
                const int i = t / P, j = t % P;
                float v = __builtin_inff();
                if (j < (int)a.cb_len[i]) {
                    const float *cb = a.cb + ((int64_t)i * P + j) * s;
                    v = sqdiff_exact(LdsLoad{s_c + i * s}, PtrLoad{cb}, s);
                }
                if (MODE == SCAN_MODE_ADC_F16)
                    reinterpret_cast<__half *>(s_x)[t] = __float2half(v);
                else
                    s_x[t] = v;
            }
            wave_sync();
        }
        for (int64_t b = e0; b < e1; b += kWave) {
            const int64_t e = b + lane;
            const bool valid = e < e1;
            float key = 0.f;
            if (valid) {
                const bool coded = !a.any_uncoded || a.has_codes[e];
                if (coded) {
                    const CodeT *ce = reinterpret_cast<const CodeT *>(a.codes) + e * m;
                    if (MODE == SCAN_MODE_EXACT) {
                        // decodePQCode + vectorAdd (pqivf/index.go:678-714), then
                        // the reference-order distance to the query
                        float *xr = s_x + lane * (d + 1);
                        for (int i = 0; i < m; ++i) {
                            const float *cb = a.cb + ((int64_t)i * P + (int)ce[i]) * s;
                            for (int t = 0; t < s; ++t) xr[i * s + t] = s_c[i * s + t] + cb[t];
                        }
                        key = sqdiff_exact(LdsLoad{s_q}, LdsLoad{xr}, d);
                    } else if (MODE == SCAN_MODE_ADC_F16) {
                        const __half *lut = reinterpret_cast<const __half *>(s_x);
                        float acc = 0.f;
                        for (int i = 0; i < m; ++i) acc = acc + __half2float(lut[i * P + (int)ce[i]]);
                        key = acc;
                    } else {
                        float acc = 0.f;
                        for (int i = 0; i < m; ++i) acc = acc + s_x[i * P + (int)ce[i]];
                        key = acc;
                    }
                } else {
                    key = sqdiff_exact(LdsLoad{s_q}, PtrLoad{a.vecs + e * d}, d);
                }
            }
            const int64_t id = a.rerank > 0 ? e : (valid ? a.ids[e] : 0);
            sel.push(key, id, valid);
        }
        wave_sync();
    }
    const int got = sel.finish();
    float *ok = a.out_key + q * a.out_stride;
    int64_t *oi = a.out_id + q * a.out_stride;
    for (int t = lane; t < a.out_stride; t += kWave) {
        const bool okk = t < got && t < selk;
        ok[t] = okk ? sel.keys[t] : __builtin_inff();
        oi[t] = okk ? sel.ids[t] : (int64_t)-1;
    }
    if (lane == 0) a.out_cnt[q] = got < selk ? got : selk;
}

// Exact FP32 rerank: cand_pos[q][0..cnt) are entry positions; the exact
// reference-order distance to the stored vector orders them by
// (distance, id); the best k are written.
__global__ void __launch_bounds__(64) k_rerank(const float *__restrict__ queries, int d,
                                               const int64_t *__restrict__ cand_pos, const int32_t *__restrict__ cand_cnt,
                                               int cand_stride, const int32_t *__restrict__ k_q, int k_default,
                                               const float *__restrict__ vecs, const int64_t *__restrict__ ids,
                                               int cap, float *__restrict__ out_key, int64_t *__restrict__ out_id,
                                               int32_t *__restrict__ out_cnt, int out_stride, int kmax) {
    extern __shared__ __align__(16) unsigned char s_raw[];
    float *s_keys = reinterpret_cast<float *>(s_raw);
    int64_t *s_ids = reinterpret_cast<int64_t *>(s_keys + cap);
    float *s_q = reinterpret_cast<float *>(s_ids + cap);
    const int lane = threadIdx.x;
    const int64_t q = blockIdx.x;
    int kq = k_q ? k_q[q] : k_default;
    kq = kq < 1 ? 1 : (kq > kmax ? kmax : kq);
    for (int e = lane; e < d; e += kWave) s_q[e] = queries[q * d + e];
    WaveTopK sel;
    sel.init(s_keys, s_ids, kq, cap);
    wave_sync();
    const int n = cand_cnt[q];
    const int64_t *cp = cand_pos + q * cand_stride;
    for (int b = 0; b < n; b += kWave) {
        const int t = b + lane;
        const bool valid = t < n;
        float key = 0.f;
        int64_t id = 0;
        if (valid) {
            const int64_t e = cp[t];
            key = sqdiff_exact(LdsLoad{s_q}, PtrLoad{vecs + e * d}, d);
            id = ids[e];
        }
        sel.push(key, id, valid);
    }
    const int got = sel.finish();
    for (int t = lane; t < out_stride; t += kWave) {
        const bool okk = t < got && t < kq;
        out_key[q * out_stride + t] = okk ? sel.keys[t] : __builtin_inff();
        out_id[q * out_stride + t] = okk ? sel.ids[t] : (int64_t)-1;
    }
    if (lane == 0) out_cnt[q] = got < kq ? got : kq;
}

// true distance: sqrt of the float32 rank, in double on the device path
__global__ void k_finalize(const float *__restrict__ key, int64_t n, float *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = __builtin_sqrtf(key[i]);
}

}  // namespace dev

size_t scan_lds_bytes(int mode, int d, int m, int pmax_cb, int cap) {
    size_t body = mode == SCAN_MODE_EXACT ? (size_t)64 * (d + 1) * 4
                                          : (mode == SCAN_MODE_ADC_F16 ? (size_t)m * pmax_cb * 2
                                                                       : (size_t)m * pmax_cb * 4);
    body = (body + 15) & ~(size_t)15;
    return (size_t)2 * d * 4 + body + (size_t)cap * 12 + 16;
}

void launch_scan(int mode, int code_bytes, const ScanArgs &a, int64_t nq, hipStream_t st) {
    if (nq == 0) return;
    const size_t shm = scan_lds_bytes(mode, a.d, a.m, a.pmax_cb, a.cap);
    if (shm > 160 * 1024)
        fail(HANN_ERR_UNSUPPORTED, strf("scan needs %zu bytes of LDS (d=%d, m=%d, pq_k=%d, k=%d); the limit is 160 KiB",
                                        shm, a.d, a.m, a.pmax_cb, a.cap / 2));
#define HANN_SCAN_LAUNCH(MODE, CT)                                                                      \
    do {                                                                                                \
        HANN_HIP(hipFuncSetAttribute((const void *)dev::k_scan<MODE, CT>,                               \
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));            \
        for (int64_t q0 = 0; q0 < nq; q0 += 65535) {                                                    \
            ScanArgs b = a;                                                                             \
            int64_t qn = nq - q0 < 65535 ? nq - q0 : 65535;                                             \
            b.queries += q0 * a.d;                                                                      \
            b.probes += q0 * a.probe_stride;                                                            \
            if (b.nprobe_q) b.nprobe_q += q0;                                                           \
            if (b.k_q) b.k_q += q0;                                                                     \
            b.out_key += q0 * a.out_stride;                                                             \
            b.out_id += q0 * a.out_stride;                                                              \
            b.out_cnt += q0;                                                                            \
            if (b.fell_back) b.fell_back += q0;                                                         \
            hipLaunchKernelGGL((dev::k_scan<MODE, CT>), dim3((unsigned)qn), dim3(64), shm, st, b);       \
            HANN_LAUNCH_CHECK();                                                                        \
        }                                                                                               \
    } while (0)
    if (code_bytes == 1) {
        if (mode == SCAN_MODE_EXACT) HANN_SCAN_LAUNCH(SCAN_MODE_EXACT, uint8_t);
        else if (mode == SCAN_MODE_ADC_F32) HANN_SCAN_LAUNCH(SCAN_MODE_ADC_F32, uint8_t);
        else HANN_SCAN_LAUNCH(SCAN_MODE_ADC_F16, uint8_t);
    } else {
        if (mode == SCAN_MODE_EXACT) HANN_SCAN_LAUNCH(SCAN_MODE_EXACT, uint16_t);
        else if (mode == SCAN_MODE_ADC_F32) HANN_SCAN_LAUNCH(SCAN_MODE_ADC_F32, uint16_t);
        else HANN_SCAN_LAUNCH(SCAN_MODE_ADC_F16, uint16_t);
    }
#undef HANN_SCAN_LAUNCH
}

void launch_scan_ordered(int mode, int code_bytes, const ScanArgs &a, int64_t nq_max, const int32_t *order,
                         const int32_t *order_n, hipStream_t st) {
    if (nq_max == 0) return;
    const size_t shm = scan_lds_bytes(mode, a.d, a.m, a.pmax_cb, a.cap);
    if (shm > 160 * 1024)
        fail(HANN_ERR_UNSUPPORTED, strf("scan needs %zu bytes of LDS (d=%d, m=%d, pq_k=%d, k=%d); the limit is 160 KiB",
                                        shm, a.d, a.m, a.pmax_cb, a.cap / 2));
#define HANN_SCAN_ORDERED(MODE, CT)                                                                     \
    do {                                                                                                \
        HANN_HIP(hipFuncSetAttribute((const void *)dev::k_scan<MODE, CT>,                               \
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));            \
        for (int64_t b0 = 0; b0 < nq_max; b0 += 65535) {                                                \
            ScanArgs b = a;                                                                             \
            b.q_order = order;                                                                          \
            b.q_order_n = order_n;                                                                      \
            b.order_base = (int)b0;                                                                     \
            const int64_t bn = nq_max - b0 < 65535 ? nq_max - b0 : 65535;                               \
            hipLaunchKernelGGL((dev::k_scan<MODE, CT>), dim3((unsigned)bn), dim3(64), shm, st, b);       \
            HANN_LAUNCH_CHECK();                                                                        \
        }                                                                                               \
    } while (0)
    if (code_bytes == 1) {
        if (mode == SCAN_MODE_EXACT) HANN_SCAN_ORDERED(SCAN_MODE_EXACT, uint8_t);
        else if (mode == SCAN_MODE_ADC_F32) HANN_SCAN_ORDERED(SCAN_MODE_ADC_F32, uint8_t);
        else HANN_SCAN_ORDERED(SCAN_MODE_ADC_F16, uint8_t);
    } else {
        if (mode == SCAN_MODE_EXACT) HANN_SCAN_ORDERED(SCAN_MODE_EXACT, uint16_t);
        else if (mode == SCAN_MODE_ADC_F32) HANN_SCAN_ORDERED(SCAN_MODE_ADC_F32, uint16_t);
        else HANN_SCAN_ORDERED(SCAN_MODE_ADC_F16, uint16_t);
    }
#undef HANN_SCAN_ORDERED
}

void launch_rerank(const float *queries, int d, int64_t nq, const int64_t *cand_pos, const int32_t *cand_cnt,
                   int cand_stride, const int32_t *k_q, int k_default, int kmax, const float *vecs,
                   const int64_t *ids, float *out_key, int64_t *out_id, int32_t *out_cnt, int out_stride,
                   hipStream_t st) {
    if (nq == 0) return;
    const int cap = 2 * kp_for(kmax);
    const size_t shm = (size_t)cap * 12 + (size_t)d * 4 + 16;
    HANN_HIP(hipFuncSetAttribute((const void *)dev::k_rerank, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    for (int64_t q0 = 0; q0 < nq; q0 += 65535) {
        int64_t qn = nq - q0 < 65535 ? nq - q0 : 65535;
        hipLaunchKernelGGL(dev::k_rerank, dim3((unsigned)qn), dim3(64), shm, st, queries + q0 * d, d,
                           cand_pos + q0 * cand_stride, cand_cnt + q0, cand_stride, k_q ? k_q + q0 : nullptr,
                           k_default, vecs, ids, cap, out_key + q0 * out_stride, out_id + q0 * out_stride,
                           out_cnt + q0, out_stride, kmax);
        HANN_LAUNCH_CHECK();
    }
}

void launch_finalize(const float *key, int64_t n, float *out, hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL(dev::k_finalize, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, key, n, out);
    HANN_LAUNCH_CHECK();
}

}  // namespace hann
