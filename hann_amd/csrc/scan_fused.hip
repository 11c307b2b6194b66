// scan_fused.hip -- persistent ADC list scan over a query work queue (gfx950).
//
// The per-(query, list) lookup table uses the expansion
//   |q - c_l - cb_ij|^2 = |q - c_l|^2 + sum_i ( T_l[i][j] + S_q[i][j] )
//   T_l[i][j] = |cb_ij|^2 + 2 <c_l,i , cb_ij>      (per list, built at sync)
//   S_q[i][j] = -2 <q_i , cb_ij>                   (per query, built once)
// so a list's table costs m*P additions instead of m*P*s multiply-adds.
// |q - c_l|^2 is the coarse step's exact rank of list l (or is recomputed
// for lists reached by the fallback).
//
// Workgroups are persistent: each pulls the next query from an atomic
// counter over a cost-ordered queue (largest nprobe first), so the ragged
// per-query work of mixed nprobe and k balances across the chip. The WPQ
// waves of a workgroup split the query's probed lists, keep their own top-k
// and merge at the end.
#include "common.h"
#include "ivf.h"
#include "scan_fused.h"

namespace hann {
namespace dev {

// T[l][i][j] for every list (grid: K x m)
__global__ void k_list_tables(const float *__restrict__ coarse, int d, int m, int s, int P,
                              const float *__restrict__ cb, const int64_t *__restrict__ cb_len,
                              float *__restrict__ T) {
    const int l = blockIdx.x, i = blockIdx.y;
    const float *c = coarse + (int64_t)l * d + i * s;
    for (int j = threadIdx.x; j < P; j += blockDim.x) {
        float v = __builtin_inff();
        if (j < (int)cb_len[i]) {
            const float *w = cb + ((int64_t)i * P + j) * s;
            float nn = 0.f, dc = 0.f;
            for (int t = 0; t < s; ++t) {
                nn = __builtin_fmaf(w[t], w[t], nn);
                dc = __builtin_fmaf(c[t], w[t], dc);
            }
            v = nn + 2.f * dc;
        }
        T[((int64_t)l * m + i) * P + j] = v;
    }
}

// queue order: queries by descending nprobe (counting sort, one workgroup)
__global__ void __launch_bounds__(1024) k_cost_order(const int32_t *__restrict__ nprobe_q, int nprobe_default,
                                                     int nq, int maxp, int32_t *__restrict__ order,
                                                     int32_t *__restrict__ counter) {
    __shared__ int32_t s_cnt[1025];
    const int t = threadIdx.x;
    const int B = maxp < 1024 ? maxp + 1 : 1025;
    for (int b = t; b < B; b += 1024) s_cnt[b] = 0;
    __syncthreads();
    for (int q = t; q < nq; q += 1024) {
        int p = nprobe_q ? nprobe_q[q] : nprobe_default;
        p = p < 0 ? 0 : (p >= B ? B - 1 : p);
        atomicAdd(&s_cnt[p], 1);
    }
    __syncthreads();
    if (t == 0) {
        // exclusive prefix from the largest bucket down
        int run = 0;
        for (int b = B - 1; b >= 0; --b) {
            int c = s_cnt[b];
            s_cnt[b] = run;
            run += c;
        }
        *counter = 0;
    }
    __syncthreads();
    for (int q = t; q < nq; q += 1024) {
        int p = nprobe_q ? nprobe_q[q] : nprobe_default;
        p = p < 0 ? 0 : (p >= B ? B - 1 : p);
        order[atomicAdd(&s_cnt[p], 1)] = q;
    }
}

template <typename LutT>
__device__ __forceinline__ float lut_get(const LutT *lut, int idx);
template <>
__device__ __forceinline__ float lut_get<float>(const float *lut, int idx) { return lut[idx]; }
template <>
__device__ __forceinline__ float lut_get<__half>(const __half *lut, int idx) { return __half2float(lut[idx]); }

template <typename LutT>
__device__ __forceinline__ LutT lut_make(float v);
template <>
__device__ __forceinline__ float lut_make<float>(float v) { return v; }
template <>
__device__ __forceinline__ __half lut_make<__half>(float v) { return __float2half(v); }

// Sum of an entry's table values in subquantizer order, starting at qc.
// M > 0 (uint8 codes, M % 4 == 0): the codes are read as M/4 dwords.
template <int M, typename CodeT, typename LutT>
__device__ __forceinline__ float adc_sum(const void *codes, int64_t e, int m, int P, const LutT *lut, float qc) {
    float acc = qc;
    if constexpr (M > 0 && sizeof(CodeT) == 1 && M % 4 == 0) {
        const uint32_t *cw = reinterpret_cast<const uint32_t *>(codes) + e * (M / 4);
        uint32_t w[M / 4];
#pragma unroll
        for (int j = 0; j < M / 4; ++j) w[j] = cw[j];
#pragma unroll
        for (int j = 0; j < M / 4; ++j)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc = acc + lut_get<LutT>(lut, (4 * j + b) * P + (int)((w[j] >> (8 * b)) & 0xFFu));
    } else {
        const CodeT *ce = reinterpret_cast<const CodeT *>(codes) + e * m;
        for (int i = 0; i < m; ++i) acc = acc + lut_get<LutT>(lut, i * P + (int)ce[i]);
    }
    return acc;
}

// LDS: q[d] | S_q[m*P] (LutT) | per wave: lut[m*P] (LutT), keys[cap], ids[cap]
template <int WPQ, typename CodeT, typename LutT, int M>
__global__ void __launch_bounds__(64 * WPQ) k_scan_fused(ScanArgs a, FusedArgs f) {
    extern __shared__ __align__(16) unsigned char s_raw[];
    __shared__ int s_item;
    const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
    const int d = a.d, m = a.m, s = a.s, P = a.pmax_cb, MP = m * P;
    float *s_q = reinterpret_cast<float *>(s_raw);
    size_t off = ((size_t)d * 4 + 15) & ~(size_t)15;
    LutT *s_S = reinterpret_cast<LutT *>(s_raw + off);
    off += ((size_t)MP * sizeof(LutT) + 15) & ~(size_t)15;
    const size_t per_wave = (((size_t)MP * sizeof(LutT) + 15) & ~(size_t)15) + (size_t)a.cap * 12;
    unsigned char *wbase = s_raw + off + wave * per_wave;
    LutT *s_lut = reinterpret_cast<LutT *>(wbase);
    float *s_keys = reinterpret_cast<float *>(wbase + (((size_t)MP * sizeof(LutT) + 15) & ~(size_t)15));
    int64_t *s_ids = reinterpret_cast<int64_t *>(s_keys + a.cap);

    for (;;) {
        __syncthreads();
        if (tid == 0) s_item = atomicAdd(f.counter, 1);
        __syncthreads();
        const int item = s_item;
        if (item >= f.nq) break;
        const int64_t q = f.order[item];

        int kq = a.k_q ? a.k_q[q] : a.k_default;
        kq = kq < 1 ? 1 : (kq > a.k_default ? a.k_default : kq);
        const int selk = a.rerank > 0 ? kq * a.rerank : kq;
        int npq = a.nprobe_q ? a.nprobe_q[q] : a.nprobe_default;
        if (npq > a.K) npq = a.K;
        if (npq > a.probe_cap) npq = a.probe_cap;
        if (npq < 0) npq = 0;
        const int64_t *probes = a.probes + q * a.probe_stride;

        for (int e = tid; e < d; e += blockDim.x) s_q[e] = a.queries[q * d + e];
        __syncthreads();
        // S_q[i][j] = -2 <q_i, cb_ij>
        for (int t = tid; t < MP; t += blockDim.x) {
            const int i = t / P, j = t % P;
            const float *w = a.cb + ((int64_t)i * P + j) * s;
            float dc = 0.f;
            for (int u = 0; u < s; ++u) dc = __builtin_fmaf(s_q[i * s + u], w[u], dc);
            s_S[t] = lut_make<LutT>(-2.f * dc);
        }
        int64_t total = 0;
        for (int j = tid; j < npq; j += blockDim.x) total += a.list_len[probes[j]];
        for (int o = 32; o > 0; o >>= 1) total += __shfl_xor(total, o);
        // sum the per-wave partial counts through LDS
        __shared__ int64_t s_tot[16];
        if (lane == 0) s_tot[wave] = total;
        __syncthreads();
        total = 0;
        for (int w = 0; w < WPQ; ++w) total += s_tot[w];
        const bool fb = total < kq && a.allow_fallback;
        if (tid == 0 && fb && a.fb_counter) atomicAdd(a.fb_counter, 1ull);
        const int nl = fb ? a.K : npq;

        WaveTopK sel;
        sel.init(s_keys, s_ids, selk, a.cap);
        wave_sync();
        for (int li = wave; li < nl; li += WPQ) {
            const int64_t l = fb ? li : probes[li];
            const int64_t e0 = a.list_start[l], e1 = e0 + a.list_len[l];
            if (e1 == e0) continue;
            // |q - c_l|^2: the coarse rank when the list was probed
            float qc;
            if (!fb && f.probe_rank) {
                qc = (float)f.probe_rank[q * a.probe_stride + li];
            } else {
                qc = sqdiff_exact(PtrLoad{s_q}, PtrLoad{a.coarse + l * d}, d);
            }
            const float *Tl = f.T + l * (int64_t)MP;
            for (int t = lane; t < MP; t += kWave) s_lut[t] = lut_make<LutT>(Tl[t] + lut_get<LutT>(s_S, t));
            wave_sync();
            for (int64_t b = e0; b < e1; b += kWave) {
                const int64_t e = b + lane;
                const bool valid = e < e1;
                float key = 0.f;
                if (valid) {
                    if (!a.any_uncoded || a.has_codes[e]) {
                        key = adc_sum<M, CodeT, LutT>(a.codes, e, m, P, s_lut, qc);
                    } else {
                        key = sqdiff_exact(PtrLoad{s_q}, PtrLoad{a.vecs + e * d}, d);
                    }
                }
                const int64_t id = a.rerank > 0 ? e : (valid ? a.ids[e] : 0);
                sel.push(key, id, valid);
            }
            wave_sync();
        }
        int got = sel.finish();
        // merge the waves' lists into wave 0's
        __shared__ int s_got[16];
        if (lane == 0) s_got[wave] = got;
        __syncthreads();
        if (wave == 0) {
            for (int w = 1; w < WPQ; ++w) {
                const float *ok = reinterpret_cast<const float *>(
                    s_raw + off + w * per_wave + (((size_t)MP * sizeof(LutT) + 15) & ~(size_t)15));
                const int64_t *oi = reinterpret_cast<const int64_t *>(ok + a.cap);
                const int n = s_got[w];
                for (int b = 0; b < n; b += kWave) {
                    const int t = b + lane;
                    sel.push(t < n ? ok[t] : 0.f, t < n ? oi[t] : 0, t < n);
                }
            }
            got = sel.finish();
            float *okey = a.out_key + q * a.out_stride;
            int64_t *oid = a.out_id + q * a.out_stride;
            // every slot of the output row is written: the result, then (inf, -1)
            for (int t = lane; t < a.out_stride; t += kWave) {
                const bool ok = t < got && t < selk;
                okey[t] = ok ? sel.keys[t] : __builtin_inff();
                oid[t] = ok ? sel.ids[t] : (int64_t)-1;
            }
            if (lane == 0) a.out_cnt[q] = got < selk ? got : selk;
        }
    }
}

}  // namespace dev

void launch_list_tables(const float *coarse, int K, int d, int m, int P, const float *cb, const int64_t *cb_len,
                        float *T, hipStream_t st) {
    if (K == 0) return;
    hipLaunchKernelGGL(dev::k_list_tables, dim3((unsigned)K, (unsigned)m), dim3(256), 0, st, coarse, d, m, d / m, P,
                       cb, cb_len, T);
    HANN_LAUNCH_CHECK();
}

size_t fused_lds_bytes(int wpq, int lut_bytes, int d, int m, int P, int cap) {
    const size_t mp = (((size_t)m * P * lut_bytes) + 15) & ~(size_t)15;
    return (((size_t)d * 4 + 15) & ~(size_t)15) + mp + (size_t)wpq * (mp + (size_t)cap * 12) + 16;
}

void launch_scan_fused(const ScanArgs &a, FusedArgs f, int64_t nq, int wpq, int lut16, int code_bytes, int maxp,
                       int grid, hipStream_t st) {
    if (nq == 0) return;
    f.nq = (int)nq;
    hipLaunchKernelGGL(dev::k_cost_order, dim3(1), dim3(1024), 0, st, a.nprobe_q, a.nprobe_default, (int)nq, maxp,
                       f.order, f.counter);
    HANN_LAUNCH_CHECK();
    const size_t shm = fused_lds_bytes(wpq, lut16 ? 2 : 4, a.d, a.m, a.pmax_cb, a.cap);
    if (shm > 160 * 1024) fail(HANN_ERR_UNSUPPORTED, strf("fused scan needs %zu bytes of LDS", shm));
#define HANN_FUSED(W, CT, LT, M)                                                                          \
    do {                                                                                                  \
        HANN_HIP(hipFuncSetAttribute((const void *)dev::k_scan_fused<W, CT, LT, M>,                       \
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));              \
        hipLaunchKernelGGL((dev::k_scan_fused<W, CT, LT, M>), dim3((unsigned)grid), dim3(64 * W), shm, st, a, f); \
        HANN_LAUNCH_CHECK();                                                                              \
    } while (0)
#define HANN_FUSED_W(CT, LT, M)                         \
    do {                                                \
        if (wpq == 1) HANN_FUSED(1, CT, LT, M);         \
        else if (wpq == 2) HANN_FUSED(2, CT, LT, M);    \
        else HANN_FUSED(4, CT, LT, M);                  \
    } while (0)
#define HANN_FUSED_M(LT)                                                    \
    do {                                                                    \
        switch (a.m) {                                                      \
        case 8: HANN_FUSED_W(uint8_t, LT, 8); break;                        \
        case 16: HANN_FUSED_W(uint8_t, LT, 16); break;                      \
        case 24: HANN_FUSED_W(uint8_t, LT, 24); break;                      \
        case 32: HANN_FUSED_W(uint8_t, LT, 32); break;                      \
        case 48: HANN_FUSED_W(uint8_t, LT, 48); break;                      \
        case 64: HANN_FUSED_W(uint8_t, LT, 64); break;                      \
        default: HANN_FUSED_W(uint8_t, LT, 0); break;                       \
        }                                                                   \
    } while (0)
    if (code_bytes == 1) {
        if (lut16) HANN_FUSED_M(__half);
        else HANN_FUSED_M(float);
    } else {
        if (lut16) HANN_FUSED_W(uint16_t, __half, 0);
        else HANN_FUSED_W(uint16_t, float, 0);
    }
#undef HANN_FUSED_M
#undef HANN_FUSED_W
#undef HANN_FUSED
}

}  // namespace hann
