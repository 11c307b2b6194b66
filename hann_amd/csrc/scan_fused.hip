// scan_fused.hip -- persistent ADC list scan over a query work queue (gfx950).
//
// The scan never builds a per-(query, list) lookup table. With y_e the PQ
// reconstruction of entry e of list l (code_e,i picks cb_i,code),
//   |q - c_l - y_e|^2 = |q - c_l|^2 + t_e + sum_i S_q[i][code_e,i]
//   t_e       = sum_i ( |cb_i,code|^2 + 2 <c_l,i , cb_i,code> )   (per entry, at sync)
//   S_q[i][j] = -2 <q_i , cb_ij>                                  (per query)
// so the only table is the query's S_q, built once per query (m*P entries,
// FP32 in LDS) and read by every list the query probes. |q - c_l|^2 is the
// coarse step's exact rank of list l (or is recomputed for lists reached by
// the fallback). S_q stays FP32: its entries are not small beside the
// distance, so rounding them to FP16 would cost accuracy.
//
// Workgroups are persistent: each pulls the next query from an atomic
// counter over a cost-ordered queue (largest nprobe first), so the ragged
// per-query work of mixed nprobe and k balances across the chip. The WPQ
// waves of a workgroup split the query's probed lists, keep their own top-k
// and merge at the end.
#include "common.h"
#include "ivf.h"
#include "scan_fused.h"

#include <type_traits>

namespace hann {
namespace dev {

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

// t_e for the entries of the given lists (grid: x over entries, y over lists)
template <typename CodeT>
__global__ void k_entry_terms(const int64_t *__restrict__ lists, int64_t j0, const int64_t *__restrict__ list_start,
                              const int64_t *__restrict__ list_len, const CodeT *__restrict__ codes,
                              const uint8_t *__restrict__ has, int d, int m, int P, const float *__restrict__ coarse,
                              const float *__restrict__ cb, float *__restrict__ term,
                              unsigned *__restrict__ ymax_tmax) {
    const int64_t j = j0 + blockIdx.y;
    const int64_t l = lists ? lists[j] : j;
    const int s = d / m;
    const float *c = coarse + l * d;
    const int64_t e0 = list_start[l], n = list_len[l];
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = e0 + r;
        float t = 0.f, yy = 0.f;
        if (has[e]) {
            for (int i = 0; i < m; ++i) {
                const float *w = cb + ((int64_t)i * P + (int)codes[e * m + i]) * s;
                float nn = 0.f, dc = 0.f;
                for (int u = 0; u < s; ++u) {
                    nn = __builtin_fmaf(w[u], w[u], nn);
                    dc = __builtin_fmaf(c[i * s + u], w[u], dc);
                }
                t = t + (nn + 2.f * dc);
                yy = yy + nn;
            }
        }
        term[e] = t;
        // upper bounds of |y_e| and |t_e| over the index (bits of
        // non-negative floats order like the floats)
        if (ymax_tmax) {
            atomicMax(&ymax_tmax[0], __float_as_uint(__builtin_sqrtf(yy) * 1.0001f));
            atomicMax(&ymax_tmax[1], __float_as_uint(__builtin_fabsf(t) * 1.0001f + 1e-30f));
        }
    }
}

// S[q][i][j] = -2 <q_i, cb_ij> (the same FMA chain as the in-kernel build);
// grid: (ceil(nq / 16), m), codebook i staged in LDS
__global__ void __launch_bounds__(256) k_query_tables(const float *__restrict__ queries, int64_t nq, int d, int m,
                                                      int P, const float *__restrict__ cb, float *__restrict__ S) {
    extern __shared__ __align__(16) unsigned char s_raw[];
    float *s_cb = reinterpret_cast<float *>(s_raw);
    const int i = blockIdx.y, s = d / m;
    const int64_t q0 = (int64_t)blockIdx.x * 16;
    for (int t = threadIdx.x; t < P * s; t += 256) s_cb[t] = cb[(int64_t)i * P * s + t];
    __syncthreads();
    for (int t = threadIdx.x; t < 16 * P; t += 256) {
        const int ql = t / P, j = t - ql * P;
        const int64_t q = q0 + ql;
        if (q >= nq) break;
        const float *qi = queries + q * d + i * s;
        float dc = 0.f;
        for (int u = 0; u < s; ++u) dc = __builtin_fmaf(qi[u], s_cb[j * s + u], dc);
        S[(q * m + i) * P + j] = -2.f * dc;
    }
}

// acc + sum_i S[i][code_i] in subquantizer order (uint8 codes)
template <int M>
__device__ __forceinline__ float split_sum(const uint8_t *codes, int64_t e, int m, int P, const float *S, float acc) {
    if constexpr (M > 0 && M % 8 == 0) {
        const uint2 *cw = reinterpret_cast<const uint2 *>(codes) + e * (M / 8);
        uint2 w[M / 8];
#pragma unroll
        for (int j = 0; j < M / 8; ++j) w[j] = cw[j];
#pragma unroll
        for (int j = 0; j < M / 8; ++j) {
#pragma unroll
            for (int b = 0; b < 4; ++b) acc = acc + S[(8 * j + b) * P + (int)((w[j].x >> (8 * b)) & 0xFFu)];
#pragma unroll
            for (int b = 0; b < 4; ++b) acc = acc + S[(8 * j + 4 + b) * P + (int)((w[j].y >> (8 * b)) & 0xFFu)];
        }
    } else if constexpr (M > 0 && M % 4 == 0) {
        const uint32_t *cw = reinterpret_cast<const uint32_t *>(codes) + e * (M / 4);
        uint32_t w[M / 4];
#pragma unroll
        for (int j = 0; j < M / 4; ++j) w[j] = cw[j];
#pragma unroll
        for (int j = 0; j < M / 4; ++j)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc = acc + S[(4 * j + b) * P + (int)((w[j] >> (8 * b)) & 0xFFu)];
    } else {
        const uint8_t *ce = codes + e * m;
        for (int i = 0; i < m; ++i) acc = acc + S[i * P + (int)ce[i]];
    }
    return acc;
}

template <int M>
__device__ __forceinline__ float split_sum16(const uint16_t *codes, int64_t e, int m, int P, const float *S,
                                             float acc) {
    const uint16_t *ce = codes + e * m;
    for (int i = 0; i < m; ++i) acc = acc + S[i * P + (int)ce[i]];
    return acc;
}

// 4 * byte B of w in one VALU op (SDWA operand select): the LDS byte offset
// of code B of a packed code word in a 256-entry FP32 table
template <int B>
__device__ __forceinline__ uint32_t byte_x4(uint32_t w) {
    uint32_t r;
    if constexpr (B == 0)
        asm("v_lshlrev_b32_sdwa %0, 2, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_0"
            : "=v"(r) : "v"(w));
    else if constexpr (B == 1)
        asm("v_lshlrev_b32_sdwa %0, 2, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_1"
            : "=v"(r) : "v"(w));
    else if constexpr (B == 2)
        asm("v_lshlrev_b32_sdwa %0, 2, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_2"
            : "=v"(r) : "v"(w));
    else
        asm("v_lshlrev_b32_sdwa %0, 2, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_3"
            : "=v"(r) : "v"(w));
    return r;
}

// b known once the caller's loop is unrolled
__device__ __forceinline__ uint32_t byte_x4_at(uint32_t w, int b) {
    switch (b) {
    case 0: return byte_x4<0>(w);
    case 1: return byte_x4<1>(w);
    case 2: return byte_x4<2>(w);
    default: return byte_x4<3>(w);
    }
}

typedef __attribute__((address_space(3))) const float lds_f32;

// the FP32 word at LDS byte address off (an LDS address, not a pointer into
// the dynamic area)
__device__ __forceinline__ float lds_at(uint32_t off) { return *(lds_f32 *)(uintptr_t)off; }

__device__ __forceinline__ int64_t uniform_i64(int64_t v) {
    const uint64_t u = (uint64_t)v;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(u >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// One query's top-k shared by the whole workgroup: LDS keys/ids[cap], the
// fill n and the threshold (tk, ti) are block-uniform. Rounds keep every
// wave in step, so compaction can use block barriers.
struct BlockTopK {
    float *keys;
    int64_t *ids;
    int cap, k, n;
    float tk;
    int64_t ti;

    __device__ void init(float *kb, int64_t *ib, int k_, int cap_) {
        keys = kb;
        ids = ib;
        k = k_;
        cap = cap_;
        n = 0;
        tk = __builtin_inff();
        ti = INT64_MAX;
    }

    // sort the filled slots (padded with +inf up to a power of two), keep the
    // best k; every thread of the block calls it
    __device__ void compact() {
        const int tid = threadIdx.x, nt = blockDim.x;
        int span = 2;
        while (span < n) span <<= 1;   // n <= cap, a power of two
        for (int t = n + tid; t < span; t += nt) {
            keys[t] = __builtin_inff();
            ids[t] = INT64_MAX;
        }
        __syncthreads();
        for (int size = 2; size <= span; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int t = tid; t < (span >> 1); t += nt) {
                    const int lo = 2 * t - (t & (stride - 1));
                    const int hi = lo + stride;
                    const bool up = (lo & size) == 0;
                    const float ka = keys[lo], kb = keys[hi];
                    const int64_t ia = ids[lo], ib = ids[hi];
                    const bool swap = up ? cand_less(kb, ib, ka, ia) : cand_less(ka, ia, kb, ib);
                    if (swap) {
                        keys[lo] = kb;
                        keys[hi] = ka;
                        ids[lo] = ib;
                        ids[hi] = ia;
                    }
                }
                __syncthreads();
            }
        }
        if (n > k) n = k;
        if (n == k) {
            // block-uniform: keep them in scalar registers
            tk = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(keys[k - 1])));
            ti = uniform_i64(ids[k - 1]);
        }
    }
};

// LDS: S_q[m*P] (FP32) | q[d] | keys[cap] | ids[cap] | per probe: first slot,
// length, list, chunks, |q-c|^2
// PC: the codebook stride P when known at compile time (256), else 0
template <int W, typename CodeT, int M, int PC>
__global__ void __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(4))) k_scan_block(ScanArgs a, FusedArgs f) {
    extern __shared__ __align__(16) unsigned char s_raw[];
    const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid / kWave), lane = tid % kWave;
    const int d = a.d, m = a.m, s = a.s, P = PC ? PC : a.pmax_cb, MP = m * P;
    const int cap = a.cap;
    // S_q first: its LDS address is then a compile-time constant, which the
    // gathers fold into their instruction offsets
    float *s_S = reinterpret_cast<float *>(s_raw);
    size_t off = ((size_t)MP * 4 + 15) & ~(size_t)15;
    float *s_q = reinterpret_cast<float *>(s_raw + off);
    off += ((size_t)d * 4 + 15) & ~(size_t)15;
    float *s_keys = reinterpret_cast<float *>(s_raw + off);
    off += ((size_t)cap * 4 + 15) & ~(size_t)15;
    int64_t *s_ids = reinterpret_cast<int64_t *>(s_raw + off);
    off += (size_t)cap * 8;
    int64_t *s_e0 = reinterpret_cast<int64_t *>(s_raw + off);   // per probe: first slot, length
    int32_t *s_len = reinterpret_cast<int32_t *>(s_e0 + f.lds_probes);
    int32_t *s_lst = s_len + f.lds_probes;
    int32_t *s_nch = s_lst + f.lds_probes;
    float *s_qc = reinterpret_cast<float *>(s_nch + f.lds_probes);
    // block bookkeeping last: the kernel has no static LDS, so S_q sits at
    // LDS address 0
    int64_t(*s_red)[W] = reinterpret_cast<int64_t(*)[W]>(
        (reinterpret_cast<uintptr_t>(s_qc + f.lds_probes) + 15) & ~(uintptr_t)15);
    int(*s_cnt)[W] = reinterpret_cast<int(*)[W]>(&s_red[2][0]);
    int &s_item = *reinterpret_cast<int *>(&s_cnt[2][0]);

    const int nq_all = f.nq_dev ? *f.nq_dev : f.nq;
    for (;;) {
        __syncthreads();
        if (tid == 0) s_item = atomicAdd(f.counter, 1);
        __syncthreads();
        const int item = s_item;
        if (item >= nq_all) break;
        const int64_t q = f.order[item];

        int kq = a.k_q ? a.k_q[q] : a.k_default;
        kq = kq < 1 ? 1 : (kq > a.k_default ? a.k_default : kq);
        const int selk = a.rerank > 0 ? kq * a.rerank : kq;
        int npq = a.nprobe_q ? a.nprobe_q[q] : a.nprobe_default;
        if (npq > a.K) npq = a.K;
        if (npq > a.probe_cap) npq = a.probe_cap;
        if (npq < 0) npq = 0;
        const int64_t *probes = a.probes + q * a.probe_stride;
        const bool in_lds = npq <= f.lds_probes;

        for (int e = tid; e < d; e += blockDim.x) s_q[e] = a.queries[q * d + e];
        __syncthreads();
        // S_q[i][j] = -2 <q_i, cb_ij>: from the batch's tables, or built here
        if (f.S) {
            const float *src = f.S + q * MP;
            if (MP % 4 == 0) {
                for (int t = tid; t < MP / 4; t += blockDim.x)
                    reinterpret_cast<float4 *>(s_S)[t] = reinterpret_cast<const float4 *>(src)[t];
            } else {
                for (int t = tid; t < MP; t += blockDim.x) s_S[t] = src[t];
            }
        } else {
            for (int t = tid; t < MP; t += blockDim.x) {
                const int i = t / P, j = t - i * P;
                const float *w = a.cb + ((int64_t)i * P + j) * s;
                float dc = 0.f;
                for (int u = 0; u < s; ++u) dc = __builtin_fmaf(s_q[i * s + u], w[u], dc);
                s_S[t] = -2.f * dc;
            }
        }
        // probed lists: entries and 64-entry chunks
        int64_t tot_e = 0, tot_c = 0;
        for (int li = tid; li < npq; li += blockDim.x) {
            const int64_t l = probes[li];
            const int64_t len = a.list_len[l];
            tot_e += len;
            tot_c += (len + kWave - 1) / kWave;
            if (in_lds) {
                s_e0[li] = a.list_start[l];
                s_len[li] = (int32_t)len;
                s_lst[li] = (int32_t)l;
                s_nch[li] = (int32_t)((len + kWave - 1) / kWave);
                s_qc[li] = f.probe_rank ? (float)f.probe_rank[q * a.probe_stride + li]
                                        : sqdiff_exact(PtrLoad{s_q}, PtrLoad{a.coarse + l * d}, d);
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            tot_e += __shfl_xor(tot_e, o);
            tot_c += __shfl_xor(tot_c, o);
        }
        if (lane == 0) {
            s_red[0][wave] = tot_e;
            s_red[1][wave] = tot_c;
        }
        __syncthreads();   // S_q, the probe table and the partial counts are complete
        tot_e = 0;
        tot_c = 0;
        for (int w = 0; w < W; ++w) {
            tot_e += s_red[0][w];
            tot_c += s_red[1][w];
        }
        const bool fb = tot_e < kq && a.allow_fallback;
        if (tid == 0 && fb && a.fb_counter) atomicAdd(a.fb_counter, 1ull);
        const int nl = fb ? a.K : npq;
        const bool tab = in_lds && !fb;
        if (fb) {
            // every list: count its chunks (rare; lists read from global memory)
            tot_c = 0;
            for (int l = tid; l < a.K; l += blockDim.x) tot_c += (a.list_len[l] + kWave - 1) / kWave;
            for (int o = 32; o > 0; o >>= 1) tot_c += __shfl_xor(tot_c, o);
            __syncthreads();
            if (lane == 0) s_red[1][wave] = tot_c;
            __syncthreads();
            tot_c = 0;
            for (int w = 0; w < W; ++w) tot_c += s_red[1][w];
        }

        BlockTopK sel;
        sel.init(s_keys, s_ids, selk, cap);
        // this wave's work item: chunk c of probe li; items go round-robin
        // over the waves, one per wave per round
        int li = -1, c = wave;
        int64_t l = 0, e0 = 0, len = 0, nch = 0;
        float qc = 0.f;
        auto enter = [&](int at) {
            li = at;
            if (li >= nl) return;
            if (tab) {
                l = s_lst[li];
                nch = s_nch[li];
                qc = s_qc[li];
                e0 = s_e0[li];
                len = s_len[li];
            } else {
                l = fb ? li : probes[li];
                nch = (a.list_len[l] + kWave - 1) / kWave;
                qc = (!fb && f.probe_rank) ? (float)f.probe_rank[q * a.probe_stride + li]
                                           : sqdiff_exact(PtrLoad{s_q}, PtrLoad{a.coarse + l * d}, d);
                e0 = a.list_start[l];
                len = a.list_len[l];
            }
        };
        enter(0);
        const int64_t rounds = (tot_c + W - 1) / W;
        auto walk = [&]() {
            while (li < nl && c >= nch) {
                c -= (int)nch;
                enter(li + 1);
            }
        };
        // append one round's survivors; the counts go through LDS so that
        // every thread knows the block's fill without atomics
        auto push_round = [&](float key, int64_t id, bool valid, int64_t r) {
            bool pass = valid && cand_less(key, id, sel.tk, sel.ti);
            uint64_t mask = __ballot(pass);
            const int par = (int)(r & 1);
            if (lane == 0) s_cnt[par][wave] = __popcll(mask);
            __syncthreads();
            int sum = 0, base = 0;
            for (int w = 0; w < W; ++w) {
                const int v = s_cnt[par][w];
                base += w < wave ? v : 0;
                sum += v;
            }
            if (sel.n + sum > sel.cap) {
                sel.compact();   // cap >= k + 64 W: the round now fits
                __syncthreads();
                pass = pass && cand_less(key, id, sel.tk, sel.ti);
                mask = __ballot(pass);
                if (lane == 0) s_cnt[par][wave] = __popcll(mask);
                __syncthreads();
                sum = 0;
                base = 0;
                for (int w = 0; w < W; ++w) {
                    const int v = s_cnt[par][w];
                    base += w < wave ? v : 0;
                    sum += v;
                }
                __syncthreads();   // the counts are read before the next round with this parity
            }
            if (pass) {
                const int pos = sel.n + base +
                                (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                               __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
                s_keys[pos] = key;
                s_ids[pos] = id;
            }
            sel.n += sum;
        };
        constexpr bool PIPE = sizeof(CodeT) == 1 && M > 0 && M % 8 == 0;
        // two chunks per wave per round (two independent gather chains, half
        // the barriers), software-pipelined: the next round's codes, terms and
        // ids are in flight while this round gathers from S_q. SLOT_IDS: the
        // ids kept are entry slots (for the rerank), not user ids.
        auto pipe_rounds = [&](auto slot_ids) {
            constexpr bool SLOT_IDS = decltype(slot_ids)::value;
            constexpr int NW8 = PIPE ? M / 8 : 1;
            const uint2 *cw = reinterpret_cast<const uint2 *>(a.codes);
            struct Slot {
                uint2 w[NW8];
                float t, qc;
                int64_t id;
                bool v;
            };
            // the next chunk of the probe table's lists (LDS reads only)
            auto walk_tab = [&]() {
                while (li < nl && c >= nch) {
                    c -= (int)nch;
                    ++li;
                    if (li < nl) {
                        nch = s_nch[li];
                        qc = s_qc[li];
                        e0 = s_e0[li];
                        len = s_len[li];
                    }
                }
            };
            // lanes past their list's end load slot 0 (it exists: some probed
            // list has entries) and are marked invalid; any code byte indexes
            // inside the block's LDS
            auto locate = [&](Slot &x) {
                walk_tab();
                const int64_t e = e0 + (int64_t)c * kWave + lane;
                x.v = li < nl && (int64_t)c * kWave + lane < len;
                x.qc = qc;
                const int64_t es = x.v ? e : 0;
#pragma unroll
                for (int j = 0; j < NW8; ++j) x.w[j] = cw[es * NW8 + j];
                x.t = f.term[es];
                if constexpr (SLOT_IDS)
                    x.id = e;
                else
                    x.id = a.ids[es];
                c += W;
            };
            // both chunks' lookups are issued before any is summed (two
            // independent chains of M LDS reads in flight); the sums keep
            // subquantizer order
            auto gather2 = [&](const Slot &x0, const Slot &x1, float &k0, float &k1) {
                constexpr int NS = NW8 * 8;
                constexpr int GS = NS <= 24 ? NS : 8;   // lookups per chain in flight
                float a0 = x0.qc + x0.t, a1 = x1.qc + x1.t;
#pragma unroll
                for (int g = 0; g < NS; g += GS) {
                    float v0[GS], v1[GS];
#pragma unroll
                    for (int u = 0; u < GS; ++u) {
                        const int i = g + u, sh = 8 * (i & 3);
                        const uint32_t w0 = (i & 4) ? x0.w[i >> 3].y : x0.w[i >> 3].x;
                        const uint32_t w1 = (i & 4) ? x1.w[i >> 3].y : x1.w[i >> 3].x;
                        if constexpr (PC == 256) {
                            // S_q at LDS address 0: table i at byte 1024 i
                            v0[u] = lds_at(byte_x4_at(w0, i & 3) + 1024u * i);
                            v1[u] = lds_at(byte_x4_at(w1, i & 3) + 1024u * i);
                        } else {
                            v0[u] = s_S[i * P + (int)((w0 >> sh) & 0xFFu)];
                            v1[u] = s_S[i * P + (int)((w1 >> sh) & 0xFFu)];
                        }
                    }
#pragma unroll
                    for (int u = 0; u < GS; ++u) {
                        a0 = a0 + v0[u];
                        a1 = a1 + v1[u];
                    }
                }
                k0 = sane_key(a0);
                k1 = sane_key(a1);
            };
            // one round: score chunks cu0 and cu1 and append their survivors
            auto round = [&](const Slot &cu0, const Slot &cu1, int64_t r) {
                float k0, k1;
                gather2(cu0, cu1, k0, k1);
                const int64_t i0 = cu0.v ? cu0.id : 0, i1 = cu1.v ? cu1.id : 0;
                // append both chunks' survivors (counts through LDS)
                bool p0 = cu0.v && cand_less(k0, i0, sel.tk, sel.ti);
                bool p1 = cu1.v && cand_less(k1, i1, sel.tk, sel.ti);
                uint64_t m0 = __ballot(p0), m1 = __ballot(p1);
                const int par = (int)(r & 1);
                if (lane == 0) s_cnt[par][wave] = __popcll(m0) + __popcll(m1);
                __syncthreads();
                int sum = 0, base = 0;
                for (int w = 0; w < W; ++w) {
                    const int v = s_cnt[par][w];
                    base += w < wave ? v : 0;
                    sum += v;
                }
                sum = __builtin_amdgcn_readfirstlane(sum);
                base = __builtin_amdgcn_readfirstlane(base);
                if (sel.n + sum > sel.cap) {
                    sel.compact();   // cap >= k + 128 W: the round now fits
                    __syncthreads();
                    p0 = p0 && cand_less(k0, i0, sel.tk, sel.ti);
                    p1 = p1 && cand_less(k1, i1, sel.tk, sel.ti);
                    m0 = __ballot(p0);
                    m1 = __ballot(p1);
                    if (lane == 0) s_cnt[par][wave] = __popcll(m0) + __popcll(m1);
                    __syncthreads();
                    sum = 0;
                    base = 0;
                    for (int w = 0; w < W; ++w) {
                        const int v = s_cnt[par][w];
                        base += w < wave ? v : 0;
                        sum += v;
                    }
                    sum = __builtin_amdgcn_readfirstlane(sum);
                    base = __builtin_amdgcn_readfirstlane(base);
                    __syncthreads();
                }
                const int below0 = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m0 >> 32),
                                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)m0, 0));
                const int below1 = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m1 >> 32),
                                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)m1, 0));
                if (p0) {
                    s_keys[sel.n + base + below0] = k0;
                    s_ids[sel.n + base + below0] = i0;
                }
                if (p1) {
                    const int pos = sel.n + base + __popcll(m0) + below1;
                    s_keys[pos] = k1;
                    s_ids[pos] = i1;
                }
                sel.n += sum;
            };
            // ping-pong slot sets: a round's loads land in the set that the
            // following round reads, with no register copies between them
            // (a copy would wait for the loads it copies)
            Slot a0, a1, b0, b1;
            locate(a0);
            locate(a1);
            const int64_t rounds2 = (tot_c + 2 * W - 1) / (2 * W);
            for (int64_t r = 0; r < rounds2; r += 2) {
                locate(b0);
                locate(b1);
                round(a0, a1, r);
                if (r + 1 >= rounds2) break;
                locate(a0);
                locate(a1);
                round(b0, b1, r + 1);
            }
        };
        // the pipelined rounds need the probe table in LDS: their list walk
        // then issues no global loads, and every round issues the same loads,
        // so each wait covers only the loads it needs
        // (the P = 256 gathers address S_q from LDS address 0)
        if (PIPE && !a.any_uncoded && tab && tot_c > 0 && (PC != 256 || __builtin_amdgcn_groupstaticsize() == 0)) {
            if (a.rerank > 0)
                pipe_rounds(std::true_type{});
            else
                pipe_rounds(std::false_type{});
        } else {
            for (int64_t r = 0; r < rounds; ++r) {
                walk();
                const int64_t e = e0 + (int64_t)c * kWave + lane;
                const bool valid = li < nl && (int64_t)c * kWave + lane < len;
                float key = 0.f;
                int64_t id = 0;
                if (valid) {
                    if (!a.any_uncoded || a.has_codes[e]) {
                        if constexpr (sizeof(CodeT) == 1)
                            key = split_sum<M>(reinterpret_cast<const uint8_t *>(a.codes), e, m, P, s_S,
                                               qc + f.term[e]);
                        else
                            key = split_sum16<M>(reinterpret_cast<const uint16_t *>(a.codes), e, m, P, s_S,
                                                 qc + f.term[e]);
                    } else {
                        key = sqdiff_exact(PtrLoad{s_q}, PtrLoad{a.vecs + e * d}, d);
                    }
                    key = sane_key(key);
                    id = a.rerank > 0 ? e : a.ids[e];
                }
                c += W;
                push_round(key, id, valid, r);
            }
        }
        __syncthreads();
        sel.compact();
        __syncthreads();
        const int got = sel.n;
        float *okey = a.out_key + q * a.out_stride;
        int64_t *oid = a.out_id + q * a.out_stride;
        // every slot of the output row is written: the result, then (inf, -1)
        for (int t = tid; t < a.out_stride; t += blockDim.x) {
            const bool ok = t < got && t < selk;
            okey[t] = ok ? s_keys[t] : __builtin_inff();
            oid[t] = ok ? s_ids[t] : (int64_t)-1;
        }
        if (tid == 0) a.out_cnt[q] = got < selk ? got : selk;
    }
}

}  // namespace dev

void launch_query_tables(const float *queries, int64_t nq, int d, int m, int P, const float *cb, float *S,
                         hipStream_t st) {
    if (nq == 0) return;
    const size_t shm = (size_t)P * (d / m) * 4;
    if (shm > 64 * 1024) fail(HANN_ERR_UNSUPPORTED, "query tables: codebook too large");
    hipLaunchKernelGGL(dev::k_query_tables, dim3((unsigned)ceil_div(nq, 16), (unsigned)m), dim3(256), shm, st, queries,
                       nq, d, m, P, cb, S);
    HANN_LAUNCH_CHECK();
}

// s_red[2][8] (int64), s_cnt[2][8], s_item, rounded up
constexpr size_t kBlockBookkeeping = 256;

size_t block_lds_bytes(int d, int m, int P, int cap, int lds_probes) {
    return (((size_t)d * 4 + 15) & ~(size_t)15) + (((size_t)m * P * 4 + 15) & ~(size_t)15) +
           (((size_t)cap * 4 + 15) & ~(size_t)15) + (size_t)cap * 8 + (((size_t)lds_probes * 24 + 15) & ~(size_t)15) +
           kBlockBookkeeping;
}

int block_cap(int kp, int w) {
    int c = 64;
    while (c < 2 * kp || c < kp + 128 * w) c <<= 1;
    return c;
}

void launch_entry_terms(const int64_t *lists, int64_t nlists, int64_t max_len, const int64_t *list_start,
                        const int64_t *list_len, const void *codes, int code_bytes, const uint8_t *has, int d, int m,
                        int P, const float *coarse, const float *cb, float *term, unsigned *ymax_tmax,
                        hipStream_t st) {
    if (nlists <= 0 || max_len <= 0) return;
    const unsigned gx = (unsigned)std::min<int64_t>(64, ceil_div(max_len, 256));
    for (int64_t j0 = 0; j0 < nlists; j0 += 65535) {
        const unsigned gy = (unsigned)std::min<int64_t>(65535, nlists - j0);
        if (code_bytes == 1)
            hipLaunchKernelGGL(dev::k_entry_terms<uint8_t>, dim3(gx, gy), dim3(256), 0, st, lists, j0, list_start,
                               list_len, reinterpret_cast<const uint8_t *>(codes), has, d, m, P, coarse, cb, term,
                               ymax_tmax);
        else
            hipLaunchKernelGGL(dev::k_entry_terms<uint16_t>, dim3(gx, gy), dim3(256), 0, st, lists, j0, list_start,
                               list_len, reinterpret_cast<const uint16_t *>(codes), has, d, m, P, coarse, cb, term,
                               ymax_tmax);
        HANN_LAUNCH_CHECK();
    }
}

void launch_scan_block(const ScanArgs &a_in, FusedArgs f, int64_t nq, int w, int kp, int code_bytes, int maxp,
                       int grid, hipStream_t st) {
    if (nq == 0) return;
    ScanArgs a = a_in;
    a.cap = block_cap(kp, w);
    f.nq = (int)nq;
    if (!f.nq_dev) {
        hipLaunchKernelGGL(dev::k_cost_order, dim3(1), dim3(1024), 0, st, a.nprobe_q, a.nprobe_default, (int)nq,
                           maxp, f.order, f.counter);
        HANN_LAUNCH_CHECK();
    }
    const size_t shm = block_lds_bytes(a.d, a.m, a.pmax_cb, a.cap, f.lds_probes);
    if (shm > 160 * 1024) fail(HANN_ERR_UNSUPPORTED, strf("list scan needs %zu bytes of LDS", shm));
#define HANN_BLOCK_P(W, CT, M, PP)                                                                        \
    do {                                                                                                  \
        HANN_HIP(hipFuncSetAttribute((const void *)dev::k_scan_block<W, CT, M, PP>,                       \
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));              \
        hipLaunchKernelGGL((dev::k_scan_block<W, CT, M, PP>), dim3((unsigned)grid), dim3(64 * W), shm, st, a, \
                           f);                                                                            \
        HANN_LAUNCH_CHECK();                                                                              \
    } while (0)
#define HANN_BLOCK(W, CT, M)                                        \
    do {                                                            \
        if (sizeof(CT) == 1 && M > 0 && a.pmax_cb == 256)           \
            HANN_BLOCK_P(W, CT, M, 256);                            \
        else                                                        \
            HANN_BLOCK_P(W, CT, M, 0);                              \
    } while (0)
#define HANN_BLOCK_W(CT, M)                         \
    do {                                            \
        if (w == 2) HANN_BLOCK(2, CT, M);           \
        else if (w == 4) HANN_BLOCK(4, CT, M);      \
        else HANN_BLOCK(8, CT, M);                  \
    } while (0)
    if (code_bytes == 1) {
        switch (a.m) {
        case 8: HANN_BLOCK_W(uint8_t, 8); break;
        case 16: HANN_BLOCK_W(uint8_t, 16); break;
        case 24: HANN_BLOCK_W(uint8_t, 24); break;
        case 32: HANN_BLOCK_W(uint8_t, 32); break;
        case 48: HANN_BLOCK_W(uint8_t, 48); break;
        case 64: HANN_BLOCK_W(uint8_t, 64); break;
        default: HANN_BLOCK_W(uint8_t, 0); break;
        }
    } else {
        HANN_BLOCK_W(uint16_t, 0);
    }
#undef HANN_BLOCK_W
#undef HANN_BLOCK
}

}  // namespace hann
