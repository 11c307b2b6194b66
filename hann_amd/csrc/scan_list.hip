// scan_list.hip -- list-major scan: decode each probed list once per batch
// and score it against every query that probes it on the matrix cores.
//
// With the entry's reconstruction y_e = concat_i cb_i[code_e,i] (residual
// part) and the per-entry list term t_e,
//   |q - c_l - y_e|^2 = |q - c_l|^2 + t_e - 2 <q, y_e>
// and for a list l probed by queries Q_l the inner products are one product
// of the decoded list [n_l x d] with Q_l^T [d x |Q_l|]. The pairs (query,
// probed list) of a batch are grouped by list; a work item is one chunk of up
// to 2,048 entries of a list with up to 16 of its queries, and belongs to one
// wave. The wave gathers 32 entries at a time from a BF16 codebook in LDS
// (BF16 hi + lo for the split mode) straight into the A operands of two
// v_mfma_f32_16x16x32_bf16 tiles and multiplies them with the item's queries,
// which stay in registers as the B operand (scaled by -2 and the query's
// key scale); the accumulator starts at the entry's and the pair's terms, so
// it ends at the scaled distance. Each list's codes are read once per item
// instead of once per (query, list).
//
// A key is the FP16 value of the distance times a power of two chosen per
// query (so that no probed entry's distance reaches FP16's range), with its
// bits mapped so that unsigned order is value order. A query's rerank
// candidates are its selk = k * R best probed entries by (key, entry
// position). They are found without writing every key:
//  1. the sample pass scores each query's nearest probed lists (at least
//     two, and until they hold 4 selk entries) and writes their keys into
//     the query's sample row;
//  2. k_theta takes, per query, theta = the selk-th smallest sample key --
//     at least selk entries have keys <= theta, and the selk best do;
//  3. the filter pass scores every entry again, with bit-identical keys, and
//     writes only the (key, entry position) records with key <= theta into
//     the query's record room (about 1.5 selk of them);
//  4. k_select_records picks the selk best records per query.
// The selection is the same as over every key; those are reranked in exact
// FP32. A query whose records overflow its room (8 selk, at least 256) is
// ranked by the table scan instead.
#include "common.h"
#include "mfma.h"
#include "scan_list.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace hann {
namespace dev {

constexpr int LQ = 16;          // queries per work item (the MFMA's N)
constexpr int LCH_SHIFT = 11;
constexpr int LCH = 1 << LCH_SHIFT;   // entries per work item (a chunk of the list; a multiple of 32)
constexpr int SEL_MAXP = 2048;  // probes per query the threshold and select kernels take

__global__ void k_fill_i32(int32_t *__restrict__ p, int64_t n, int32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__device__ __forceinline__ int clamp_nprobe(const ScanArgs &a, int64_t q) {
    int npq = a.nprobe_q ? a.nprobe_q[q] : a.nprobe_default;
    if (npq > a.K) npq = a.K;
    if (npq > a.probe_cap) npq = a.probe_cap;
    return npq < 0 ? 0 : npq;
}

// 16-bit keys of two scaled distances, packed: each is clamped on its bits
// as a signed integer to [0, +inf] (negative values, -0 and negative NaNs
// become +0, positive NaNs +inf) and converted to FP16 rounding toward zero.
// The key is thus a non-decreasing function of the value's bits read as a
// signed integer, from 0x0000 up to +inf = 0x7C00, and the FP16 bits order
// like the values as unsigned integers.
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float key_clamp(float v) {
    const int b = __float_as_int(v);
    return __int_as_float(min(max(b, 0), 0x7F800000));
}
__device__ __forceinline__ uint32_t key16x2(float a, float b) {
    return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(key_clamp(a), key_clamp(b)));
}
// the key of v (in the lower 16 bits; the upper 16 are zero)
__device__ __forceinline__ uint32_t key16(float v) {
    return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(key_clamp(v), 0.f));
}
// the largest signed bit pattern b with key(b) <= theta: key(b) <= theta
// exactly when (int)b <= key_bound(theta). Below +inf, theta + 1 is a
// representable FP16 value V, RTZ(v) <= theta exactly when v < V, and
// positive floats order like their bits; from +inf up every key passes.
__device__ __forceinline__ int key_bound(uint32_t theta) {
    if (theta >= 0x7C00u) return 0x7FFFFFFF;
    const float v = (float)__builtin_bit_cast(_Float16, (unsigned short)(theta + 1u));
    return __float_as_int(v) - 1;
}
__device__ __forceinline__ uint32_t min_u16x2(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(u16x2, a),
                                                                  __builtin_bit_cast(u16x2, b)));
}

// a pair's slots in its query's row: its list's length rounded up to 8
__device__ __forceinline__ int64_t row_slots(int64_t len) { return (len + 7) & ~(int64_t)7; }

// The sample of a query: every entry of its nearest probed lists -- the
// probes of rank < SAMP_MIN_PROBES, and on until the lists before the next
// probe hold SAMP_MULT * k * R entries. The nearest lists hold most of a
// query's best entries, so the (k R)-th smallest key of such a sample lies
// close to that of all its probed entries. (Two lists at least: a sample
// pass 19 % longer and a filter pass 4 % shorter on the 10M-entry
// benchmark, the scan stage 2 % longer.)
constexpr int SAMP_MIN_PROBES = 1;
constexpr int SAMP_MULT = 4;

// a query's rerank depth k * R (k clamped to [1, the batch's k])
__device__ __forceinline__ unsigned query_selk(const ScanArgs &a, int64_t q) {
    int kq = a.k_q ? a.k_q[q] : a.k_default;
    kq = kq < 1 ? 1 : (kq > a.k_default ? a.k_default : kq);
    return (unsigned)(kq * a.rerank);
}

// one wave per query: entries of its probed lists, the pairs of every list
// and of its sample, and the sample's probes (a prefix of the probe order)
// and row slots
__global__ void __launch_bounds__(64) k_pair_count(ScanArgs a, int32_t *__restrict__ lcnt,
                                                   int32_t *__restrict__ lscnt, int64_t *__restrict__ qtot,
                                                   int32_t *__restrict__ qsamp_np,
                                                   int64_t *__restrict__ qsamp_slots) {
    const int64_t q = blockIdx.x;
    const int lane = threadIdx.x;
    const int npq = clamp_nprobe(a, q);
    const int64_t target = (int64_t)SAMP_MULT * query_selk(a, q);
    int64_t tot = 0, ssl = 0;
    int snp = 0;
    for (int l0 = 0; l0 < npq; l0 += 64) {
        const int li = l0 + lane;
        int64_t l = 0, len = 0;
        if (li < npq) {
            l = a.probes[q * a.probe_stride + li];
            len = a.list_len[l];
        }
        int64_t incl = len;
        for (int o = 1; o < 64; o <<= 1) {
            const int64_t v = __shfl_up(incl, o);
            if (lane >= o) incl += v;
        }
        // sampled: a low rank, or fewer than target entries before it
        const bool smp = li < npq && (li < SAMP_MIN_PROBES || tot + incl - len < target);
        if (li < npq && len > 0) {
            atomicAdd(&lcnt[l], 1);
            if (smp) atomicAdd(&lscnt[l], 1);
        }
        int64_t ss = smp ? row_slots(len) : 0;
        for (int o = 32; o > 0; o >>= 1) ss += __shfl_xor(ss, o);
        snp += __popcll(__ballot(smp));
        tot += __shfl(incl, 63);
        ssl += ss;
    }
    if (lane == 0) {
        qtot[q] = tot;
        qsamp_np[q] = snp;
        qsamp_slots[q] = ssl;
    }
}

// one workgroup of 1024: chunked exclusive scan of the row sizes
__global__ void __launch_bounds__(1024) k_row_offsets(const int64_t *__restrict__ qtot, int64_t nq,
                                                      int64_t *__restrict__ row_off) {
    __shared__ int64_t s_a[1024];
    const int t = threadIdx.x;
    const int64_t per = (nq + 1023) / 1024;
    const int64_t b0 = t * per, b1 = b0 + per < nq ? b0 + per : nq;
    int64_t sa = 0;
    for (int64_t i = b0; i < b1; ++i) sa += qtot[i];
    s_a[t] = sa;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int64_t va = t >= o ? s_a[t - o] : 0;
        __syncthreads();
        s_a[t] += va;
        __syncthreads();
    }
    int64_t ra = s_a[t] - sa;
    for (int64_t i = b0; i < b1; ++i) {
        row_off[i] = ra;
        ra += qtot[i];
    }
}

// work items of a list: query groups of 2^qg times entry chunks of 2^ch
// (the list-major MFMA scan takes 32 x 2,048, the exact scan smaller items)
__device__ __forceinline__ int32_t list_items(int32_t pairs, int64_t len, ItemShape sh) {
    return pairs > 0 ? ((pairs + (1 << sh.qg) - 1) >> sh.qg) * (int32_t)((len + (1 << sh.ch) - 1) >> sh.ch) : 0;
}

// one workgroup of 1024: exclusive scans of lcnt and of the item counts, in
// spans of 16,384 lists. Thread t holds the 16 consecutive lists 16 t .. of
// the span (vector loads and stores), sums them itself, and one block scan
// of the 1,024 thread sums (wave shuffles, then the 16 wave totals by one
// wave) gives each thread its offset, with two block barriers a span
constexpr int PO_PER = 16;   // lists per thread

__global__ void __launch_bounds__(1024) k_pair_offsets(const int32_t *__restrict__ lcnt,
                                                       const int64_t *__restrict__ list_len, int K,
                                                       int32_t *__restrict__ lstart, int32_t *__restrict__ ioff,
                                                       int32_t *__restrict__ lfill, int32_t *__restrict__ lbfill,
                                                       int32_t *__restrict__ totals, ItemShape sh) {
    // the waves' totals, then their exclusive prefixes (the span's carry
    // included)
    __shared__ int32_t s_a[16], s_b[16];
    __shared__ int32_t s_ca, s_cb;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    int32_t carry_a = 0, carry_b = 0;
    for (int g0 = 0; g0 < K; g0 += 1024 * PO_PER) {
        const int l0 = g0 + t * PO_PER;
        int32_t v[PO_PER], it[PO_PER];
        const bool whole = l0 + PO_PER <= K;
        if (whole) {
            // 64-byte aligned runs (buffers from the allocator, l0 a
            // multiple of 16)
            const int4 *c4 = reinterpret_cast<const int4 *>(lcnt + l0);
            const longlong2 *n2 = reinterpret_cast<const longlong2 *>(list_len + l0);
#pragma unroll
            for (int u = 0; u < PO_PER / 4; ++u) {
                const int4 c = c4[u];
                v[4 * u] = c.x;
                v[4 * u + 1] = c.y;
                v[4 * u + 2] = c.z;
                v[4 * u + 3] = c.w;
            }
#pragma unroll
            for (int u = 0; u < PO_PER / 2; ++u) {
                const longlong2 n = n2[u];
                it[2 * u] = list_items(v[2 * u], n.x, sh);
                it[2 * u + 1] = list_items(v[2 * u + 1], n.y, sh);
            }
        } else {
#pragma unroll
            for (int r = 0; r < PO_PER; ++r) {
                const int l = l0 + r;
                v[r] = l < K ? lcnt[l] : 0;
                it[r] = l < K ? list_items(v[r], list_len[l], sh) : 0;
            }
        }
        int32_t sa = 0, sb = 0;
#pragma unroll
        for (int r = 0; r < PO_PER; ++r) {
            sa += v[r];
            sb += it[r];
        }
        // inclusive scans of the thread sums over the wave's lanes
        int32_t ia = sa, ib = sb;
        for (int o = 1; o < 64; o <<= 1) {
            const int32_t xa = __shfl_up(ia, o), xb = __shfl_up(ib, o);
            if (lane >= o) {
                ia += xa;
                ib += xb;
            }
        }
        if (lane == 63) {
            s_a[w] = ia;
            s_b[w] = ib;
        }
        __syncthreads();
        if (w == 0) {
            // lanes < 16: the waves' totals scanned, made exclusive, plus the
            // carry of the spans before
            const int32_t ra = lane < 16 ? s_a[lane] : 0, rb = lane < 16 ? s_b[lane] : 0;
            int32_t ja = ra, jb = rb;
            for (int o = 1; o < 16; o <<= 1) {
                const int32_t xa = __shfl_up(ja, o), xb = __shfl_up(jb, o);
                if (lane >= o) {
                    ja += xa;
                    jb += xb;
                }
            }
            if (lane < 16) {
                s_a[lane] = carry_a + ja - ra;
                s_b[lane] = carry_b + jb - rb;
            }
            if (lane == 15) {
                s_ca = carry_a + ja;
                s_cb = carry_b + jb;
            }
        }
        __syncthreads();
        int32_t oa = s_a[w] + ia - sa, ob = s_b[w] + ib - sb;
        int32_t xs[PO_PER], ys[PO_PER];
#pragma unroll
        for (int r = 0; r < PO_PER; ++r) {
            xs[r] = oa;
            ys[r] = ob;
            oa += v[r];
            ob += it[r];
        }
        if (whole) {
#pragma unroll
            for (int u = 0; u < PO_PER / 4; ++u) {
                reinterpret_cast<int4 *>(lstart + l0)[u] = make_int4(xs[4 * u], xs[4 * u + 1], xs[4 * u + 2], xs[4 * u + 3]);
                reinterpret_cast<int4 *>(ioff + l0)[u] = make_int4(ys[4 * u], ys[4 * u + 1], ys[4 * u + 2], ys[4 * u + 3]);
                reinterpret_cast<int4 *>(lfill + l0)[u] = make_int4(0, 0, 0, 0);
                reinterpret_cast<int4 *>(lbfill + l0)[u] = make_int4(0, 0, 0, 0);
            }
        } else {
#pragma unroll
            for (int r = 0; r < PO_PER; ++r) {
                const int l = l0 + r;
                if (l < K) {
                    lstart[l] = xs[r];
                    ioff[l] = ys[r];
                    lfill[l] = 0;
                    lbfill[l] = 0;
                }
            }
        }
        carry_a = s_ca;
        carry_b = s_cb;
        __syncthreads();   // s_a, s_b, s_ca, s_cb are rewritten by the next span
    }
    if (t == 0) {
        totals[0] = carry_a;
        totals[1] = carry_b;
        totals[2] = 0;   // the sample pass's items, counted by k_item_hist
    }
}

// one wave per query: its pairs (non-empty probed lists) into their lists'
// pair ranges -- the sample's pairs from the front, the others from the
// back -- each with its coarse rank, the query's key scale and, for a
// sample pair, where its keys start in the query's sample row (the sample
// probes' slot counts summed in probe order; 0 for the other pairs). The scale is the power of two 2^(14 - e) with
// (sqrt(max coarse rank) + max|y|)^2 < 2^e: no probed entry's distance
// (nor its BF16 estimate) reaches FP16's range once scaled, and every
// scaled distance above 2^-28 of that bound keeps FP16's full precision.
__global__ void __launch_bounds__(64) k_pair_scatter(ScanArgs a, const double *__restrict__ probe_rank,
                                                     const unsigned *__restrict__ ymax,
                                                     const int32_t *__restrict__ lstart,
                                                     const int32_t *__restrict__ lcnt,
                                                     const int32_t *__restrict__ qsamp_np,
                                                     const int64_t *__restrict__ row_off, int32_t *__restrict__ lfill,
                                                     int32_t *__restrict__ lbfill,
                                                     int32_t *__restrict__ pair_q, float *__restrict__ pair_qc,
                                                     float *__restrict__ pair_scale, int64_t *__restrict__ pair_off,
                                                     float *__restrict__ q_scale) {
    const int64_t q = blockIdx.x;
    const int lane = threadIdx.x;
    const int npq = clamp_nprobe(a, q);
    auto rank_of = [&](int li, int64_t l) -> float {
        return probe_rank ? (float)probe_rank[q * a.probe_stride + li]
                          : sqdiff_exact(PtrLoad{a.queries + q * a.d}, PtrLoad{a.coarse + l * a.d}, a.d);
    };
    float qcmax = 0.f;
    for (int li = lane; li < npq; li += 64) qcmax = fmaxf(qcmax, rank_of(li, a.probes[q * a.probe_stride + li]));
    for (int o = 32; o > 0; o >>= 1) qcmax = fmaxf(qcmax, __shfl_xor(qcmax, o));
    const float yb = ymax ? __uint_as_float(*ymax) : 0.f;
    const float r = __builtin_sqrtf(fmaxf(qcmax, 0.f)) + yb;
    const float bound = r * r;
    float scale = 1.f;
    if (bound > 0.f && bound < __builtin_inff()) {
        int e = 0;
        (void)__builtin_frexpf(bound, &e);   // bound < 2^e
        e = e < -100 ? -100 : (e > 100 ? 100 : e);
        scale = __builtin_ldexpf(1.f, 14 - e);
    }
    if (q_scale && lane == 0) q_scale[q] = scale;
    int64_t carry = row_off[q];
    const int snp = qsamp_np[q];
    for (int l0 = 0; l0 < npq; l0 += 64) {
        const int li = l0 + lane;
        int64_t l = 0, len = 0;
        if (li < npq) {
            l = a.probes[q * a.probe_stride + li];
            len = a.list_len[l];
        }
        const int64_t slots = row_slots(len);
        int64_t incl = slots;
        for (int o = 1; o < 64; o <<= 1) {
            const int64_t v = __shfl_up(incl, o);
            if (lane >= o) incl += v;
        }
        if (li < npq && len > 0) {
            const int pos = li < snp ? lstart[l] + atomicAdd(&lfill[l], 1)
                                     : lstart[l] + lcnt[l] - 1 - atomicAdd(&lbfill[l], 1);
            pair_q[pos] = (int32_t)q;
            pair_qc[pos] = rank_of(li, l);
            pair_scale[pos] = scale;
            pair_off[pos] = li < snp ? carry + incl - slots : 0;
        }
        carry += __shfl(incl, 63);
    }
}

// one thread per list: its query groups times its entry chunks
// Work items longest first: item weight is its slab count (32-entry slabs
// of its list chunk, 1 .. LCH / 32), bucket 32 - slabs. Items run in bucket
// order, so the long items start first and the short ones fill the end
// (results do not depend on the order: every pair's keys have a fixed place).
constexpr int kItemBuckets = LCH / 32;

__device__ inline int item_bucket(int64_t len, int c, ItemShape sh) {
    const int64_t ch = (int64_t)1 << sh.ch;
    const int64_t e = len - (int64_t)c * ch;
    const int slabs = (int)(((e < ch ? e : ch) + 31) >> 5);
    return kItemBuckets - slabs;
}

// per bucket, the number of items of the batch (hist) and of the sample
// pass (hist_s: items over each list's sample pairs, which lead its pair
// range; their total is added to *n_samp)
__global__ void __launch_bounds__(256) k_item_hist(const int32_t *__restrict__ lcnt,
                                                   const int32_t *__restrict__ lscnt,
                                                   const int64_t *__restrict__ list_len, int K,
                                                   int32_t *__restrict__ hist, int32_t *__restrict__ hist_s,
                                                   int32_t *__restrict__ n_samp, ItemShape sh) {
    __shared__ int32_t s_h[2 * kItemBuckets];
    if (threadIdx.x < 2 * kItemBuckets) s_h[threadIdx.x] = 0;
    __syncthreads();
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    const int gq = 1 << sh.qg;
    const int64_t ch = (int64_t)1 << sh.ch;
    if (l < K && lcnt[l] > 0) {
        const int64_t len = list_len[l];
        const int nqg = (lcnt[l] + gq - 1) / gq, nsg = (lscnt[l] + gq - 1) / gq;
        const int nch = (int)((len + ch - 1) / ch);
        for (int c = 0; c < nch; ++c) {
            atomicAdd(&s_h[item_bucket(len, c, sh)], nqg);
            if (nsg) atomicAdd(&s_h[kItemBuckets + item_bucket(len, c, sh)], nsg);
        }
    }
    __syncthreads();
    const int t = threadIdx.x;
    if (t < 2 * kItemBuckets && s_h[t]) atomicAdd(t < kItemBuckets ? &hist[t] : &hist_s[t - kItemBuckets], s_h[t]);
    if (t == 0) {
        int tot = 0;
        for (int j = 0; j < kItemBuckets; ++j) tot += s_h[kItemBuckets + j];
        if (tot) atomicAdd(n_samp, tot);
    }
}

// the items in bucket order: a block reserves its share of every bucket
// with one atomic per bucket, then places its items in LDS order (the
// batch's items, and the sample pass's in the same way)
__global__ void __launch_bounds__(256) k_make_items(const int32_t *__restrict__ lcnt,
                                                    const int32_t *__restrict__ lscnt,
                                                    const int32_t *__restrict__ lstart,
                                                    const int64_t *__restrict__ list_len, int K,
                                                    const int32_t *__restrict__ hist, int32_t *__restrict__ fill,
                                                    int4 *__restrict__ items, const int32_t *__restrict__ hist_s,
                                                    int32_t *__restrict__ fill_s, int4 *__restrict__ items_s,
                                                    ItemShape sh) {
    __shared__ int32_t s_c[2 * kItemBuckets];
    if (threadIdx.x < 2 * kItemBuckets) s_c[threadIdx.x] = 0;
    __syncthreads();
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    const int gq = 1 << sh.qg;
    const int64_t ch = (int64_t)1 << sh.ch;
    const int n = l < K ? lcnt[l] : 0;
    const int ns = n > 0 ? lscnt[l] : 0;
    const int64_t len = n > 0 ? list_len[l] : 0;
    const int nqg = (n + gq - 1) / gq, nsg = (ns + gq - 1) / gq;
    const int nch = (int)((len + ch - 1) / ch);
    for (int c = 0; c < nch; ++c) {
        atomicAdd(&s_c[item_bucket(len, c, sh)], nqg);
        if (nsg) atomicAdd(&s_c[kItemBuckets + item_bucket(len, c, sh)], nsg);
    }
    __syncthreads();
    if (threadIdx.x < 2 * kItemBuckets) {
        const int set = threadIdx.x / kItemBuckets, b = threadIdx.x % kItemBuckets;
        const int32_t *h = set ? hist_s : hist;
        int off = 0;
        for (int j = 0; j < b; ++j) off += h[j];
        const int v = s_c[threadIdx.x];
        s_c[threadIdx.x] = v ? off + atomicAdd(set ? &fill_s[b] : &fill[b], v) : 0;
    }
    __syncthreads();
    const int ls = n > 0 ? lstart[l] : 0;
    for (int c = 0; c < nch; ++c) {
        const int b = item_bucket(len, c, sh);
        for (int t = 0; t * gq < n; ++t) {
            const int o = atomicAdd(&s_c[b], 1);
            items[o] = make_int4(l, ls + t * gq, n - t * gq < gq ? n - t * gq : gq, (int)(c * ch));
        }
        for (int t = 0; t * gq < ns; ++t) {
            const int o = atomicAdd(&s_c[kItemBuckets + b], 1);
            items_s[o] = make_int4(l, ls + t * gq, ns - t * gq < gq ? ns - t * gq : gq, (int)(c * ch));
        }
    }
}

// LNW waves per workgroup share the codebook in LDS; each wave runs its own
// work items with no block barrier after the codebook is staged. s == 4
// (d = 4 M) and uint8 codes. The codebook sits in LDS as BF16 (hi, and lo
// for the split mode).
//
// A slab is 32 entries, two tiles of 16, times the item's up to 16 queries
// (a list is probed by 12.6 queries on average at the benchmark's shape, so
// a 16-query tile wastes less than a 32-query one). Operand layout of
// v_mfma_f32_16x16x32_bf16 (k-block j < M / 8 of 32 dimensions, lane l,
// group g = l / 16): the contraction runs over the dimensions in the order
// that gives group g the subquantizers g M / 4 .. (g + 1) M / 4 - 1, two per
// k-block: k-block j, group g covers subquantizers g M / 4 + 2 j and + 1,
// i.e. dimensions g M + 8 j .. + 7. So a lane needs one run of M / 4
// contiguous bytes of each of its two entries' code rows (from the
// dword-aligned words that hold it), and the B operand (column l % 16: the
// item's query of that rank) is eight contiguous floats of the query per
// k-block, loaded once per item. The A operand of lane l for tile t is the
// two codewords of entry 16 t + l % 16 of the slab, gathered straight from
// LDS.
//
// D register i of tile t of lane l holds entry 16 t + 4 g + i of the slab for
// query l % 16, so a tile's four registers are four consecutive entries of
// one query's row: in the sample pass each tile becomes one 8-byte store of
// four keys, with no transposition. The entry terms of a slab pass through a
// 32-float LDS slot of the wave, read back as float4 broadcasts.
//
// The filter pass tests a register's value against the lane's bound on the
// value bits (no key conversion), takes the wave's ballot and, where any lane
// passes, appends the passing lanes' records with their row positions to a
// ring of the wave in LDS at (fill + lanes before it); full blocks of 64 are
// written out by the whole wave, at the start of a slab before its loads
// (so the compiler's count of memory operations in flight never has to
// include a store issued after the loads it waits for).
//
// Scheduling: items are split into partitions (item i belongs to partition
// i mod n_parts) with a counter each, and a workgroup takes the items of
// partition blockIdx mod n_parts, so that no single counter takes every
// wave's atomics. Latency: the next item's index, header and pair load while
// an item runs; every slab's codes and terms load two slabs ahead through a
// ring of three register buffers, and the ring runs on into the next item's
// first slabs; waves of several workgroups per CU (the codebook is the only
// LDS) overlap each other's loads.
constexpr int LNW = 8;

// what a launch of the list-major scan writes
enum { KIND_SAMPLE = 1, KIND_FILTER = 2 };
// the filter pass's ring of pending records per wave, (key << 32 | entry
// position) and query; it holds fewer than 64 between slabs, and a
// register adds at most 64
constexpr int RING = 256;

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
// a buffer offset past any key buffer (key buffers hold less than 2^32 - 64
// bytes), for lanes whose store is to be discarded
constexpr uint32_t kNoStore = 0xFFFFFFF0u;

// F8: FP8 (e4m3) operands and v_mfma_f32_32x32x16_fp8_fp8: the codebook
// scaled by la.cb_scale (a power of two), the query by a power of two of its
// own that keeps its B operand within FP8's range (the key scale of the
// query then includes it)
__device__ __forceinline__ float fp8_clamp(float v) { return fminf(fmaxf(v, -448.f), 448.f); }
__device__ __forceinline__ uint32_t pack4_fp8(float a, float b, float c, float d) {
    int p = __builtin_amdgcn_cvt_pk_fp8_f32(fp8_clamp(a), fp8_clamp(b), 0, false);
    p = __builtin_amdgcn_cvt_pk_fp8_f32(fp8_clamp(c), fp8_clamp(d), p, true);
    return (uint32_t)p;
}

template <bool X3, int M, int KIND, bool F8>
__global__ void __launch_bounds__(64 * LNW) k_scan_list(ListScanArgs la) {
    static_assert(!(X3 && F8), "one operand mode");
    extern __shared__ __align__(16) unsigned char s_raw[];
    constexpr int D = 4 * M;
    constexpr int KB = M / 8;          // k-blocks of 32 dimensions (MFMAs per tile)
    constexpr int RUN = M / 4;         // code bytes of a lane group's run
    constexpr int ND = (RUN + (M % 16 ? 2 : 0) + 3) / 4;   // dwords that hold the run
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int P = la.P;
    bf16x4 *sCBh = reinterpret_cast<bf16x4 *>(s_raw);            // [M][P]
    bf16x4 *sCBl = sCBh + (size_t)M * P;                          // [M][P] (split mode)
    uint32_t *sCB8 = reinterpret_cast<uint32_t *>(s_raw);         // [M][P] (FP8 mode)
    const size_t cb_bytes = F8 ? (size_t)M * P * 4 : (X3 ? 2 : 1) * (size_t)M * P * 8;
    float *s_term = reinterpret_cast<float *>(s_raw + cb_bytes) + wave * 32;
    // filter pass: the wave's ring of pending records
    uint64_t *s_ring = reinterpret_cast<uint64_t *>(reinterpret_cast<float *>(s_raw + cb_bytes) +
                                                    LNW * 32) + (size_t)wave * RING;
    uint32_t *s_ringq = reinterpret_cast<uint32_t *>(s_ring + (size_t)(LNW - wave) * RING) + (size_t)wave * RING;
    const uint8_t *codes = reinterpret_cast<const uint8_t *>(la.codes);

    // the codebook, once per workgroup
    for (int t = tid; t < M * P; t += 64 * LNW) {
        const float4 w = *reinterpret_cast<const float4 *>(la.cb + (size_t)t * 4);
        if (F8) {
            const float a = la.cb_scale;
            sCB8[t] = pack4_fp8(w.x * a, w.y * a, w.z * a, w.w * a);
            continue;
        }
        bf16x4 h;
        h[0] = (__bf16)w.x;
        h[1] = (__bf16)w.y;
        h[2] = (__bf16)w.z;
        h[3] = (__bf16)w.w;
        sCBh[t] = h;
        if (X3) {
            bf16x4 g;
            g[0] = (__bf16)(w.x - (float)h[0]);
            g[1] = (__bf16)(w.y - (float)h[1]);
            g[2] = (__bf16)(w.z - (float)h[2]);
            g[3] = (__bf16)(w.w - (float)h[3]);
            sCBl[t] = g;
        }
    }
    __syncthreads();
    const int n_items = *la.n_items;
    // lane: column (query rank in the item) col, k-group grp; the group's run
    // of RUN code bytes starts at byte sb of an entry's code row, in the
    // dword-aligned words loaded from sb & ~3, shifted by sb & 3 bytes
    const int col = lane & 15, grp = lane >> 4;
    const int sb = RUN * grp, sbo = sb & 3;
    // work items of this workgroup's partition: part, part + nparts, ...
    const int nparts = la.n_parts, part = (int)(blockIdx.x % (unsigned)nparts);
    int32_t *ctr = la.item_counter + part;
    // an item's header and this lane's pair (column col of the product);
    // entry positions and row offsets fit in 32 bits (the host checks)
    struct ItemState {
        int e0, eb, ee, slot_end, ns, pc, qid, off;
        float qc, sc;
    };
    auto load_item = [&](int item) -> ItemState {
        ItemState S{0, 0, 0, 0, 0, 0, 0, 0, 0.f, 0.f};
        if (item >= n_items) return S;
        const int4 it = la.items[item];
        S.e0 = (int)la.list_start[it.x];
        const int n = (int)la.list_len[it.x];
        S.eb = it.w;
        S.ee = n < it.w + la.chunk ? n : it.w + la.chunk;
        // groups of four entries starting below slot_end are stored (the
        // last chunk pads its query runs to a multiple of 8 slots)
        S.slot_end = S.ee == n ? (int)row_slots(n) : S.ee;
        S.ns = S.ee > it.w ? (S.ee - it.w + 31) >> 5 : 1;
        S.pc = it.z;
        const int pidx = it.y + (col < it.z ? col : 0);
        S.qid = la.pair_q[pidx];
        S.qc = la.pair_qc[pidx];
        S.sc = la.pair_scale[pidx];
        S.off = (int)la.pair_off[pidx];
        return S;
    };
    // one slab's loads: this lane group's code run of entries b + col and
    // b + 16 + col (tiles 0 and 1) and their terms (rows past the chunk
    // read its last entry, or the entry before an empty chunk, and are
    // discarded, so the loads are unconditional)
    struct Slab {
        uint32_t w[2][ND];
        float t[2];
    };
    auto load_slab = [&](int e0, int b, int ee, Slab &x) {
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int e = b + 16 * t + col < ee ? b + 16 * t + col : ee - 1;
            const uint32_t *cw = reinterpret_cast<const uint32_t *>(codes + (int64_t)(e0 + e) * M + (sb & ~3));
#pragma unroll
            for (int u = 0; u < ND; ++u) x.w[t][u] = cw[u];
            x.t[t] = la.term[e0 + e];
        }
    };
    // item indices are wave-uniform (taken from lane 0 into a scalar), so
    // item headers load through the scalar unit and the slab loop's
    // control flow is uniform
    int j = 0;
    if (lane == 0) j = atomicAdd(ctr, 1);
    int item = part + __builtin_amdgcn_readfirstlane(j) * nparts;
    ItemState S = load_item(item), SN = S;
    // slab s of the stream that runs through the current item and on into
    // the next (past both: a discarded reload of the current item's last)
    auto load_at = [&](int s, Slab &x) {
        const bool cur = s < S.ns, nxt = !cur && s - S.ns < SN.ns;
        const int e0 = nxt ? SN.e0 : S.e0;
        const int b = cur ? S.eb + 32 * s : (nxt ? SN.eb + 32 * (s - S.ns) : S.ee - 1);
        const int ee = nxt ? SN.ee : S.ee;
        load_slab(e0, b, ee, x);
    };
    Slab b0, b1, b2;
    if (item < n_items) {
        load_at(0, b0);
        load_at(1, b1);
    }

    // the key rows as a buffer resource of dist_bytes bytes (raw, gfx9 layout)
    const __amdgpu_buffer_rsrc_t drs =
        __builtin_amdgcn_make_buffer_rsrc(la.dist, (short)0, (int)la.dist_bytes, 0x00020000);
    // the ring's pending records are [head, fill) (wave-uniform counters)
    uint32_t fill = 0, head = 0;
    // the whole wave writes n <= 64 pending records (lane i the i-th), each
    // at the next place of its query's record room (kept while in the room).
    // Lanes are grouped by query first, so that one atomic a query reserves
    // the places of all its records (device-wide atomics on one address
    // are served one at a time).
    auto flush = [&](uint32_t n) {
        wave_sync();   // the ring's appends by other lanes
        const uint32_t slot = (head + (uint32_t)lane) & (RING - 1);
        const uint64_t v = s_ring[slot];
        const uint32_t vq = s_ringq[slot];
        const bool has = (uint32_t)lane < n;
        int lead = 0;
        uint32_t rank = 0, gcnt = 0;
        for (uint64_t todo = __builtin_amdgcn_ballot_w64(has); todo;) {
            const int l0 = __builtin_ctzll(todo);
            const uint32_t ql = (uint32_t)__builtin_amdgcn_readlane((int)vq, l0);
            const uint64_t grp = __builtin_amdgcn_ballot_w64(vq == ql) & todo;
            if (__builtin_amdgcn_inverse_ballot_w64(grp)) {
                lead = l0;
                rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(grp >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)grp, 0u));
            }
            if (lane == l0) gcnt = (uint32_t)__builtin_popcountll(grp);
            todo &= ~grp;
        }
        uint32_t p0 = 0;
        if (has && lane == lead) p0 = atomicAdd(&la.qcnt[vq], gcnt);
        const uint32_t p = (uint32_t)__shfl((int)p0, lead) + rank;
        if (has && p < la.capq) la.rec[(uint64_t)vq * la.capq + p] = v;
        head = (uint32_t)__builtin_amdgcn_readfirstlane((int)(head + n));
        wave_sync();   // the slots are read before they are appended to again
    };

    while (item < n_items) {
        if (lane == 0) j = atomicAdd(ctr, 1);   // the next item, in flight with this one's loads
        const int eb = S.eb, ee = S.ee, slot_end = S.slot_end;
        const bool qv = col < S.pc;
        const int qid = S.qid;
        float sc = S.sc;
        const int off = S.off;
        // filter pass: values with key <= theta survive, i.e. those whose bits
        // are <= tb as a signed integer
        const int tb = (KIND == KIND_FILTER && qv) ? key_bound(la.theta[qid]) : 0;
        const uint64_t qmask = __builtin_amdgcn_ballot_w64(qv);
        const int e0 = S.e0;
        // B operand of k-step jj: the query's dimensions 2 h M + 8 jj .. + 7
        // times -2 sc (a power of two, so the BF16 rounding is scaled
        // exactly); the accumulator starts at (t + |q - c|^2) sc, so it ends
        // at the scaled key ((qc + t) - 2 <q, y>) sc
        bf16x8 bq[KB], bl[X3 ? KB : 1];
        long bq8[F8 ? KB : 1];
        if (F8) {
            const float4 *qr = reinterpret_cast<const float4 *>(la.queries + (int64_t)qid * D) + grp * (M / 4);
            float mx = 0.f;
#pragma unroll
            for (int jj = 0; jj < M / 4; ++jj) {
                const float4 v = qr[jj];
                mx = fmaxf(mx, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))));
            }
            // the query's largest |value| over the four groups; its scale:
            // the largest power of two k <= 1 with 2 sc k max|q| / cb_scale <= 224
            mx = fmaxf(mx, __shfl_xor(mx, 16));
            mx = fmaxf(mx, __shfl_xor(mx, 32));
            const float lim = 224.f * la.cb_scale / (2.f * sc * mx);
            const float kap = mx > 0.f ? fminf(1.f, __builtin_exp2f(__builtin_floorf(__builtin_log2f(lim)))) : 1.f;
            sc *= kap;
            const uint32_t keep = qv ? 0xFFFFFFFFu : 0u;
            const float m2 = -2.f * sc / la.cb_scale;
            auto z = [&](float v) { return __uint_as_float(__float_as_uint(v * m2) & keep); };
#pragma unroll
            for (int jj = 0; jj < KB; ++jj) {
                const float4 x0 = qr[2 * jj], x1 = qr[2 * jj + 1];
                const uint32_t lo = pack4_fp8(z(x0.x), z(x0.y), z(x0.z), z(x0.w));
                const uint32_t hi = pack4_fp8(z(x1.x), z(x1.y), z(x1.z), z(x1.w));
                bq8[jj] = (long)((uint64_t)lo | ((uint64_t)hi << 32));
            }
        } else {
            const float4 *qr = reinterpret_cast<const float4 *>(la.queries + (int64_t)qid * D) + grp * (M / 4);
            // lanes without a query take zeros through a bit mask (a branch
            // that overwrote the loaded registers would first wait for every
            // memory operation of the wave)
            const uint32_t keep = qv ? 0xFFFFFFFFu : 0u;
            const float m2sc = -2.f * sc;
            auto z = [&](float v) { return __uint_as_float(__float_as_uint(v * m2sc) & keep); };
#pragma unroll
            for (int jj = 0; jj < KB; ++jj) {
                const float4 x0 = qr[2 * jj], x1 = qr[2 * jj + 1];
                const float xs[8] = {z(x0.x), z(x0.y), z(x0.z), z(x0.w), z(x1.x), z(x1.y), z(x1.z), z(x1.w)};
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    bq[jj][u] = (__bf16)xs[u];
                    if (X3) bl[jj][u] = (__bf16)(xs[u] - (float)bq[jj][u]);
                }
            }
        }
        const float qcs = S.qc * sc;
        const int next_item = part + __builtin_amdgcn_readfirstlane(j) * nparts;
        SN = load_item(next_item);
        // b1 holds stream slab 1: this item's slab 1, or, for a one-slab
        // item, the next item's slab 0, which was not known when the ring
        // was filled
        if (S.ns == 1) load_at(1, b1);
        // the filter pass's survivors of a slab: register i of tile t holds
        // entry 16 t + 4 grp + i of the slab; CHECK: the slab runs past the
        // chunk's end
        auto survivors = [&](const f32x4 (&acc)[2], int base, auto check) {
            // entry in chunk of tile 0's register 0, and its entry position
            const uint32_t e0r = (uint32_t)(base - eb) + 4u * (uint32_t)grp;
            const uint32_t p0r = (uint32_t)(e0 + base) + 4u * (uint32_t)grp;
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                const float av = acc[r >> 2][r & 3];
                const uint32_t cr = 16u * (uint32_t)(r >> 2) + (uint32_t)(r & 3);
                // the wave's mask of passing lanes (compares straight into
                // a lane mask: 41 and 40 are signed <= and <)
                uint64_t mk = __builtin_amdgcn_sicmp(__float_as_int(av), tb, 41) & qmask;
                if (decltype(check)::value) mk &= __builtin_amdgcn_sicmp((int)(e0r + cr), ee - eb, 40);
                if (mk) {
                    // room for 64 more (taken only when a slab has many)
                    if (fill - head > (uint32_t)(RING - 64)) flush(64);
                    if (__builtin_amdgcn_inverse_ballot_w64(mk)) {
                        const uint32_t pos = __builtin_amdgcn_mbcnt_hi(
                            (uint32_t)(mk >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mk, fill));
                        s_ring[pos & (RING - 1)] = ((uint64_t)key16(av) << 32) | (p0r + cr);
                        s_ringq[pos & (RING - 1)] = (uint32_t)qid;
                    }
                    fill = (uint32_t)__builtin_amdgcn_readfirstlane((int)(fill + (uint32_t)__builtin_popcountll(mk)));
                }
            }
        };
        auto slab = [&](const Slab &x, int s) {
            const int base = eb + 32 * s;
            if (grp == 0) {
                s_term[col] = x.t[0];
                s_term[16 + col] = x.t[1];
            }
            wave_sync();   // s_term
            // register i of tile t of lane l: entry 16 t + 4 (l / 16) + i of
            // the slab, for query l % 16
            f32x4 acc[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const float4 t4 = *reinterpret_cast<const float4 *>(&s_term[16 * t + 4 * grp]);
                acc[t][0] = __builtin_fmaf(t4.x, sc, qcs);
                acc[t][1] = __builtin_fmaf(t4.y, sc, qcs);
                acc[t][2] = __builtin_fmaf(t4.z, sc, qcs);
                acc[t][3] = __builtin_fmaf(t4.w, sc, qcs);
            }
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                // the run's bytes from byte 0: words of four codes
                uint32_t r[ND];
#pragma unroll
                for (int u = 0; u < ND; ++u)
                    r[u] = u + 1 < ND ? __builtin_amdgcn_alignbyte(x.w[t][u + 1], x.w[t][u], (uint32_t)sbo)
                                      : x.w[t][u] >> (8 * sbo);
#pragma unroll
                for (int jj = 0; jj < KB; ++jj) {
                    const int i0 = RUN * grp + 2 * jj;
                    const uint32_t wv = r[jj >> 1] >> (16 * (jj & 1));
                    const uint32_t c0 = wv & 0xFFu, c1 = (wv >> 8) & 0xFFu;
                    if (F8) {
                        const uint64_t a8 = (uint64_t)sCB8[i0 * P + c0] | ((uint64_t)sCB8[(i0 + 1) * P + c1] << 32);
                        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8((long)a8, bq8[jj], acc[t], 0, 0, 0);
                        continue;
                    }
                    const bf16x4 x0 = sCBh[i0 * P + c0], x1 = sCBh[(i0 + 1) * P + c1];
                    const bf16x8 a = {x0[0], x0[1], x0[2], x0[3], x1[0], x1[1], x1[2], x1[3]};
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bq[jj], acc[t], 0, 0, 0);
                    if (X3) {
                        const bf16x4 y0 = sCBl[i0 * P + c0], y1 = sCBl[(i0 + 1) * P + c1];
                        const bf16x8 al = {y0[0], y0[1], y0[2], y0[3], y1[0], y1[1], y1[2], y1[3]};
                        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bq[jj], acc[t], 0, 0, 0);
                        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bl[jj], acc[t], 0, 0, 0);
                    }
                }
            }
            // every entry of the slab lies in the chunk except in its last
            // slab (a wave-uniform test)
            const bool full = base + 32 <= ee;
            if (KIND == KIND_SAMPLE) {
                // the 16-bit keys, two per word: tile g's words hold entries
                // 16 g + 4 grp + (0, 1) and (2, 3)
#pragma unroll
                for (int g = 0; g < 2; ++g) {
                    const int e4 = base + 16 * g + 4 * grp;
                    uint32_t p0 = key16x2(acc[g][0], acc[g][1]), p1 = key16x2(acc[g][2], acc[g][3]);
                    if (!full) {
                        // entries past the chunk: the padding key 0xFFFF,
                        // above every key (keys are at most +inf = 0x7C00)
                        const uint32_t v0 = (e4 + 0 < ee ? 0x0000FFFFu : 0u) | (e4 + 1 < ee ? 0xFFFF0000u : 0u);
                        const uint32_t v1 = (e4 + 2 < ee ? 0x0000FFFFu : 0u) | (e4 + 3 < ee ? 0xFFFF0000u : 0u);
                        p0 = (p0 & v0) | ~v0;
                        p1 = (p1 & v1) | ~v1;
                    }
                    // a buffer store for every lane: a lane without a key to
                    // write takes an offset past the buffer's range, which the
                    // hardware discards. With no branch around the store the
                    // compiler counts the wave's memory operations exactly and
                    // waits only for the loads it needs, not for every store.
                    const uint32_t boff = (qv && e4 < slot_end) ? 2u * (uint32_t)(off + e4) : kNoStore;
                    __builtin_amdgcn_raw_buffer_store_b64(u32x2{p0, p1}, drs, (int)boff, 0, 0);
                }
            } else if (full) {
                survivors(acc, base, std::false_type{});
            } else {
                survivors(acc, base, std::true_type{});
            }
            wave_sync();   // s_term is rewritten by the next slab
        };
        // filter pass: full blocks of pending records, written before the
        // next slab's loads are issued
        auto drain = [&]() {
            if (KIND == KIND_FILTER)
                while (fill - head >= 64u) flush(64);
        };
        // the ring: b0 holds stream slab 0 and b1 stream slab 1; each step
        // loads slab s + 2 into the free buffer and computes slab s; on exit
        // the next item's slabs 0 and 1 are moved into b0 and b1
        const int ns = S.ns;
        int s = 0;
        for (;;) {
            drain();
            load_at(s + 2, b2);
            slab(b0, s);
            if (++s == ns) {
                b0 = b1;
                b1 = b2;
                break;
            }
            drain();
            load_at(s + 2, b0);
            slab(b1, s);
            if (++s == ns) {
                b1 = b0;
                b0 = b2;
                break;
            }
            drain();
            load_at(s + 2, b1);
            slab(b2, s);
            if (++s == ns) break;
        }
        item = next_item;
        S = SN;
    }
    // the last pending records
    if (KIND == KIND_FILTER)
        while (fill != head) flush(fill - head < 64u ? fill - head : 64u);
}

// The list-major exact scan. A work item is a chunk of up to la.chunk
// entries of one list with up to 8 of its queries, and belongs to one wave;
// the workgroup's waves share the FP32 codebook in LDS. Lane l takes entry
// b + l of each 64-entry slab of the chunk and decodes it once into
// registers as decodePQCode + vectorAdd do (pqivf/index.go:678-714): x[t] =
// c[t] + cb_i[code_i][t] in float32. Then, for each query of the item, the
// squared distance in the reference's AVX2 order (32 accumulator slots,
// blocks of 32 elements, d = q - x rounded, fused d * d + acc; slots
// combined (acc0 + acc1) + (acc2 + acc3) per lane, then the horizontal sum),
// bit-identical to the reference and to k_scan's exact mode. Query values
// are wave-uniform and load through the scalar unit (they are the VALU
// instructions' scalar operands).
// The sample pass stores each sample pair's 16-bit keys (of the distance
// times the query's scale) in its run of the query's sample row; the filter
// pass keeps the records (distance bits << 32 | entry position) whose
// scaled distance has a key <= theta, through the same per-wave ring as the
// MFMA scan.
constexpr int XNW = 8;

// wave-uniform reads of buffers the kernel does not write, through the
// constant address space, so that they load through the scalar unit
typedef const __attribute__((address_space(4))) float cfloat;
__device__ __forceinline__ cfloat *as_const(const float *p) {
    return (cfloat *)(const __attribute__((address_space(4))) void *)(uintptr_t)p;
}

// squared distance of a lane's row x and a wave-uniform query q in the
// reference's AVX2 order (32 slots, d = q - x rounded, fused d * d + acc,
// (acc0 + acc1) + (acc2 + acc3) per lane, the horizontal sum); D a multiple
// of 32, q's values scalar operands
template <int D>
__device__ __forceinline__ float sqdiff_scalar_q(const float (&x)[D], cfloat *qp) {
    float acc[32];
#pragma unroll
    for (int c = 0; c < 32; ++c) acc[c] = 0.f;
#pragma unroll
    for (int blk = 0; blk < D / 32; ++blk) {
#pragma unroll
        for (int c = 0; c < 32; ++c) {
            const float dd = qp[32 * blk + c] - x[32 * blk + c];
            acc[c] = __builtin_fmaf(dd, dd, acc[c]);
        }
    }
    float s8[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) s8[u] = (acc[u] + acc[8 + u]) + (acc[16 + u] + acc[24 + u]);
    return hsum8(s8);
}

template <int D, int S, int KIND>
__global__ void __launch_bounds__(64 * XNW) k_scan_exact(ListScanArgs la) {
    extern __shared__ __align__(16) unsigned char s_raw[];
    constexpr int M = D / S;
    static_assert(D % 32 == 0 && (S == 4 || S == 8) && M % 4 == 0, "shape");
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int P = la.P;
    float *sCB = reinterpret_cast<float *>(s_raw);   // [M][P][S]
    uint64_t *ring0 = reinterpret_cast<uint64_t *>(sCB + (size_t)M * P * S);
    uint64_t *s_ring = ring0 + (size_t)wave * RING;
    uint32_t *s_ringq = reinterpret_cast<uint32_t *>(ring0 + (size_t)XNW * RING) + (size_t)wave * RING;
    const uint8_t *codes = reinterpret_cast<const uint8_t *>(la.codes);
    {
        const float4 *src = reinterpret_cast<const float4 *>(la.cb);
        float4 *dst = reinterpret_cast<float4 *>(sCB);
        for (int t = tid; t < M * P * S / 4; t += 64 * XNW) dst[t] = src[t];
    }
    __syncthreads();
    const int n_items = *la.n_items;
    const int nparts = la.n_parts, part = (int)(blockIdx.x % (unsigned)nparts);
    int32_t *ctr = la.item_counter + part;
    const __amdgpu_buffer_rsrc_t drs =
        __builtin_amdgcn_make_buffer_rsrc(la.dist, (short)0, (int)la.dist_bytes, 0x00020000);
    uint32_t fill = 0, head = 0;
    // the whole wave writes n <= 64 pending records (lane i the i-th), each
    // at the next place of its query's record room, one atomic a query
    auto flush = [&](uint32_t n) {
        wave_sync();
        const uint32_t slot = (head + (uint32_t)lane) & (RING - 1);
        const uint64_t v = s_ring[slot];
        const uint32_t vq = s_ringq[slot];
        const bool has = (uint32_t)lane < n;
        int lead = 0;
        uint32_t rank = 0, gcnt = 0;
        for (uint64_t todo = __builtin_amdgcn_ballot_w64(has); todo;) {
            const int l0 = __builtin_ctzll(todo);
            const uint32_t ql = (uint32_t)__builtin_amdgcn_readlane((int)vq, l0);
            const uint64_t grp = __builtin_amdgcn_ballot_w64(vq == ql) & todo;
            if (__builtin_amdgcn_inverse_ballot_w64(grp)) {
                lead = l0;
                rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(grp >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)grp, 0u));
            }
            if (lane == l0) gcnt = (uint32_t)__builtin_popcountll(grp);
            todo &= ~grp;
        }
        uint32_t p0 = 0;
        if (has && lane == lead) p0 = atomicAdd(&la.qcnt[vq], gcnt);
        const uint32_t pp = (uint32_t)__shfl((int)p0, lead) + rank;
        if (has && pp < la.capq) la.rec[(uint64_t)vq * la.capq + pp] = v;
        head = (uint32_t)__builtin_amdgcn_readfirstlane((int)(head + n));
        wave_sync();
    };
    int j = 0;
    if (lane == 0) j = atomicAdd(ctr, 1);
    int item = part + __builtin_amdgcn_readfirstlane(j) * nparts;
    while (item < n_items) {
        if (lane == 0) j = atomicAdd(ctr, 1);
        const int4 it = la.items[item];
        const int l = __builtin_amdgcn_readfirstlane(it.x);
        const int pc = __builtin_amdgcn_readfirstlane(it.z);
        const int eb = __builtin_amdgcn_readfirstlane(it.w);
        const int e0 = (int)la.list_start[l];
        const int n = (int)la.list_len[l];
        const int ee = n < eb + la.chunk ? n : eb + la.chunk;
        const int slot_end = ee == n ? (int)row_slots(n) : ee;
        // lane jq < pc holds the item's query jq: its id, key scale, and its
        // run's start in the sample row or its bound
        const int pidx = it.y + (lane < pc ? lane : 0);
        const int vq = la.pair_q[pidx];
        const float vsc = la.pair_scale[pidx];
        const int vaux = KIND == KIND_SAMPLE ? (int)la.pair_off[pidx] : key_bound(la.theta[vq]);
        const float *cent = la.coarse + (int64_t)l * D;
        for (int b = eb; b < ee; b += 64) {
            if (KIND == KIND_FILTER)
                while (fill - head >= 64u) flush(64);
            const int e = b + lane;
            const bool valid = e < ee;
            const int er = valid ? e : ee - 1;
            float x[D];
            {
                const uint32_t *cw = reinterpret_cast<const uint32_t *>(codes + (int64_t)(e0 + er) * M);
                uint32_t w[M / 4];
#pragma unroll
                for (int u = 0; u < M / 4; ++u) w[u] = cw[u];
#pragma unroll
                for (int i = 0; i < M; ++i) {
                    const uint32_t c = (w[i >> 2] >> (8 * (i & 3))) & 0xFFu;
                    const float4 *cb = reinterpret_cast<const float4 *>(sCB + ((size_t)i * P + c) * S);
#pragma unroll
                    for (int h = 0; h < S / 4; ++h) {
                        const float4 v = cb[h];
                        const int t = i * S + 4 * h;
                        x[t + 0] = cent[t + 0] + v.x;
                        x[t + 1] = cent[t + 1] + v.y;
                        x[t + 2] = cent[t + 2] + v.z;
                        x[t + 3] = cent[t + 3] + v.w;
                    }
                }
            }
            for (int jq = 0; jq < pc; ++jq) {
                const int qid = __builtin_amdgcn_readlane(vq, jq);
                const float sc = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(vsc), jq));
                const int aux = __builtin_amdgcn_readlane(vaux, jq);
                const float key = sqdiff_scalar_q<D>(x, as_const(la.queries + (int64_t)qid * D));
                const float ks = key * sc;
                if (KIND == KIND_SAMPLE) {
                    // entries past the chunk up to the run's slot end take
                    // the padding key 0xFFFF; other lanes' stores are discarded
                    const uint32_t k16 = valid ? key16(ks) : 0xFFFFu;
                    const uint32_t boff = e < slot_end ? 2u * (uint32_t)(aux + e) : kNoStore;
                    __builtin_amdgcn_raw_buffer_store_b16((unsigned short)k16, drs, (int)boff, 0, 0);
                } else {
                    const uint64_t mk = __builtin_amdgcn_ballot_w64(valid && __float_as_int(ks) <= aux);
                    if (mk) {
                        if (fill - head > (uint32_t)(RING - 64)) flush(64);
                        if (__builtin_amdgcn_inverse_ballot_w64(mk)) {
                            const uint32_t pos = __builtin_amdgcn_mbcnt_hi(
                                (uint32_t)(mk >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mk, fill));
                            s_ring[pos & (RING - 1)] = ((uint64_t)__float_as_uint(key) << 32) | (uint32_t)(e0 + e);
                            s_ringq[pos & (RING - 1)] = (uint32_t)qid;
                        }
                        fill = (uint32_t)__builtin_amdgcn_readfirstlane(
                            (int)(fill + (uint32_t)__builtin_popcountll(mk)));
                    }
                }
            }
        }
        item = part + __builtin_amdgcn_readfirstlane(j) * nparts;
    }
    if (KIND == KIND_FILTER)
        while (fill != head) flush(fill - head < 64u ? fill - head : 64u);
}

// ------------------------------------------------- exact top-k over rows --
// The exact FP32 top-k of a dense row matrix (the plain path's coarse step:
// every query's nearest centroids) in the same four steps as the list scans:
// a sample pass over every stride-th row writes each query's 16-bit keys of
// (distance times its scale) into its sample row; k_theta takes the k-th
// smallest; a filter pass over every row keeps the records (distance bits <<
// 32 | row) whose keys are at most that; the exact-record select picks the k
// best by (distance, row); then the rank kernel orders them (a query whose
// records overflowed its room ranks every row). Work item i is rows chunk
// i / ngroups times query group i % ngroups, so that the items in flight at
// once share their rows in L2. Lane l takes row b + l of each 64-row slab
// into registers; the distance of each of the item's queries, in the
// reference's order with the query's values as scalar operands, is the
// reference-order rank bit for bit.
struct TopkRowsArgs {
    const float *queries;     // [nq][D]
    const float *base;        // [nb][D]
    int64_t nq, nb;
    const float *q_scale;     // [nq] key scale (a power of two)
    int64_t stride, ns, ns_pad;   // sample: rows r * stride, r < ns; rows of ns_pad slots
    uint16_t *skeys;          // [nq][ns_pad]
    uint32_t skey_bytes;      // below 2^32 - 64
    const uint32_t *theta;    // [nq]
    uint64_t *rec;            // [nq][capq]
    uint32_t *qcnt;           // [nq], zeroed
    uint32_t capq;
    int qg, chunk;            // queries per item (<= 64), rows per item (a multiple of 64)
    int64_t ngroups;
    int n_items;
    int32_t *item_counter;    // [n_parts], zeroed
    int n_parts;
};

template <int D, int KIND>
__global__ void __launch_bounds__(64 * XNW) k_topk_rows(TopkRowsArgs t) {
    __shared__ uint64_t s_ring0[XNW * RING];
    __shared__ uint32_t s_ringq0[XNW * RING];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    uint64_t *s_ring = s_ring0 + (size_t)wave * RING;
    uint32_t *s_ringq = s_ringq0 + (size_t)wave * RING;
    const int nparts = t.n_parts, part = (int)(blockIdx.x % (unsigned)nparts);
    int32_t *ctr = t.item_counter + part;
    const __amdgpu_buffer_rsrc_t drs =
        __builtin_amdgcn_make_buffer_rsrc(t.skeys, (short)0, (int)t.skey_bytes, 0x00020000);
    const int64_t nrows = KIND == KIND_SAMPLE ? t.ns_pad : t.nb;
    uint32_t fill = 0, head = 0;
    auto flush = [&](uint32_t n) {
        wave_sync();
        const uint32_t slot = (head + (uint32_t)lane) & (RING - 1);
        const uint64_t v = s_ring[slot];
        const uint32_t vq = s_ringq[slot];
        const bool has = (uint32_t)lane < n;
        int lead = 0;
        uint32_t rank = 0, gcnt = 0;
        for (uint64_t todo = __builtin_amdgcn_ballot_w64(has); todo;) {
            const int l0 = __builtin_ctzll(todo);
            const uint32_t ql = (uint32_t)__builtin_amdgcn_readlane((int)vq, l0);
            const uint64_t grp = __builtin_amdgcn_ballot_w64(vq == ql) & todo;
            if (__builtin_amdgcn_inverse_ballot_w64(grp)) {
                lead = l0;
                rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(grp >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)grp, 0u));
            }
            if (lane == l0) gcnt = (uint32_t)__builtin_popcountll(grp);
            todo &= ~grp;
        }
        uint32_t p0 = 0;
        if (has && lane == lead) p0 = atomicAdd(&t.qcnt[vq], gcnt);
        const uint32_t pp = (uint32_t)__shfl((int)p0, lead) + rank;
        if (has && pp < t.capq) t.rec[(uint64_t)vq * t.capq + pp] = v;
        head = (uint32_t)__builtin_amdgcn_readfirstlane((int)(head + n));
        wave_sync();
    };
    int j = 0;
    if (lane == 0) j = atomicAdd(ctr, 1);
    int item = part + __builtin_amdgcn_readfirstlane(j) * nparts;
    while (item < t.n_items) {
        if (lane == 0) j = atomicAdd(ctr, 1);
        const int64_t c = item / t.ngroups, g = item % t.ngroups;
        const int64_t q0 = g * t.qg;
        const int pc = (int)(t.nq - q0 < t.qg ? t.nq - q0 : t.qg);
        const int64_t rb = c * t.chunk, re = rb + t.chunk < nrows ? rb + t.chunk : nrows;
        // lane jq < pc holds query q0 + jq: its scale and bound
        const int vq = (int)(q0 + (lane < pc ? lane : 0));
        const float vsc = t.q_scale[vq];
        const int vaux = KIND == KIND_SAMPLE ? 0 : key_bound(t.theta[vq]);
        for (int64_t b = rb; b < re; b += 64) {
            if (KIND == KIND_FILTER)
                while (fill - head >= 64u) flush(64);
            const int64_t r = b + lane;
            const bool valid = KIND == KIND_SAMPLE ? r < t.ns : r < re;
            const int64_t row = KIND == KIND_SAMPLE ? (r < t.ns ? r : t.ns - 1) * t.stride : (r < re ? r : re - 1);
            float x[D];
            {
                const float4 *xr = reinterpret_cast<const float4 *>(t.base + row * D);
#pragma unroll
                for (int u = 0; u < D / 4; ++u) {
                    const float4 v = xr[u];
                    x[4 * u + 0] = v.x;
                    x[4 * u + 1] = v.y;
                    x[4 * u + 2] = v.z;
                    x[4 * u + 3] = v.w;
                }
            }
            for (int jq = 0; jq < pc; ++jq) {
                const int qid = __builtin_amdgcn_readlane(vq, jq);
                const float sc = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(vsc), jq));
                const float key = sqdiff_scalar_q<D>(x, as_const(t.queries + (int64_t)qid * D));
                const float ks = key * sc;
                if (KIND == KIND_SAMPLE) {
                    // slots past the sample hold the padding key 0xFFFF
                    const uint32_t k16 = valid ? key16(ks) : 0xFFFFu;
                    const uint32_t boff = r < re ? 2u * (uint32_t)((int64_t)qid * t.ns_pad + r) : kNoStore;
                    __builtin_amdgcn_raw_buffer_store_b16((unsigned short)k16, drs, (int)boff, 0, 0);
                } else {
                    const int aux = __builtin_amdgcn_readlane(vaux, jq);
                    const uint64_t mk = __builtin_amdgcn_ballot_w64(valid && __float_as_int(ks) <= aux);
                    if (mk) {
                        if (fill - head > (uint32_t)(RING - 64)) flush(64);
                        if (__builtin_amdgcn_inverse_ballot_w64(mk)) {
                            const uint32_t pos = __builtin_amdgcn_mbcnt_hi(
                                (uint32_t)(mk >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mk, fill));
                            s_ring[pos & (RING - 1)] = ((uint64_t)__float_as_uint(key) << 32) | (uint32_t)row;
                            s_ringq[pos & (RING - 1)] = (uint32_t)qid;
                        }
                        fill = (uint32_t)__builtin_amdgcn_readfirstlane(
                            (int)(fill + (uint32_t)__builtin_popcountll(mk)));
                    }
                }
            }
        }
        item = part + __builtin_amdgcn_readfirstlane(j) * nparts;
    }
    if (KIND == KIND_FILTER)
        while (fill != head) flush(fill - head < 64u ? fill - head : 64u);
}

// the largest row norm of the base (as float bits; *out zeroed before)
// one wave per row in turn (lanes read consecutive columns), 64 rows per
// workgroup; the norms are widened by 2^-10, which covers their FP32 error
constexpr int NORM_ROWS = 64;
__global__ void __launch_bounds__(256) k_row_norm_max(const float *__restrict__ base, int64_t nb, int d,
                                                      unsigned *__restrict__ out) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    float m = 0.f;
    for (int i = wave; i < NORM_ROWS; i += 4) {
        const int64_t r = (int64_t)blockIdx.x * NORM_ROWS + i;
        if (r >= nb) break;
        float s = 0.f;
        for (int e = lane; e < d; e += 64) s = __builtin_fmaf(base[r * d + e], base[r * d + e], s);
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        m = fmaxf(m, __builtin_sqrtf(s) * (1.f + 1.f / 1024.f));
    }
    if (lane == 0) atomicMax(out, __float_as_uint(m));
}

// per query: the key scale 2^(14 - e) with (|q| + max |x|)^2 < 2^e, and the
// select's per-query arrays (row offsets and slot counts of the sample rows,
// the row count)
__global__ void __launch_bounds__(256) k_topk_rows_prep(const float *__restrict__ queries, int64_t nq, int d,
                                                        const unsigned *__restrict__ max_norm, int64_t ns_pad,
                                                        int64_t nb, float *__restrict__ q_scale,
                                                        int64_t *__restrict__ row_off, int64_t *__restrict__ qslots,
                                                        int64_t *__restrict__ qtot) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= nq) return;
    float s = 0.f;
    for (int e = 0; e < d; ++e) s = __builtin_fmaf(queries[q * d + e], queries[q * d + e], s);
    const float r = __builtin_sqrtf(s) * (1.f + 1.f / 1024.f) + __uint_as_float(*max_norm);
    const float bound = r * r;
    float scale = 1.f;
    if (bound > 0.f && bound < __builtin_inff()) {
        int e = 0;
        (void)__builtin_frexpf(bound, &e);
        e = e < -100 ? -100 : (e > 100 ? 100 : e);
        scale = __builtin_ldexpf(1.f, 14 - e);
    }
    q_scale[q] = scale;
    row_off[q] = q * ns_pad;
    qslots[q] = ns_pad;
    qtot[q] = nb;
}

// Selection helpers of the threshold and select kernels: histograms over
// key bits, and the bin where a running count reaches a rank.
// the bin of hist[0 .. nb) where the count reaches need: s_bin, with s_below
// entries before it (wave 0: a block of nb / 64 bins a lane, then the bins
// of the block that holds it, one a lane; nb / 64 <= 64)
__device__ __forceinline__ void bins_find(const unsigned *hist, int nb, unsigned need, int wave, int lane, int *s_bin,
                                          unsigned *s_below) {
    if (wave == 0) {
        const int per = nb / 64;
        unsigned own = 0;
        for (int b = 0; b < per; ++b) own += hist[lane * per + b];
        unsigned incl = own;
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned v = __shfl_up(incl, o);
            if (lane >= o) incl += v;
        }
        const uint64_t hit = __ballot(incl >= need);
        const int blk = hit ? __builtin_ctzll(hit) : 63;
        const unsigned before = __shfl(incl - own, blk);
        const unsigned h = lane < per ? hist[blk * per + lane] : 0u;
        unsigned ih = h;
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned v = __shfl_up(ih, o);
            if (lane >= o) ih += v;
        }
        const uint64_t hit2 = __ballot(lane < per && before + ih >= need);
        const int b = hit2 ? __builtin_ctzll(hit2) : per - 1;
        if (lane == b) {
            *s_bin = blk * per + b;
            *s_below = before + ih - h;
        }
    }
    __syncthreads();
}

constexpr int QB = 8;   // loads in flight per thread in the select's first pass

// the threshold's first histogram: 2,048 bins over key >> 5
constexpr int TH_BINS = 2048;

// One workgroup of 256 per query: theta[q] = the selk-th smallest key of the
// query's sample, the prefix of its row of qsamp_slots[q] slots that the
// sample pass wrote (pad slots hold 0xFFFF, above every key): a histogram
// over key >> 5, then one over the low 5 bits of the bin where the count
// reaches selk. 0xFFFF (every key) when the sample holds fewer than selk
// keys or the query takes the brute-force fallback. The filter pass
// recomputes the sample's keys bit for bit, so at least selk entries of the
// query have keys <= theta[q], and every entry of its selk best (by key,
// then entry position) does.
__global__ void __launch_bounds__(256) k_theta(ScanArgs a, int64_t nq, const uint16_t *__restrict__ keys,
                                               const int64_t *__restrict__ row_off, const int64_t *__restrict__ qtot,
                                               const int64_t *__restrict__ qsamp_slots,
                                               uint32_t *__restrict__ theta) {
    __shared__ unsigned hist[TH_BINS];
    __shared__ unsigned h2[32];
    __shared__ int s_bin;
    __shared__ unsigned s_below, s_cnt;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t q = blockIdx.x;
    const unsigned selk = query_selk(a, q);
    if (a.allow_fallback && qtot[q] < (int64_t)(selk / (unsigned)a.rerank)) {
        if (tid == 0) theta[q] = 0xFFFFu;
        return;
    }
    for (int b = tid; b < TH_BINS; b += 256) hist[b] = 0;
    if (tid < 32) h2[tid] = 0;
    if (tid == 0) s_cnt = 0;
    __syncthreads();
    // rows start at multiples of 8 slots, and the sample is whole lists' runs
    const uint4 *row = reinterpret_cast<const uint4 *>(keys + row_off[q]);
    const int64_t n8 = qsamp_slots[q] >> 3;
    auto each_key = [&](const uint4 &v, auto f) {
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int u = 0; u < 8; ++u) f((w[u >> 1] >> (16 * (u & 1))) & 0xFFFFu);
    };
    unsigned c = 0;
    for (int64_t i = tid; i < n8; i += 512) {
        const uint4 v0 = row[i];
        const uint4 v1 = i + 256 < n8 ? row[i + 256] : make_uint4(~0u, ~0u, ~0u, ~0u);
        auto add = [&](uint32_t k) {
            if (k != 0xFFFFu) {
                atomicAdd(&hist[k >> 5], 1u);
                ++c;
            }
        };
        each_key(v0, add);
        each_key(v1, add);
    }
    atomicAdd(&s_cnt, c);
    __syncthreads();
    if (s_cnt < selk) {
        if (tid == 0) theta[q] = 0xFFFFu;
        return;
    }
    bins_find(hist, TH_BINS, selk, wave, lane, &s_bin, &s_below);
    const uint32_t bin = (uint32_t)s_bin;
    for (int64_t i = tid; i < n8; i += 256)
        each_key(row[i], [&](uint32_t k) {
            if (k >> 5 == bin) atomicAdd(&h2[k & 31u], 1u);
        });
    __syncthreads();
    if (wave == 0) {
        const unsigned h = lane < 32 ? h2[lane] : 0u;
        unsigned incl = h;
        for (int o = 1; o < 32; o <<= 1) {
            const unsigned v = __shfl_up(incl, o);
            if (lane >= o) incl += v;
        }
        const uint64_t hit = __ballot(lane < 32 && s_below + incl >= selk);
        if (lane == 0) theta[q] = (bin << 5) | (uint32_t)(hit ? __builtin_ctzll(hit) : 31);
    }
}

constexpr int SELR_ROOM = 1024;    // records of the boundary bin the select keeps in LDS
constexpr int SELR_RANK = 512;     // boundary bins up to this size are ranked by counting
constexpr int SELR_CACHE = 2048;   // records a query keeps in LDS (more are read again from the row)
constexpr int SELR_BINS = 2048;    // histogram over key >> 5

// One workgroup of 256 per query: the sk = min(k_q * R, probed entries)
// best entries of the query by (key, entry position), from the records the
// filter pass kept (every entry with key <= theta[q]; the sk best are among
// them), held contiguously in the query's record room. A record is
// (key << 32 | entry position): the 16-bit key of the MFMA scan, or, with
// q_scale, the exact FP32 distance's bits (positive floats order like their
// bits), binned by the 16-bit key of the distance times the query's scale
// (a non-decreasing function of it). A 2,048-bin histogram over key >> 5
// gives the bin B where the count reaches sk; records below B are selected,
// and those in B are ranked by their whole value -- by counting for up to
// SELR_RANK of them, else by a sort, or for more than SELR_ROOM by a radix
// select over the 64-bit value. Output: entry positions for the rerank.
// Queries that take the brute-force fallback, and those whose records
// overflowed their room, go to ovf.
__global__ void __launch_bounds__(256) k_select_records(ScanArgs a, int64_t nq, const uint64_t *__restrict__ rec,
                                                        const uint32_t *__restrict__ qcnt, uint32_t capq,
                                                        const int64_t *__restrict__ qtot,
                                                        int64_t *__restrict__ out_pos, int32_t *__restrict__ out_cnt,
                                                        int out_stride, int32_t *__restrict__ ovf,
                                                        int32_t *__restrict__ ovf_cnt,
                                                        const float *__restrict__ q_scale) {
    __shared__ uint64_t cache[SELR_CACHE];   // (key << 32 | entry position)
    __shared__ uint64_t tie[SELR_ROOM];      // the boundary bin's records
    __shared__ unsigned hist[SELR_BINS];
    __shared__ int s_bin, s_out, s_fill;
    __shared__ unsigned s_below;
    __shared__ uint64_t s_prefix;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t q = blockIdx.x;
    int kq = a.k_q ? a.k_q[q] : a.k_default;
    kq = kq < 1 ? 1 : (kq > a.k_default ? a.k_default : kq);
    const int selk = kq * a.rerank;
    const int64_t Lr = qtot[q];
    int64_t *orow = out_pos + q * out_stride;
    const uint32_t T = qcnt[q];
    if ((a.allow_fallback && Lr < kq) || T > capq) {
        if (tid == 0) ovf[atomicAdd(ovf_cnt, 1)] = (int32_t)q;
        return;
    }
    const int sk = (int)(Lr < selk ? Lr : selk);
    for (int t = sk + tid; t < out_stride; t += 256) orow[t] = -1;
    if (tid == 0) out_cnt[q] = sk;
    if (sk == 0) return;
    const float qs = q_scale ? q_scale[q] : 0.f;
    auto bin_of = [&](uint64_t v) -> unsigned {
        if (!q_scale) return (unsigned)(v >> 37);
        return key16(__uint_as_float((uint32_t)(v >> 32)) * qs) >> 5;
    };
    for (int b = tid; b < SELR_BINS; b += 256) hist[b] = 0;
    if (tid == 0) {
        s_out = 0;
        s_fill = 0;
    }
    __syncthreads();
    const uint64_t *rrow = rec + (uint64_t)q * capq;
    auto rec_at = [&](uint32_t t) -> uint64_t { return rrow[t]; };
    // pass 0: the records (eight loads in flight a thread) into the cache
    // and the histogram
    for (uint32_t t0 = tid; t0 < T; t0 += 256 * QB) {
        uint64_t v[QB];
#pragma unroll
        for (int u = 0; u < QB; ++u) {
            const uint32_t t = t0 + 256 * u;
            v[u] = t < T ? rec_at(t) : 0ull;
        }
#pragma unroll
        for (int u = 0; u < QB; ++u) {
            const uint32_t t = t0 + 256 * u;
            if (t >= T) continue;
            if (t < (uint32_t)SELR_CACHE) cache[t] = v[u];
            atomicAdd(&hist[bin_of(v[u])], 1u);
        }
    }
    __syncthreads();
    auto for_rec = [&](auto &&fn) {
        for (uint32_t t = tid; t < T; t += 256) fn(t < (uint32_t)SELR_CACHE ? cache[t] : rec_at(t));
    };
    if (T <= (uint32_t)sk) {
        // every record is selected (T < sk cannot happen: at least
        // min(probed entries, selk) keys are <= theta)
        for_rec([&](uint64_t v) { orow[atomicAdd(&s_out, 1)] = (int64_t)(uint32_t)v; });
        if (tid == 0) out_cnt[q] = (int32_t)T;
        return;
    }
    bins_find(hist, SELR_BINS, (unsigned)sk, wave, lane, &s_bin, &s_below);
    const unsigned B = (unsigned)s_bin;
    const int need = sk - (int)s_below;   // records of bin B to take, 1 .. hist[B]
    const int nb = (int)hist[B];
    const bool room = nb <= SELR_ROOM;
    // pass 1: records below bin B are selected; bin B's go to LDS
    for_rec([&](uint64_t v) {
        const unsigned b = bin_of(v);
        if (b < B)
            orow[atomicAdd(&s_out, 1)] = (int64_t)(uint32_t)v;
        else if (b == B && room)
            tie[atomicAdd(&s_fill, 1)] = v;
    });
    __syncthreads();
    if (room && nb <= SELR_RANK) {
        // a record of bin B is taken when fewer than need of the others
        // come before it by (key, entry position)
        const int o0 = s_out;
        for (int i = tid; i < nb; i += 256) {
            const uint64_t vi = tie[i];
            int rank = 0;
            for (int j2 = 0; j2 < nb; ++j2) rank += tie[j2] < vi ? 1 : 0;
            if (rank < need) orow[o0 + rank] = (int64_t)(uint32_t)vi;
        }
        return;
    }
    if (room) {
        // sort bin B by (key, entry position); the first need complete the selection
        int span = 1;
        while (span < nb) span <<= 1;
        for (int t = nb + tid; t < span; t += 256) tie[t] = ~0ull;
        __syncthreads();
        for (int size = 2; size <= span; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int t = tid; t < (span >> 1); t += 256) {
                    const int i0 = 2 * t - (t & (stride - 1)), i1 = i0 + stride;
                    const bool up = (i0 & size) == 0;
                    const uint64_t va = tie[i0], vb = tie[i1];
                    if (up ? vb < va : va < vb) {
                        tie[i0] = vb;
                        tie[i1] = va;
                    }
                }
                __syncthreads();
            }
        }
        const int o0 = s_out;
        for (int t = tid; t < need; t += 256) orow[o0 + t] = (int64_t)(uint32_t)tie[t];
        return;
    }
    // more than SELR_ROOM records in bin B: the need-th smallest value V of
    // the bin by a radix select over the 64-bit record value, eight bits a
    // pass from the top (values are distinct: entry positions are), then
    // every record of the bin up to V
    uint64_t prefix = 0, mask = 0;
    int kk = need;
    for (int pass = 0; pass < 8; ++pass) {
        const int shift = 56 - 8 * pass;
        __syncthreads();
        hist[tid] = 0;
        __syncthreads();
        for_rec([&](uint64_t v) {
            if (bin_of(v) == B && (v & mask) == prefix) atomicAdd(&hist[(unsigned)(v >> shift) & 255u], 1u);
        });
        __syncthreads();
        if (tid == 0) {
            unsigned run = 0;
            int b = 0;
            while (run + hist[b] < (unsigned)kk) run += hist[b++];
            s_prefix = prefix | ((uint64_t)b << shift);
            s_below = run;
        }
        __syncthreads();
        prefix = s_prefix;
        kk -= (int)s_below;
        mask |= (uint64_t)255 << shift;
    }
    for_rec([&](uint64_t v) {
        if (bin_of(v) == B && v <= prefix) orow[atomicAdd(&s_out, 1)] = (int64_t)(uint32_t)v;
    });
}

}  // namespace dev

void launch_fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(dev::k_fill_i32, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, p, n, v);
    HANN_LAUNCH_CHECK();
}

void launch_pair_count(const ScanArgs &a, int64_t nq, int32_t *lcnt, int32_t *lscnt, int64_t *qtot,
                       int32_t *qsamp_np, int64_t *qsamp_slots, hipStream_t st) {
    if (nq == 0) return;
    hipLaunchKernelGGL(dev::k_pair_count, dim3((unsigned)nq), dim3(64), 0, st, a, lcnt, lscnt, qtot, qsamp_np,
                       qsamp_slots);
    HANN_LAUNCH_CHECK();
}

void launch_row_offsets(const int64_t *qslots, int64_t nq, int64_t *row_off, hipStream_t st) {
    if (nq == 0) return;
    hipLaunchKernelGGL(dev::k_row_offsets, dim3(1), dim3(1024), 0, st, qslots, nq, row_off);
    HANN_LAUNCH_CHECK();
}

void launch_pair_offsets(const int32_t *lcnt, const int64_t *list_len, int K, int32_t *lstart, int32_t *ioff,
                         int32_t *lfill, int32_t *lbfill, int32_t *totals, ItemShape sh, hipStream_t st) {
    hipLaunchKernelGGL(dev::k_pair_offsets, dim3(1), dim3(1024), 0, st, lcnt, list_len, K, lstart, ioff, lfill, lbfill,
                       totals, sh);
    HANN_LAUNCH_CHECK();
}

void launch_pair_scatter(const ScanArgs &a, int64_t nq, const double *probe_rank, const unsigned *ymax,
                         const int32_t *lstart, const int32_t *lcnt, const int32_t *qsamp_np, const int64_t *row_off,
                         int32_t *lfill, int32_t *lbfill, int32_t *pair_q, float *pair_qc, float *pair_scale,
                         int64_t *pair_off, float *q_scale, hipStream_t st) {
    if (nq == 0) return;
    hipLaunchKernelGGL(dev::k_pair_scatter, dim3((unsigned)nq), dim3(64), 0, st, a, probe_rank, ymax, lstart, lcnt,
                       qsamp_np, row_off, lfill, lbfill, pair_q, pair_qc, pair_scale, pair_off, q_scale);
    HANN_LAUNCH_CHECK();
}

void launch_make_items(const int32_t *lcnt, const int32_t *lscnt, const int32_t *lstart, const int64_t *list_len,
                       int K, int32_t *order, int4 *items, int4 *items_s, int32_t *n_samp, ItemShape sh,
                       hipStream_t st) {
    if (sh.qg < 0 || sh.qg > 5 || sh.ch < 6 || sh.ch > dev::LCH_SHIFT)
        fail(HANN_ERR_UNSUPPORTED, "list-major scan: unsupported work-item shape");
    constexpr int B = dev::kItemBuckets;
    int32_t *hist = order, *fill = order + B, *hist_s = order + 2 * B, *fill_s = order + 3 * B;
    hipLaunchKernelGGL(dev::k_item_hist, dim3((unsigned)ceil_div(K, 256)), dim3(256), 0, st, lcnt, lscnt, list_len, K,
                       hist, hist_s, n_samp, sh);
    HANN_LAUNCH_CHECK();
    hipLaunchKernelGGL(dev::k_make_items, dim3((unsigned)ceil_div(K, 256)), dim3(256), 0, st, lcnt, lscnt, lstart,
                       list_len, K, hist, fill, items, hist_s, fill_s, items_s, sh);
    HANN_LAUNCH_CHECK();
}

int list_item_order_ints() { return 4 * dev::kItemBuckets; }

int list_scan_chunk() { return dev::LCH; }

bool list_scan_supported(int d, int m, int P, int code_bytes) {
    if (code_bytes != 1 || d != 4 * m) return false;
    if (m != 8 && m != 16 && m != 24 && m != 32 && m != 48 && m != 64) return false;
    return list_scan_lds_bytes_kind(0, m, P, dev::KIND_FILTER) <= 160 * 1024;
}

void launch_scan_list_kind(const ListScanArgs &la, int x3, int kind, int grid, hipStream_t st) {
    const size_t shm = list_scan_lds_bytes_kind(x3, la.m, la.P, kind);
    if (shm > 160 * 1024) fail(HANN_ERR_UNSUPPORTED, strf("list-major scan needs %zu bytes of LDS", shm));
#define HANN_LIST(X, MM, KK)                                                                                 \
    do {                                                                                                     \
        constexpr bool F8_ = (X) == 2, X3_ = (X) == 1;                                                       \
        HANN_HIP(hipFuncSetAttribute((const void *)dev::k_scan_list<X3_, MM, KK, F8_>,                       \
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));                 \
        hipLaunchKernelGGL((dev::k_scan_list<X3_, MM, KK, F8_>), dim3((unsigned)grid), dim3(64 * dev::LNW),  \
                           shm, st, la);                                                                     \
        HANN_LAUNCH_CHECK();                                                                                 \
    } while (0)
#define HANN_LIST_K(X, MM)                                                             \
    do {                                                                               \
        if (kind == dev::KIND_SAMPLE) HANN_LIST(X, MM, dev::KIND_SAMPLE);              \
        else HANN_LIST(X, MM, dev::KIND_FILTER);                                       \
    } while (0)
#define HANN_LIST_M(X)                                                            \
    do {                                                                          \
        switch (la.m) {                                                           \
        case 8: HANN_LIST_K(X, 8); break;                                         \
        case 16: HANN_LIST_K(X, 16); break;                                       \
        case 24: HANN_LIST_K(X, 24); break;                                       \
        case 32: HANN_LIST_K(X, 32); break;                                       \
        case 48: HANN_LIST_K(X, 48); break;                                       \
        case 64: HANN_LIST_K(X, 64); break;                                       \
        default: fail(HANN_ERR_UNSUPPORTED, "list-major scan: unsupported m");    \
        }                                                                         \
    } while (0)
    if (la.d != 4 * la.m) fail(HANN_ERR_UNSUPPORTED, "list-major scan needs 4-dimensional subvectors");
    // x3: 0 BF16, 1 split BF16, 3 FP8
    if (x3 == 1)
        HANN_LIST_M(1);
    else if (x3 == 3)
        HANN_LIST_M(2);
    else
        HANN_LIST_M(0);
#undef HANN_LIST_M
#undef HANN_LIST_K
#undef HANN_LIST
}

size_t list_scan_lds_bytes_kind(int x3, int m, int P, int kind) {
    // the codebook: BF16 (8 bytes a codeword), split BF16 (16), FP8 (4)
    const size_t per = x3 == 1 ? 16 : (x3 == 3 ? 4 : 8);
    size_t b = per * m * P + (size_t)dev::LNW * 32 * 4;
    if (kind == dev::KIND_FILTER) b += (size_t)dev::LNW * dev::RING * 12;
    return b;
}

bool exact_scan_supported(int d, int m, int P, int code_bytes) {
    if (code_bytes != 1 || m <= 0 || d % m != 0) return false;
    const int s = d / m;
    if (s != 4 && s != 8) return false;
    if (d != 64 && d != 96 && d != 128) return false;
    return exact_scan_lds_bytes(d, m, P) <= 160 * 1024;
}

size_t exact_scan_lds_bytes(int d, int m, int P) {
    return (size_t)m * P * (d / m) * 4 + (size_t)dev::XNW * dev::RING * 12;
}

void launch_scan_exact_kind(const ListScanArgs &la, int kind, int grid, hipStream_t st) {
    if (!exact_scan_supported(la.d, la.m, la.P, 1) || la.chunk % 64 != 0)
        fail(HANN_ERR_UNSUPPORTED, "exact list scan: unsupported shape");
    const size_t shm = exact_scan_lds_bytes(la.d, la.m, la.P);
#define HANN_EXACT(DD, SS, KK)                                                                               \
    do {                                                                                                     \
        HANN_HIP(hipFuncSetAttribute((const void *)dev::k_scan_exact<DD, SS, KK>,                            \
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));                 \
        hipLaunchKernelGGL((dev::k_scan_exact<DD, SS, KK>), dim3((unsigned)grid), dim3(64 * dev::XNW), shm,  \
                           st, la);                                                                          \
        HANN_LAUNCH_CHECK();                                                                                 \
    } while (0)
#define HANN_EXACT_K(DD, SS)                                                        \
    do {                                                                            \
        if (kind == dev::KIND_SAMPLE) HANN_EXACT(DD, SS, dev::KIND_SAMPLE);         \
        else HANN_EXACT(DD, SS, dev::KIND_FILTER);                                  \
    } while (0)
    const int s = la.d / la.m;
    if (la.d == 64 && s == 4) HANN_EXACT_K(64, 4);
    else if (la.d == 64) HANN_EXACT_K(64, 8);
    else if (la.d == 96 && s == 4) HANN_EXACT_K(96, 4);
    else if (la.d == 96) HANN_EXACT_K(96, 8);
    else if (la.d == 128 && s == 4) HANN_EXACT_K(128, 4);
    else HANN_EXACT_K(128, 8);
#undef HANN_EXACT_K
#undef HANN_EXACT
}

void launch_theta(const ScanArgs &a, int64_t nq, const uint16_t *keys, const int64_t *row_off, const int64_t *qtot,
                  const int64_t *qsamp_slots, uint32_t *theta, hipStream_t st) {
    if (nq == 0) return;
    hipLaunchKernelGGL(dev::k_theta, dim3((unsigned)nq), dim3(256), 0, st, a, nq, keys, row_off, qtot, qsamp_slots,
                       theta);
    HANN_LAUNCH_CHECK();
}

size_t select_records_lds_bytes() {
    return (size_t)(dev::SELR_CACHE + dev::SELR_ROOM) * 8 + (size_t)dev::SELR_BINS * 4 + 64;
}

void launch_select_records(const ScanArgs &a, int64_t nq, const uint64_t *rec, const uint32_t *qcnt, uint32_t capq,
                           const int64_t *qtot, int64_t *out_pos, int32_t *out_cnt, int out_stride, int32_t *ovf,
                           int32_t *ovf_cnt, const float *q_scale, hipStream_t st) {
    if (nq == 0) return;
    hipLaunchKernelGGL(dev::k_select_records, dim3((unsigned)nq), dim3(256), 0, st, a, nq, rec, qcnt, capq, qtot,
                       out_pos, out_cnt, out_stride, ovf, ovf_cnt, q_scale);
    HANN_LAUNCH_CHECK();
}

// ---------------------------------------------------- exact top-k rows --
namespace {
struct RowsPlan {
    int64_t stride = 1, ns = 0, ns_pad = 0, capq = 0, qc = 0;
};
RowsPlan rows_plan(int64_t nq, int64_t nb, int k) {
    RowsPlan P;
    // a sample of about a quarter of the rows (at least 8 k, at most 65,536):
    // a query's k-th sample key bounds about its (k stride)-th distance
    const int64_t want = std::max<int64_t>(8 * (int64_t)k, std::min<int64_t>(nb / 4, 65536));
    P.stride = std::max<int64_t>(1, nb / std::max<int64_t>(1, want));
    P.ns = ceil_div(nb, P.stride);
    P.ns_pad = (P.ns + 7) & ~(int64_t)7;
    P.capq = std::max<int64_t>(512, 4 * (int64_t)k * P.stride);
    // queries per round: sample rows and record rooms within 512 MB
    const int64_t per = P.ns_pad * 2 + P.capq * 8 + (int64_t)k * 8 + 64;
    P.qc = std::max<int64_t>(1, std::min<int64_t>(nq, (int64_t(512) << 20) / per));
    P.qc = std::min<int64_t>(P.qc, ((int64_t(1) << 32) - 64) / (2 * P.ns_pad));
    return P;
}
}  // namespace

bool topk_rows_supported(int metric, int64_t nb, int dim, int k) {
    return (metric == HANN_METRIC_SQEUCLIDEAN || metric == HANN_METRIC_EUCLIDEAN) &&
           (dim == 64 || dim == 96 || dim == 128) && k >= 1 && kp_for(k) <= 512 && nb < (int64_t(1) << 31) &&
           nb >= 16 * (int64_t)k;
}

// bases up to this many rows rank an overflowed query's every row on the
// device; larger ones hand such queries to the tiled kernel from the host
constexpr int64_t kRowsRankEveryMax = 65536;

size_t topk_rows_scratch_bytes(int64_t nq, int64_t nb, int k) {
    const RowsPlan P = rows_plan(nq, nb, k);
    return (size_t)P.qc * (size_t)(P.ns_pad * 2 + P.capq * 8 + (int64_t)k * 8 + 64) + 8192 + 3 * 256;
}

void launch_topk_rows(int metric, const float *queries, int64_t nq, const float *base, int64_t nb, int dim, int k,
                      int64_t *out_ids, double *out_dist, Workspace &ws, hipStream_t st) {
    if (nq == 0) return;
    if (!topk_rows_supported(metric, nb, dim, k)) fail(HANN_ERR_UNSUPPORTED, "exact top-k rows: unsupported shape");
    const RowsPlan P = rows_plan(nq, nb, k);
    const int64_t qc = P.qc;
    // arena: the zeroed block (record counts, counters, partitions, the
    // norm bound), then the per-query arrays
    size_t off = 0;
    auto carve = [&](size_t b) {
        const size_t o = off;
        off += (b + 255) & ~(size_t)255;
        return o;
    };
    const size_t o_qcnt = carve((size_t)qc * 4), o_ctr = carve(16), o_ps = carve(kListScanParts * 4),
                 o_pf = carve(kListScanParts * 4), o_norm = carve(16), o_zero_end = off;
    const size_t o_scale = carve((size_t)qc * 4), o_theta = carve((size_t)qc * 4), o_roff = carve((size_t)qc * 8),
                 o_qsl = carve((size_t)qc * 8), o_qtot = carve((size_t)qc * 8), o_cnt = carve((size_t)qc * 4),
                 o_ovf = carve((size_t)qc * 4);
    char *A = ws.get<char>(21, off);
    uint32_t *qcnt = reinterpret_cast<uint32_t *>(A + o_qcnt);
    int32_t *ctr = reinterpret_cast<int32_t *>(A + o_ctr);
    int32_t *parts_s = reinterpret_cast<int32_t *>(A + o_ps);
    int32_t *parts_f = reinterpret_cast<int32_t *>(A + o_pf);
    unsigned *norm = reinterpret_cast<unsigned *>(A + o_norm);
    float *q_scale = reinterpret_cast<float *>(A + o_scale);
    uint32_t *theta = reinterpret_cast<uint32_t *>(A + o_theta);
    int64_t *row_off = reinterpret_cast<int64_t *>(A + o_roff);
    int64_t *qslots = reinterpret_cast<int64_t *>(A + o_qsl);
    int64_t *qtot = reinterpret_cast<int64_t *>(A + o_qtot);
    int32_t *cnt = reinterpret_cast<int32_t *>(A + o_cnt);
    int32_t *ovf = reinterpret_cast<int32_t *>(A + o_ovf);
    uint16_t *skeys = ws.get<uint16_t>(22, (size_t)(qc * P.ns_pad));
    uint64_t *rec = ws.get<uint64_t>(23, (size_t)(qc * P.capq));
    int64_t *pos = ws.get<int64_t>(14, (size_t)(qc * k));
    int dev = 0, cus = 0;
    HANN_HIP(hipGetDevice(&dev));
    HANN_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int grid = cus * 2;
    for (int64_t q0 = 0; q0 < nq; q0 += qc) {
        const int64_t qn = std::min(qc, nq - q0);
        const float *qs = queries + q0 * dim;
        // the zeroed block (the norm bound once, in the first round)
        launch_fill_i32(reinterpret_cast<int32_t *>(A), (int64_t)((q0 == 0 ? o_zero_end : o_norm) / 4), 0, st);
        if (q0 == 0) {
            hipLaunchKernelGGL(dev::k_row_norm_max, dim3((unsigned)ceil_div(nb, (int64_t)dev::NORM_ROWS)), dim3(256),
                               0, st, base, nb, dim, norm);
            HANN_LAUNCH_CHECK();
        }
        hipLaunchKernelGGL(dev::k_topk_rows_prep, dim3((unsigned)ceil_div(qn, 256)), dim3(256), 0, st, qs, qn, dim,
                           norm, P.ns_pad, nb, q_scale, row_off, qslots, qtot);
        HANN_LAUNCH_CHECK();
        dev::TopkRowsArgs t{};
        t.queries = qs;
        t.base = base;
        t.nq = qn;
        t.nb = nb;
        t.q_scale = q_scale;
        t.stride = P.stride;
        t.ns = P.ns;
        t.ns_pad = P.ns_pad;
        t.skeys = skeys;
        t.skey_bytes = (uint32_t)(qn * P.ns_pad * 2);
        t.theta = theta;
        t.rec = rec;
        t.qcnt = qcnt;
        t.capq = (uint32_t)P.capq;
        t.qg = 8;
        t.chunk = 256;
        t.ngroups = ceil_div(qn, (int64_t)t.qg);
        t.n_parts = std::min(kListScanParts, grid);
#define HANN_ROWS(DD, KK, NR, CTR)                                                                   \
    do {                                                                                             \
        t.n_items = (int)(ceil_div((int64_t)(NR), (int64_t)t.chunk) * t.ngroups);                    \
        t.item_counter = CTR;                                                                        \
        hipLaunchKernelGGL((dev::k_topk_rows<DD, KK>), dim3((unsigned)grid), dim3(64 * dev::XNW), 0, \
                           st, t);                                                                   \
        HANN_LAUNCH_CHECK();                                                                         \
    } while (0)
#define HANN_ROWS_D(KK, NR, CTR)                             \
    do {                                                     \
        if (dim == 64) HANN_ROWS(64, KK, NR, CTR);           \
        else if (dim == 96) HANN_ROWS(96, KK, NR, CTR);      \
        else HANN_ROWS(128, KK, NR, CTR);                    \
    } while (0)
        // the sample pass, each query's bound, the filter pass
        HANN_ROWS_D(dev::KIND_SAMPLE, P.ns_pad, parts_s);
        ScanArgs a{};
        a.k_q = nullptr;
        a.k_default = k;
        a.rerank = 1;
        a.allow_fallback = 0;
        launch_theta(a, qn, skeys, row_off, qtot, qslots, theta, st);
        HANN_ROWS_D(dev::KIND_FILTER, nb, parts_f);
#undef HANN_ROWS_D
#undef HANN_ROWS
        // the k best records by (distance, row); a query whose records
        // overflowed keeps the count INT32_MAX, and the rank kernel ranks
        // every row for it
        launch_fill_i32(cnt, qn, INT32_MAX, st);
        launch_select_records(a, qn, rec, qcnt, (uint32_t)P.capq, qtot, pos, cnt, k, ovf, ctr + 1, q_scale, st);
        const bool every = nb <= kRowsRankEveryMax;
        launch_rank_rows64(qs, qn, dim, base, every ? nb : 0, cnt, pos, k, k, nullptr,
                           metric == HANN_METRIC_EUCLIDEAN ? 1 : 0, out_ids + q0 * k, out_dist + q0 * k, st);
        if (every) continue;
        // a large base: the queries whose records overflowed (left untouched
        // by the rank kernel) take the tiled kernel
        std::vector<int32_t> hc((size_t)qn);
        HANN_HIP(hipMemcpyAsync(hc.data(), cnt, (size_t)qn * sizeof(int32_t), hipMemcpyDeviceToHost, st));
        HANN_HIP(hipStreamSynchronize(st));
        std::vector<int64_t> redo;
        for (int64_t q = 0; q < qn; ++q)
            if (hc[q] > k) redo.push_back(q0 + q);
        if (!redo.empty()) topk_tiled_redo(metric, queries, redo, base, nb, dim, k, out_ids, out_dist, ws, st);
    }
}

int select_rows_max_probes() { return dev::SEL_MAXP; }

}  // namespace hann
