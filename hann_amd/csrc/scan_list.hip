// scan_list.hip -- list-major scan: decode each probed list once per batch
// and score it against every query that probes it on the matrix cores.
//
// With the entry's reconstruction y_e = concat_i cb_i[code_e,i] (residual
// part) and the per-entry list term t_e,
//   |q - c_l - y_e|^2 = |q - c_l|^2 + t_e - 2 <q, y_e>
// and for a list l probed by queries Q_l the inner products are one product
// of the decoded list [n_l x d] with Q_l^T [d x |Q_l|]. The pairs (query,
// probed list) of a batch are grouped by list; a work item is one chunk of up
// to 1,024 entries of a list with up to 32 of its queries, and belongs to one
// wave. The wave gathers 32 entries at a time from a BF16 codebook in LDS
// (BF16 hi + lo for the split mode) straight into the A operand of
// v_mfma_f32_32x32x16_bf16 and multiplies them with the item's queries,
// which stay in registers as the B operand. Each list's codes are read once
// per item instead of once per (query, list).
//
// Every (query, entry) distance of the batch is written as a 16-bit key:
// query q owns a row of its probed lists' entries, probes in order, so each
// pair's keys land at a fixed offset and no candidate test, threshold or
// atomic append is involved. A key is the FP16 value of the distance times a
// power of two chosen per query (so that the largest distance any of its
// probed entries can have stays far below FP16's range), with its bits mapped
// so that unsigned order is value order. The scan also keeps each row's
// smallest and largest key. A select kernel then takes, per query, the
// selk = k * R best row entries by (key, position) with a range histogram over
// the row (refined while the boundary bin is too large to sort), and those are
// reranked in exact FP32.
#include "common.h"
#include "scan_list.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

namespace hann {
namespace dev {

constexpr int LQ = 32;          // queries per work item (the MFMA's N)
constexpr int LCH_SHIFT = 11;
constexpr int LCH = 1 << LCH_SHIFT;   // entries per work item (a chunk of the list; a multiple of 32)
constexpr int SEL_BINS = 2048;  // histogram bins of the row select
constexpr int SEL_ROOM = 1024;  // boundary-bin entries the select sorts in LDS (8 workgroups per CU)
constexpr int SEL_MAXP = 2048;  // probes per query in the select's LDS table
constexpr int SEL_SAMP = 8;     // the select's sample: every 8th 128-byte block of a row

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

// 16-bit keys of two scaled distances, packed: each is clamped to
// [0, +inf] on its bits (negative values and -0 become +0, NaN becomes
// +inf) and converted to FP16 rounding toward zero, a monotone map. The
// FP16 bits of non-negative values order like the values as unsigned
// integers, from 0x0000 up to +inf = 0x7C00.
typedef _Float16 half2v __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float key_clamp(float v) {
    const int b = __float_as_int(v);
    const unsigned c = (unsigned)(b > 0 ? b : 0);
    return __uint_as_float(c < 0x7F800000u ? c : 0x7F800000u);
}
__device__ __forceinline__ uint32_t key16x2(float a, float b) {
    return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(key_clamp(a), key_clamp(b)));
}
__device__ __forceinline__ uint32_t min_u16x2(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(u16x2, a),
                                                                  __builtin_bit_cast(u16x2, b)));
}
__device__ __forceinline__ uint32_t max_u16x2(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(u16x2, a),
                                                                  __builtin_bit_cast(u16x2, b)));
}

// a pair's slots in its query's row: its list's length rounded up to 8
__device__ __forceinline__ int64_t row_slots(int64_t len) { return (len + 7) & ~(int64_t)7; }

// one wave per query: entries and row slots of its probed lists
__global__ void __launch_bounds__(64) k_pair_count(ScanArgs a, int32_t *__restrict__ lcnt,
                                                   int64_t *__restrict__ qtot, int64_t *__restrict__ qslots) {
    const int64_t q = blockIdx.x;
    const int lane = threadIdx.x;
    const int npq = clamp_nprobe(a, q);
    int64_t tot = 0, slots = 0;
    for (int li = lane; li < npq; li += 64) {
        const int64_t l = a.probes[q * a.probe_stride + li];
        const int64_t len = a.list_len[l];
        tot += len;
        slots += row_slots(len);
        if (len > 0) atomicAdd(&lcnt[l], 1);
    }
    for (int o = 32; o > 0; o >>= 1) {
        tot += __shfl_xor(tot, o);
        slots += __shfl_xor(slots, o);
    }
    if (lane == 0) {
        qtot[q] = tot;
        qslots[q] = slots;
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

// work items of a list: query groups of 32 times entry chunks
__device__ __forceinline__ int32_t list_items(int32_t pairs, int64_t len) {
    static_assert(LQ == 32 && LCH >= 32, "shifts below");
    return pairs > 0 ? ((pairs + LQ - 1) >> 5) * (int32_t)((len + LCH - 1) >> LCH_SHIFT) : 0;
}

// one workgroup of 1024: exclusive scans of lcnt and of the item counts,
// in rounds of 1024 lists (thread t holds list r * 1024 + t of round r, all
// rounds' loads issued first)
constexpr int PO_ROUNDS = 16;   // lists up to 16,384 in registers; more loop over rounds of that

__global__ void __launch_bounds__(1024) k_pair_offsets(const int32_t *__restrict__ lcnt,
                                                       const int64_t *__restrict__ list_len, int K,
                                                       int32_t *__restrict__ lstart, int32_t *__restrict__ ioff,
                                                       int32_t *__restrict__ lfill, int32_t *__restrict__ totals) {
    __shared__ int32_t s_a[16], s_b[16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    int32_t carry_a = 0, carry_b = 0;
    for (int g0 = 0; g0 < K; g0 += 1024 * PO_ROUNDS) {
        int32_t v[PO_ROUNDS], it[PO_ROUNDS];
#pragma unroll
        for (int r = 0; r < PO_ROUNDS; ++r) {
            const int l = g0 + r * 1024 + t;
            const int lc = l < K ? l : K - 1;
            const int32_t c = lcnt[lc];
            const int64_t n = list_len[lc];
            v[r] = l < K ? c : 0;
            it[r] = l < K ? list_items(c, n) : 0;
        }
#pragma unroll
        for (int r = 0; r < PO_ROUNDS; ++r) {
            // block-wide inclusive scan of (v, it) over the round
            int32_t a = v[r], b = it[r];
            for (int o = 1; o < 64; o <<= 1) {
                const int32_t xa = __shfl_up(a, o), xb = __shfl_up(b, o);
                if (lane >= o) {
                    a += xa;
                    b += xb;
                }
            }
            if (lane == 63) {
                s_a[w] = a;
                s_b[w] = b;
            }
            __syncthreads();
            int32_t wa = 0, wb = 0, ta = 0, tb = 0;
            for (int u = 0; u < 16; ++u) {
                const int32_t ua = s_a[u], ub = s_b[u];
                if (u < w) {
                    wa += ua;
                    wb += ub;
                }
                ta += ua;
                tb += ub;
            }
            __syncthreads();
            const int l = g0 + r * 1024 + t;
            if (l < K) {
                lstart[l] = carry_a + wa + a - v[r];
                ioff[l] = carry_b + wb + b - it[r];
                lfill[l] = 0;
            }
            carry_a += ta;
            carry_b += tb;
        }
    }
    if (t == 0) {
        totals[0] = carry_a;
        totals[1] = carry_b;
    }
}

// one wave per query: its pairs (non-empty probed lists) into their lists'
// pair ranges, each with its coarse rank, the query's key scale and where
// its keys start in the query's row (the probes' slot counts summed in
// probe order). The scale is the power of two 2^(14 - e) with
// (sqrt(max coarse rank) + max|y|)^2 < 2^e: no probed entry's distance
// (nor its BF16 estimate) reaches FP16's range once scaled, and every
// scaled distance above 2^-28 of that bound keeps FP16's full precision.
__global__ void __launch_bounds__(64) k_pair_scatter(ScanArgs a, const double *__restrict__ probe_rank,
                                                     const unsigned *__restrict__ ymax,
                                                     const int32_t *__restrict__ lstart,
                                                     const int64_t *__restrict__ row_off, int32_t *__restrict__ lfill,
                                                     int32_t *__restrict__ pair_q, float *__restrict__ pair_qc,
                                                     float *__restrict__ pair_scale, int64_t *__restrict__ pair_off) {
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
    int64_t carry = row_off[q];
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
            const int pos = lstart[l] + atomicAdd(&lfill[l], 1);
            pair_q[pos] = (int32_t)q;
            pair_qc[pos] = rank_of(li, l);
            pair_scale[pos] = scale;
            pair_off[pos] = carry + incl - slots;
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

__device__ inline int item_bucket(int64_t len, int c) {
    const int64_t e = len - (int64_t)c * LCH;
    const int slabs = (int)(((e < LCH ? e : LCH) + 31) >> 5);
    return kItemBuckets - slabs;
}

// per bucket, the number of items of the batch
__global__ void __launch_bounds__(256) k_item_hist(const int32_t *__restrict__ lcnt,
                                                   const int64_t *__restrict__ list_len, int K,
                                                   int32_t *__restrict__ hist) {
    __shared__ int32_t s_h[kItemBuckets];
    if (threadIdx.x < kItemBuckets) s_h[threadIdx.x] = 0;
    __syncthreads();
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < K && lcnt[l] > 0) {
        const int64_t len = list_len[l];
        const int nqg = (lcnt[l] + LQ - 1) / LQ;
        const int nch = (int)((len + LCH - 1) / LCH);
        for (int c = 0; c < nch; ++c) atomicAdd(&s_h[item_bucket(len, c)], nqg);
    }
    __syncthreads();
    if (threadIdx.x < kItemBuckets && s_h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], s_h[threadIdx.x]);
}

// the items in bucket order: a block reserves its share of every bucket
// with one atomic per bucket, then places its items in LDS order
__global__ void __launch_bounds__(256) k_make_items(const int32_t *__restrict__ lcnt,
                                                    const int32_t *__restrict__ lstart,
                                                    const int64_t *__restrict__ list_len, int K,
                                                    const int32_t *__restrict__ hist, int32_t *__restrict__ fill,
                                                    int4 *__restrict__ items) {
    __shared__ int32_t s_c[kItemBuckets];
    if (threadIdx.x < kItemBuckets) s_c[threadIdx.x] = 0;
    __syncthreads();
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = l < K ? lcnt[l] : 0;
    const int64_t len = n > 0 ? list_len[l] : 0;
    const int nqg = (n + LQ - 1) / LQ;
    const int nch = (int)((len + LCH - 1) / LCH);
    for (int c = 0; c < nch; ++c) atomicAdd(&s_c[item_bucket(len, c)], nqg);
    __syncthreads();
    if (threadIdx.x < kItemBuckets) {
        const int b = threadIdx.x;
        int off = 0;
        for (int j = 0; j < b; ++j) off += hist[j];
        s_c[b] = s_c[b] ? off + atomicAdd(&fill[b], s_c[b]) : 0;
    }
    __syncthreads();
    for (int c = 0; c < nch; ++c) {
        const int b = item_bucket(len, c);
        for (int t = 0; t * LQ < n; ++t) {
            const int o = atomicAdd(&s_c[b], 1);
            items[o] = make_int4(l, lstart[l] + t * LQ, n - t * LQ < LQ ? n - t * LQ : LQ, c * LCH);
        }
    }
}

// LNW waves per workgroup share the codebook in LDS; each wave runs its own
// work items with no block barrier after the codebook is staged. s == 4
// (d = 4 M) and uint8 codes. The codebook sits in LDS as BF16 (hi, and lo
// for the split mode).
//
// Operand layout of v_mfma_f32_32x32x16_bf16 (k-step j < M / 4, lane l,
// h = l / 32): the contraction runs over the dimensions in the order that
// gives half h the subquantizers h M / 2 .. (h + 1) M / 2 - 1, two per
// k-step: k-step j, half h covers subquantizers h M / 2 + 2 j and + 1, i.e.
// dimensions 2 h M + 8 j .. + 7. So a lane needs one contiguous half of its
// entry's code row (M / 2 bytes, one load), and the B operand (column
// l % 32: the item's query of that rank) is eight contiguous floats of the
// query per k-step, loaded once per item. The A operand of lane l is the
// two codewords of entry l % 32 of the slab, gathered straight from LDS.
//
// D register r of lane l holds entry 8 (r / 4) + 4 h + r % 4 of the slab for
// query l % 32, so the four registers r = 4g .. 4g + 3 are four consecutive
// entries of one query's row: each becomes one 8-byte store of four keys,
// with no transposition. The entry terms of a slab pass through a 32-float
// LDS slot of the wave, read back as float4 broadcasts.
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

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
// a buffer offset past any key buffer (key buffers hold less than 2^32 - 64
// bytes), for lanes whose store is to be discarded
constexpr uint32_t kNoStore = 0xFFFFFFF0u;

template <bool X3, int M>
__global__ void __launch_bounds__(64 * LNW) k_scan_list(ListScanArgs la) {
    extern __shared__ __align__(16) unsigned char s_raw[];
    constexpr int D = 4 * M;
    constexpr int HW = M / 8;   // 32-bit words of half a code row
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int P = la.P;
    bf16x4 *sCBh = reinterpret_cast<bf16x4 *>(s_raw);            // [M][P]
    bf16x4 *sCBl = sCBh + (size_t)M * P;                          // [M][P] (split mode)
    float *s_term = reinterpret_cast<float *>(sCBh + (X3 ? 2 : 1) * (size_t)M * P) + wave * 32;
    const uint8_t *codes = reinterpret_cast<const uint8_t *>(la.codes);

    // the codebook, once per workgroup
    for (int t = tid; t < M * P; t += 64 * LNW) {
        const float4 w = *reinterpret_cast<const float4 *>(la.cb + (size_t)t * 4);
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
    const int col = lane & 31, half = lane >> 5;
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
        S.ee = n < it.w + LCH ? n : it.w + LCH;
        // groups of four entries starting below slot_end are stored (the
        // last chunk pads its query runs to a multiple of 8 slots)
        S.slot_end = S.ee == n ? (int)row_slots(n) : S.ee;
        S.ns = (S.ee - it.w + 31) >> 5;
        S.pc = it.z;
        const int pidx = it.y + (col < it.z ? col : 0);
        S.qid = la.pair_q[pidx];
        S.qc = la.pair_qc[pidx];
        S.sc = la.pair_scale[pidx];
        S.off = (int)la.pair_off[pidx];
        return S;
    };
    // one slab's loads: this lane's half code row of entry b + col and
    // that entry's term (rows past the chunk read its last entry and are
    // discarded, so the loads are unconditional)
    struct Slab {
        uint32_t w[HW];
        float t;
    };
    auto load_slab = [&](int e0, int b, int ee, Slab &x) {
        const int e = b + col < ee ? b + col : ee - 1;
        const uint32_t *cw = reinterpret_cast<const uint32_t *>(codes + (int64_t)(e0 + e) * M + half * (M / 2));
#pragma unroll
        for (int u = 0; u < HW; ++u) x.w[u] = cw[u];
        x.t = la.term[e0 + e];
    };
    int j = 0;
    if (lane == 0) j = atomicAdd(ctr, 1);
    int item = part + __shfl(j, 0) * nparts;
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

    while (item < n_items) {
        if (lane == 0) j = atomicAdd(ctr, 1);   // the next item, in flight with this one's loads
        const int pc = S.pc, eb = S.eb, ee = S.ee, slot_end = S.slot_end;
        const bool qv = col < pc;
        const int qid = S.qid;
        const float qc = S.qc, sc = S.sc;
        const int off = S.off;
        // B operand of k-step jj: the query's dimensions 2 h M + 8 jj .. + 7
        bf16x8 bq[M / 4], bl[X3 ? M / 4 : 1];
        {
            const float4 *qr = reinterpret_cast<const float4 *>(la.queries + (int64_t)qid * D) + half * (M / 2);
            // lanes without a query take zeros through a bit mask (a branch
            // that overwrote the loaded registers would first wait for every
            // memory operation of the wave)
            const uint32_t keep = qv ? 0xFFFFFFFFu : 0u;
            auto z = [&](float v) { return __uint_as_float(__float_as_uint(v) & keep); };
#pragma unroll
            for (int jj = 0; jj < M / 4; ++jj) {
                const float4 x0 = qr[2 * jj], x1 = qr[2 * jj + 1];
                const float xs[8] = {z(x0.x), z(x0.y), z(x0.z), z(x0.w), z(x1.x), z(x1.y), z(x1.z), z(x1.w)};
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    bq[jj][u] = (__bf16)xs[u];
                    if (X3) bl[jj][u] = (__bf16)(xs[u] - (float)bq[jj][u]);
                }
            }
        }
        const int next_item = part + __shfl(j, 0) * nparts;
        SN = load_item(next_item);
        // b1 holds stream slab 1: this item's slab 1, or, for a one-slab
        // item, the next item's slab 0, which was not known when the ring
        // was filled
        if (S.ns == 1) load_at(1, b1);
        // key = ((qc + t) - 2 <q, y>) * sc, as fma(-2 sc, acc, t sc + qc sc)
        const float qcs = qc * sc, m2sc = -2.f * sc;
        // smallest and largest key of this lane's query, two halves each
        uint32_t kmin2 = 0xFFFFFFFFu, kmax2 = 0u;
        auto slab = [&](const Slab &x, int s) {
            const int base = eb + 32 * s;
            if (half == 0) s_term[col] = x.t;
            f32x16 acc;
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[r] = 0.f;
#pragma unroll
            for (int jj = 0; jj < M / 4; ++jj) {
                const int i0 = half * (M / 2) + 2 * jj;
                const uint32_t wv = x.w[jj >> 1] >> (16 * (jj & 1));
                const uint32_t c0 = wv & 0xFFu, c1 = (wv >> 8) & 0xFFu;
                const bf16x4 x0 = sCBh[i0 * P + c0], x1 = sCBh[(i0 + 1) * P + c1];
                const bf16x8 a = {x0[0], x0[1], x0[2], x0[3], x1[0], x1[1], x1[2], x1[3]};
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bq[jj], acc, 0, 0, 0);
                if (X3) {
                    const bf16x4 y0 = sCBl[i0 * P + c0], y1 = sCBl[(i0 + 1) * P + c1];
                    const bf16x8 al = {y0[0], y0[1], y0[2], y0[3], y1[0], y1[1], y1[2], y1[3]};
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bq[jj], acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bl[jj], acc, 0, 0, 0);
                }
            }
            wave_sync();   // s_term
            // every entry of the slab lies in the chunk except in its last
            // slab (a wave-uniform test)
            const bool full = base + 32 <= ee;
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const float4 t4 = *reinterpret_cast<const float4 *>(&s_term[8 * g + 4 * half]);
                const int e4 = base + 8 * g + 4 * half;
                uint32_t p0 = key16x2(__builtin_fmaf(m2sc, acc[4 * g + 0], __builtin_fmaf(t4.x, sc, qcs)),
                                      __builtin_fmaf(m2sc, acc[4 * g + 1], __builtin_fmaf(t4.y, sc, qcs)));
                uint32_t p1 = key16x2(__builtin_fmaf(m2sc, acc[4 * g + 2], __builtin_fmaf(t4.z, sc, qcs)),
                                      __builtin_fmaf(m2sc, acc[4 * g + 3], __builtin_fmaf(t4.w, sc, qcs)));
                if (full) {
                    kmin2 = min_u16x2(kmin2, min_u16x2(p0, p1));
                    kmax2 = max_u16x2(kmax2, max_u16x2(p0, p1));
                } else {
                    // entries past the chunk: the padding key 0xFFFF (the
                    // bits of an FP16 NaN, above every stored key, which are
                    // non-negative and at most +inf), and no part in the range
                    const uint32_t v0 = (e4 + 0 < ee ? 0x0000FFFFu : 0u) | (e4 + 1 < ee ? 0xFFFF0000u : 0u);
                    const uint32_t v1 = (e4 + 2 < ee ? 0x0000FFFFu : 0u) | (e4 + 3 < ee ? 0xFFFF0000u : 0u);
                    p0 = (p0 & v0) | ~v0;
                    p1 = (p1 & v1) | ~v1;
                    kmin2 = min_u16x2(kmin2, min_u16x2(p0, p1));
                    kmax2 = max_u16x2(kmax2, max_u16x2(p0 & v0, p1 & v1));
                }
                // a buffer store for every lane: a lane without a key to
                // write takes an offset past the buffer's range, which the
                // hardware discards. With no branch around the store the
                // compiler counts the wave's memory operations exactly and
                // waits only for the loads it needs, not for every store.
                const uint32_t boff = (qv && e4 < slot_end) ? 2u * (uint32_t)(off + e4) : kNoStore;
                __builtin_amdgcn_raw_buffer_store_b64(u32x2{p0, p1}, drs, (int)boff, 0, 0);
            }
            wave_sync();   // s_term is rewritten by the next slab
        };
        // the ring: b0 holds stream slab 0 and b1 stream slab 1; each step
        // loads slab s + 2 into the free buffer and computes slab s; on exit
        // the next item's slabs 0 and 1 are moved into b0 and b1
        const int ns = S.ns;
        int s = 0;
        for (;;) {
            load_at(s + 2, b2);
            slab(b0, s);
            if (++s == ns) {
                b0 = b1;
                b1 = b2;
                break;
            }
            load_at(s + 2, b0);
            slab(b1, s);
            if (++s == ns) {
                b1 = b0;
                b0 = b2;
                break;
            }
            load_at(s + 2, b1);
            slab(b2, s);
            if (++s == ns) break;
        }
        // the row's key range: lanes col and col + 32 share a query
        unsigned kmin = kmin2 & 0xFFFFu, kmax = kmax2 & 0xFFFFu;
        kmin = (kmin2 >> 16) < kmin ? (kmin2 >> 16) : kmin;
        kmax = (kmax2 >> 16) > kmax ? (kmax2 >> 16) : kmax;
        const unsigned omin = __shfl_xor(kmin, 32), omax = __shfl_xor(kmax, 32);
        kmin = omin < kmin ? omin : kmin;
        kmax = omax > kmax ? omax : kmax;
        if (half == 0 && qv && kmin <= kmax) {
            atomicMin(&la.qrange[qid], kmin);
            atomicMax(&la.qrange[la.nq + qid], kmax);
        }
        item = next_item;
        S = SN;
    }
}

// One workgroup of 256 per query: the sk = min(k_q * R, row length) best
// entries of the query's row by (key bits, position).
//  1. A histogram of SEL_BINS bins over the row's key range (from the scan)
//     finds the bin B where the count reaches sk; entries below B are all
//     selected.
//  2. If bin B holds at most SEL_ROOM entries they are sorted in LDS by
//     (key bits, position) and the first ones complete the selection;
//     otherwise the range narrows to bin B and step 1 repeats.
//  3. A bin of one key value with more than SEL_ROOM entries (exact ties)
//     is resolved by a radix select over the positions.
// Positions come from the query's probe table (row offset and first slot
// of each probed list) by binary search.
__global__ void __launch_bounds__(256) k_select_rows(ScanArgs a, int64_t nq, const uint16_t *__restrict__ dist,
                                                     const int64_t *__restrict__ row_off,
                                                     const int64_t *__restrict__ qtot,
                                                     const int64_t *__restrict__ qslots,
                                                     const unsigned *__restrict__ qrange,
                                                     int64_t *__restrict__ out_pos, int32_t *__restrict__ out_cnt,
                                                     int out_stride, int32_t *__restrict__ ovf,
                                                     int32_t *__restrict__ ovf_cnt) {
    extern __shared__ __align__(16) unsigned char s_raw[];
    unsigned *hist = reinterpret_cast<unsigned *>(s_raw);       // [SEL_BINS]
    unsigned *bkey = hist + SEL_BINS;                           // [SEL_ROOM]
    int32_t *bpos = reinterpret_cast<int32_t *>(bkey + SEL_ROOM);   // [SEL_ROOM] entry positions (< 2^31)
    int64_t *poff = reinterpret_cast<int64_t *>(bpos + SEL_ROOM);   // [probe_cap + 1]
    int64_t *pst = poff + a.probe_cap + 1;                      // [probe_cap]
    __shared__ int s_out, s_fill, s_B, s_below, s_nb;
    __shared__ unsigned s_prefix;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t q = blockIdx.x;
    int kq = a.k_q ? a.k_q[q] : a.k_default;
    kq = kq < 1 ? 1 : (kq > a.k_default ? a.k_default : kq);
    const int selk = kq * a.rerank;
    const int64_t Lr = qtot[q], L = qslots[q];   // entries, row slots (a multiple of 8)
    int64_t *orow = out_pos + q * out_stride;
    if (a.allow_fallback && Lr < kq) {
        // the reference's brute-force fallback: the table scan redoes it
        if (tid == 0) ovf[atomicAdd(ovf_cnt, 1)] = (int32_t)q;
        return;
    }
    const int sk = (int)(Lr < selk ? Lr : selk);
    for (int t = sk + tid; t < out_stride; t += 256) orow[t] = -1;
    if (tid == 0) out_cnt[q] = sk;
    if (sk == 0) return;
    const int npq = clamp_nprobe(a, q);
    if (wave == 0) {
        int64_t carry = 0;
        for (int l0 = 0; l0 < npq; l0 += 64) {
            const int li = l0 + lane;
            int64_t len = 0, first = 0;
            if (li < npq) {
                const int64_t l = a.probes[q * a.probe_stride + li];
                len = row_slots(a.list_len[l]);
                first = a.list_start[l];
            }
            int64_t incl = len;
            for (int o = 1; o < 64; o <<= 1) {
                const int64_t v = __shfl_up(incl, o);
                if (lane >= o) incl += v;
            }
            if (li < npq) {
                poff[li] = carry + incl - len;
                pst[li] = first;
            }
            carry += __shfl(incl, 63);
        }
        if (lane == 0) poff[npq] = carry;
    }
    if (tid == 0) s_out = 0;
    __syncthreads();
    const uint16_t *row = dist + row_off[q];
    // slot i of the row -> entry position (the last probe whose row offset is <= i)
    auto pos_of = [&](int64_t i) -> int64_t {
        int lo_j = 0, hi_j = npq - 1;
        while (lo_j < hi_j) {
            const int mid = (lo_j + hi_j + 1) >> 1;
            if (poff[mid] <= i)
                lo_j = mid;
            else
                hi_j = mid - 1;
        }
        return pst[lo_j] + (i - poff[lo_j]);
    };
    // fn(slot, key) over the row (16-byte aligned, a multiple of 8 slots):
    // four 16-byte loads in flight per thread
    constexpr int SU = 4;
    const uint4 *row8 = reinterpret_cast<const uint4 *>(row);
    const int64_t L8 = L / 8;
    auto for_row = [&](auto &&fn) {
        for (int64_t j0 = 0; j0 < L8; j0 += 256 * SU) {
            uint4 v[SU];
#pragma unroll
            for (int u = 0; u < SU; ++u) {
                const int64_t j = j0 + u * 256 + tid;
                v[u] = row8[j < L8 ? j : L8 - 1];
            }
#pragma unroll
            for (int u = 0; u < SU; ++u) {
                const int64_t j = j0 + u * 256 + tid;
                if (j < L8) {
                    const unsigned w[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
#pragma unroll
                    for (int h = 0; h < 4; ++h) {
                        fn(8 * j + 2 * h, w[h] & 0xFFFFu);
                        fn(8 * j + 2 * h + 1, w[h] >> 16);
                    }
                }
            }
        }
    };
    unsigned lo = qrange[q], hi = qrange[nq + q];
    // sort the nb (key, position) pairs in LDS and emit the first need
    // after the s_out already emitted
    auto sort_emit = [&](int nb, int need) {
        __syncthreads();
        int span = 1;
        while (span < nb) span <<= 1;
        for (int t = nb + tid; t < span; t += 256) {
            bkey[t] = 0xFFFFFFFFu;
            bpos[t] = INT32_MAX;
        }
        __syncthreads();
        for (int size = 2; size <= span; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int t = tid; t < (span >> 1); t += 256) {
                    const int i0 = 2 * t - (t & (stride - 1));
                    const int i1 = i0 + stride;
                    const bool up = (i0 & size) == 0;
                    const unsigned ka = bkey[i0], kb = bkey[i1];
                    const int32_t pa = bpos[i0], pb = bpos[i1];
                    const bool b_less = kb < ka || (kb == ka && pb < pa);
                    const bool a_less = ka < kb || (ka == kb && pa < pb);
                    if (up ? b_less : a_less) {
                        bkey[i0] = kb;
                        bkey[i1] = ka;
                        bpos[i0] = pb;
                        bpos[i1] = pa;
                    }
                }
                __syncthreads();
            }
        }
        const int o0 = s_out;
        for (int t = tid; t < need; t += 256) orow[o0 + t] = bpos[t];
    };
    // the bin of a SEL_BINS histogram over [lo, hi] (shift sh) where the
    // cumulative count reaches need: s_B, with s_below entries before it
    // and s_nb in it (wave 0: 32 bins a lane)
    auto find_bin = [&](unsigned need) {
        if (wave == 0) {
            unsigned own = 0;
            for (int b = 0; b < SEL_BINS / 64; ++b) own += hist[lane * (SEL_BINS / 64) + b];
            unsigned incl = own;
            for (int o = 1; o < 64; o <<= 1) {
                const unsigned v = __shfl_up(incl, o);
                if (lane >= o) incl += v;
            }
            const unsigned excl = incl - own;
            if (excl < need && need <= incl) {
                unsigned run = excl;
                int b = lane * (SEL_BINS / 64);
                while (run + hist[b] < need) run += hist[b++];
                s_B = b;
                s_below = (int)run;
                s_nb = (int)hist[b];
            }
        }
        __syncthreads();
    };
    // Sampled threshold: a histogram of every SEL_SAMP-th 128-byte block of
    // the row puts an estimate T above the sk-th smallest key (with margin);
    // one pass then gathers every key <= T into LDS. If at least sk keys
    // (and at most SEL_ROOM) are <= T they hold the sk best by (key,
    // position), which the sort picks; otherwise the exact method below
    // runs. Either way the selection is the same.
    if (2 * sk <= SEL_ROOM && L8 >= 4 * SEL_SAMP && lo <= hi) {
        const unsigned width = hi - lo;
        int sh = 0;
        while ((width >> sh) >= (unsigned)SEL_BINS) ++sh;
        for (int b = tid; b < SEL_BINS; b += 256) hist[b] = 0;
        if (tid == 0) s_B = SEL_BINS - 1;
        __syncthreads();
        // whole 128-byte blocks (8 groups), every SEL_SAMP-th block
        for (int64_t j = (int64_t)(tid >> 3) * 8 * SEL_SAMP + (tid & 7); j < L8; j += 32 * 8 * SEL_SAMP) {
            const uint4 v = row8[j];
            const unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const unsigned u0 = w[h] & 0xFFFFu, u1 = w[h] >> 16;
                if (u0 >= lo && u0 <= hi) atomicAdd(&hist[(u0 - lo) >> sh], 1u);
                if (u1 >= lo && u1 <= hi) atomicAdd(&hist[(u1 - lo) >> sh], 1u);
            }
        }
        __syncthreads();
        // the sample holds about tot / Lr of the row's entries: aim at 1.5
        // times the sk-th's expected rank in it, plus 16
        unsigned tot = 0;
        for (int b = lane; b < SEL_BINS; b += 64) tot += hist[b];
        for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o);
        const uint64_t want = ((uint64_t)3 * sk * tot + 2 * (uint64_t)Lr - 1) / (2 * (uint64_t)Lr) + 16;
        find_bin((unsigned)(want < tot ? want : tot));
        if (tot > 0) {
            const unsigned bspan = (sh >= 32) ? 0xFFFFFFFFu : ((1u << sh) - 1u);
            const unsigned tlo = lo + ((unsigned)s_B << sh);
            const unsigned thr = (hi - tlo) < bspan ? hi : tlo + bspan;
            if (tid == 0) s_fill = 0;
            __syncthreads();
            for_row([&](int64_t i, unsigned u) {
                if (u > thr) return;
                const int slot = atomicAdd(&s_fill, 1);
                if (slot < SEL_ROOM) {
                    bkey[slot] = u;
                    bpos[slot] = (int32_t)pos_of(i);
                }
            });
            __syncthreads();
            const int nb = s_fill;
            if (nb >= sk && nb <= SEL_ROOM) {
                sort_emit(nb, sk);
                return;
            }
        }
        __syncthreads();
    }
    int need = sk;
    for (;;) {
        const unsigned width = hi - lo;
        int sh = 0;
        while ((width >> sh) >= (unsigned)SEL_BINS) ++sh;
        for (int b = tid; b < SEL_BINS; b += 256) hist[b] = 0;
        __syncthreads();
        for_row([&](int64_t, unsigned u) {
            if (u >= lo && u <= hi) atomicAdd(&hist[(u - lo) >> sh], 1u);
        });
        __syncthreads();
        find_bin((unsigned)need);
        const int B = s_B, below = s_below, nb = s_nb;
        const bool fits = nb <= SEL_ROOM;
        if (tid == 0) s_fill = 0;
        __syncthreads();
        // one pass: entries below bin B are selected; bin B goes to LDS when
        // it fits there
        if (below > 0 || fits) {
            for_row([&](int64_t i, unsigned u) {
                if (u < lo || u > hi) return;
                const int b = (int)((u - lo) >> sh);
                if (b < B) {
                    orow[atomicAdd(&s_out, 1)] = pos_of(i);
                } else if (fits && b == B) {
                    const int slot = atomicAdd(&s_fill, 1);
                    bkey[slot] = u;
                    bpos[slot] = (int32_t)pos_of(i);
                }
            });
        }
        need -= below;
        const unsigned blo = lo + ((unsigned)B << sh);
        const unsigned bspan = (sh >= 32) ? 0xFFFFFFFFu : ((1u << sh) - 1u);
        const unsigned bhi = (hi - blo) < bspan ? hi : blo + bspan;
        if (fits) {
            // sort bin B by (key bits, position); the first need complete the row
            sort_emit(nb, need);
            return;
        }
        if (blo == bhi) {
            // more than SEL_ROOM entries share one key: the need smallest
            // positions among them (positions < 2^31: 8 bits a pass)
            unsigned prefix = 0, mask = 0;
            int kk = need;
            for (int pass = 0; pass < 4; ++pass) {
                const int shift = 24 - 8 * pass;
                for (int b = tid; b < 256; b += 256) hist[b] = 0;
                __syncthreads();
                for_row([&](int64_t i, unsigned u) {
                    if (u != blo) return;
                    const unsigned p = (unsigned)pos_of(i);
                    if ((p & mask) == prefix) atomicAdd(&hist[(p >> shift) & 255u], 1u);
                });
                __syncthreads();
                if (tid == 0) {
                    unsigned run = 0;
                    int b = 0;
                    while (run + hist[b] < (unsigned)kk) run += hist[b++];
                    s_prefix = prefix | ((unsigned)b << shift);
                    s_below = (int)run;
                }
                __syncthreads();
                prefix = s_prefix;
                kk -= s_below;
                mask |= 255u << shift;
                __syncthreads();
            }
            // positions are distinct: exactly need of them are <= prefix
            for_row([&](int64_t i, unsigned u) {
                if (u != blo) return;
                const int64_t p = pos_of(i);
                if ((unsigned)p <= prefix) orow[atomicAdd(&s_out, 1)] = p;
            });
            return;
        }
        lo = blo;
        hi = bhi;
        __syncthreads();
    }
}

}  // namespace dev

void launch_fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(dev::k_fill_i32, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, p, n, v);
    HANN_LAUNCH_CHECK();
}

void launch_pair_count(const ScanArgs &a, int64_t nq, int32_t *lcnt, int64_t *qtot, int64_t *qslots,
                       hipStream_t st) {
    if (nq == 0) return;
    hipLaunchKernelGGL(dev::k_pair_count, dim3((unsigned)nq), dim3(64), 0, st, a, lcnt, qtot, qslots);
    HANN_LAUNCH_CHECK();
}

void launch_row_offsets(const int64_t *qslots, int64_t nq, int64_t *row_off, hipStream_t st) {
    if (nq == 0) return;
    hipLaunchKernelGGL(dev::k_row_offsets, dim3(1), dim3(1024), 0, st, qslots, nq, row_off);
    HANN_LAUNCH_CHECK();
}

void launch_pair_offsets(const int32_t *lcnt, const int64_t *list_len, int K, int32_t *lstart, int32_t *ioff,
                         int32_t *lfill, int32_t *totals, hipStream_t st) {
    hipLaunchKernelGGL(dev::k_pair_offsets, dim3(1), dim3(1024), 0, st, lcnt, list_len, K, lstart, ioff, lfill, totals);
    HANN_LAUNCH_CHECK();
}

void launch_pair_scatter(const ScanArgs &a, int64_t nq, const double *probe_rank, const unsigned *ymax,
                         const int32_t *lstart, const int64_t *row_off, int32_t *lfill, int32_t *pair_q,
                         float *pair_qc, float *pair_scale, int64_t *pair_off, hipStream_t st) {
    if (nq == 0) return;
    hipLaunchKernelGGL(dev::k_pair_scatter, dim3((unsigned)nq), dim3(64), 0, st, a, probe_rank, ymax, lstart, row_off,
                       lfill, pair_q, pair_qc, pair_scale, pair_off);
    HANN_LAUNCH_CHECK();
}

void launch_make_items(const int32_t *lcnt, const int32_t *lstart, const int64_t *list_len, int K, int32_t *order,
                       int4 *items, hipStream_t st) {
    int32_t *hist = order, *fill = order + dev::kItemBuckets;
    launch_fill_i32(order, 2 * dev::kItemBuckets, 0, st);
    hipLaunchKernelGGL(dev::k_item_hist, dim3((unsigned)ceil_div(K, 256)), dim3(256), 0, st, lcnt, list_len, K, hist);
    HANN_LAUNCH_CHECK();
    hipLaunchKernelGGL(dev::k_make_items, dim3((unsigned)ceil_div(K, 256)), dim3(256), 0, st, lcnt, lstart, list_len,
                       K, hist, fill, items);
    HANN_LAUNCH_CHECK();
}

int list_item_order_ints() { return 2 * dev::kItemBuckets; }

int list_scan_chunk() { return dev::LCH; }

bool list_scan_supported(int d, int m, int P, int code_bytes) {
    if (code_bytes != 1 || d != 4 * m) return false;
    if (m != 8 && m != 16 && m != 24 && m != 32 && m != 48 && m != 64) return false;
    return list_scan_lds_bytes(1, d, m, P) <= 160 * 1024;
}

size_t list_scan_lds_bytes(int x3, int d, int m, int P) {
    (void)d;
    const int nb = x3 ? 2 : 1;
    return (size_t)nb * m * P * 8 + (size_t)dev::LNW * 32 * 4;
}

int list_scan_threads() { return 64 * dev::LNW; }

void launch_scan_list(const ListScanArgs &la, int x3, int grid, hipStream_t st) {
    const size_t shm = list_scan_lds_bytes(x3, la.d, la.m, la.P);
    if (shm > 160 * 1024) fail(HANN_ERR_UNSUPPORTED, strf("list-major scan needs %zu bytes of LDS", shm));
#define HANN_LIST(X, MM)                                                                                 \
    do {                                                                                                 \
        HANN_HIP(hipFuncSetAttribute((const void *)dev::k_scan_list<X, MM>,                              \
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));             \
        hipLaunchKernelGGL((dev::k_scan_list<X, MM>), dim3((unsigned)grid), dim3(64 * dev::LNW), shm, st, \
                           la);                                                                          \
        HANN_LAUNCH_CHECK();                                                                             \
    } while (0)
#define HANN_LIST_M(X)                                                            \
    do {                                                                          \
        switch (la.m) {                                                           \
        case 8: HANN_LIST(X, 8); break;                                           \
        case 16: HANN_LIST(X, 16); break;                                         \
        case 24: HANN_LIST(X, 24); break;                                         \
        case 32: HANN_LIST(X, 32); break;                                         \
        case 48: HANN_LIST(X, 48); break;                                         \
        case 64: HANN_LIST(X, 64); break;                                         \
        default: fail(HANN_ERR_UNSUPPORTED, "list-major scan: unsupported m");    \
        }                                                                         \
    } while (0)
    if (la.d != 4 * la.m) fail(HANN_ERR_UNSUPPORTED, "list-major scan needs 4-dimensional subvectors");
    if (x3)
        HANN_LIST_M(true);
    else
        HANN_LIST_M(false);
#undef HANN_LIST_M
#undef HANN_LIST
}

int select_rows_max_probes() { return dev::SEL_MAXP; }

void launch_select_rows(const ScanArgs &a, int64_t nq, const uint16_t *dist, const int64_t *row_off,
                        const int64_t *qtot, const int64_t *qslots, const unsigned *qrange, int64_t *out_pos,
                        int32_t *out_cnt, int out_stride, int32_t *ovf, int32_t *ovf_cnt, hipStream_t st) {
    if (nq == 0) return;
    if (a.probe_cap > dev::SEL_MAXP) fail(HANN_ERR_UNSUPPORTED, "list-major select: too many probes per query");
    const size_t shm = (size_t)dev::SEL_BINS * 4 + (size_t)dev::SEL_ROOM * 8 + (size_t)(2 * a.probe_cap + 1) * 8;
    HANN_HIP(hipFuncSetAttribute((const void *)dev::k_select_rows, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)shm));
    hipLaunchKernelGGL(dev::k_select_rows, dim3((unsigned)nq), dim3(256), shm, st, a, nq, dist, row_off, qtot, qslots,
                       qrange, out_pos, out_cnt, out_stride, ovf, ovf_cnt);
    HANN_LAUNCH_CHECK();
}

}  // namespace hann
