// scan_list.hip -- list-major scan: decode each probed list once per batch
// and score it against every query that probes it on the matrix cores.
//
// With the entry's reconstruction y_e = concat_i cb_i[code_e,i] (residual
// part) and the per-entry list term t_e,
//   |q - c_l - y_e|^2 = |q - c_l|^2 + t_e - 2 <q, y_e>
// and for a list l probed by queries Q_l the inner products are one product
// of the decoded list [n_l x d] with Q_l^T [d x |Q_l|]. The pairs (query,
// probed list) of a batch are grouped by list; a work item is one list with
// up to 32 of its queries. A workgroup gathers 32 entries per wave from a
// BF16 codebook in LDS (BF16 hi + lo for the split mode) straight into the
// A operand of v_mfma_f32_32x32x16_bf16 and multiplies them with the item's
// queries. Each list's codes are read once per item instead of once per
// (query, list).
//
// Every (query, entry) distance of the batch is written: query q owns a row
// of its probed lists' entries, probes in order, so each pair's distances
// land at a fixed offset and no candidate test, threshold or atomic append
// is involved. The scan also keeps each row's smallest and largest key. A
// select kernel then takes, per query, the selk = k * R best row entries by
// (key, position) with a range histogram over the row (refined while the
// boundary bin is too large to sort), and those are reranked in exact FP32.
#include "common.h"
#include "scan_list.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

namespace hann {
namespace dev {

constexpr int LQ = 32;          // queries per work item (the MFMA's N)
constexpr int SEL_BINS = 2048;  // histogram bins of the row select
constexpr int SEL_ROOM = 2048;  // boundary-bin entries the select sorts in LDS
constexpr int SEL_MAXP = 2048;  // probes per query in the select's LDS table
// row slots past a list's end: a NaN whose ordered bits (0xFFFFFFFF) exceed
// every stored key's (keys are finite or +inf), so no key range holds it
__device__ constexpr float kPadKey = __builtin_bit_cast(float, 0x7FFFFFFFu);

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

// ordered bits of a key: unsigned order = float order (-0 before +0); the
// row keys are never NaN except the padding sentinel, which maps to ~0u
__device__ __forceinline__ unsigned key_bits(float k) {
    const unsigned u = __float_as_uint(k);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// a pair's slots in its query's row: its list's length rounded up to 4
__device__ __forceinline__ int64_t row_slots(int64_t len) { return (len + 3) & ~(int64_t)3; }

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

// one workgroup of 1024: chunked exclusive scans of lcnt and of the item
// counts ceil(lcnt / 32)
__global__ void __launch_bounds__(1024) k_pair_offsets(const int32_t *__restrict__ lcnt, int K,
                                                       int32_t *__restrict__ lstart, int32_t *__restrict__ ioff,
                                                       int32_t *__restrict__ lfill, int32_t *__restrict__ totals) {
    __shared__ int32_t s_a[1024], s_b[1024];
    const int t = threadIdx.x;
    const int per = (K + 1023) / 1024;
    const int b0 = t * per, b1 = b0 + per < K ? b0 + per : K;
    int32_t sa = 0, sb = 0;
    // 16 loads in flight per thread
    for (int l0 = b0; l0 < b1; l0 += 16) {
        int32_t v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = l0 + u < b1 ? lcnt[l0 + u] : 0;
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            sa += v[u];
            sb += (v[u] + LQ - 1) / LQ;
        }
    }
    s_a[t] = sa;
    s_b[t] = sb;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int32_t va = t >= o ? s_a[t - o] : 0, vb = t >= o ? s_b[t - o] : 0;
        __syncthreads();
        s_a[t] += va;
        s_b[t] += vb;
        __syncthreads();
    }
    int32_t ra = s_a[t] - sa, rb = s_b[t] - sb;
    for (int l0 = b0; l0 < b1; l0 += 16) {
        int32_t v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = l0 + u < b1 ? lcnt[l0 + u] : 0;
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            if (l0 + u < b1) {
                lstart[l0 + u] = ra;
                ioff[l0 + u] = rb;
                lfill[l0 + u] = 0;
            }
            ra += v[u];
            rb += (v[u] + LQ - 1) / LQ;
        }
    }
    if (t == 1023) {
        totals[0] = s_a[1023];
        totals[1] = s_b[1023];
    }
}

// one wave per query: its pairs (non-empty probed lists) into their lists'
// pair ranges, each with its coarse rank and where its distances start in
// the query's row (the probes' slot counts summed in probe order)
__global__ void __launch_bounds__(64) k_pair_scatter(ScanArgs a, const double *__restrict__ probe_rank,
                                                     const int32_t *__restrict__ lstart,
                                                     const int64_t *__restrict__ row_off, int32_t *__restrict__ lfill,
                                                     int32_t *__restrict__ pair_q, float *__restrict__ pair_qc,
                                                     int64_t *__restrict__ pair_off) {
    const int64_t q = blockIdx.x;
    const int lane = threadIdx.x;
    const int npq = clamp_nprobe(a, q);
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
            pair_qc[pos] = probe_rank ? (float)probe_rank[q * a.probe_stride + li]
                                      : sqdiff_exact(PtrLoad{a.queries + q * a.d}, PtrLoad{a.coarse + l * a.d}, a.d);
            pair_off[pos] = carry + incl - slots;
        }
        carry += __shfl(incl, 63);
    }
}

__global__ void k_make_items(const int32_t *__restrict__ lcnt, const int32_t *__restrict__ lstart,
                             const int32_t *__restrict__ ioff, int K, int4 *__restrict__ items) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= K) return;
    const int n = lcnt[l];
    for (int t = 0; t * LQ < n; ++t)
        items[ioff[l] + t] = make_int4(l, lstart[l] + t * LQ, n - t * LQ < LQ ? n - t * LQ : LQ, 0);
}

// LNW = 4 waves per workgroup; s == 4 (d = 4 M) and uint8 codes. The
// codebook sits in LDS as BF16 (hi, and lo for the split mode) for the
// kernel's lifetime; the MFMA's A operand of lane l for k-step j is the two
// codewords of subquantizers 4j + 2h and 4j + 2h + 1 (h = l / 32) of entry
// row l % 32, gathered straight from LDS into registers, and its B operand
// (the item's queries) stays in registers for the whole item.
//
// Each wave turns its 32 x 32 accumulator (entries x queries) around in an
// LDS tile, so that its stores write each query's 32 distances as one run.
//
// Latency: a workgroup takes its next item from the queue while it starts
// the current one, and every thread prefetches its share of the next item
// (its query row slice, the pair's rank and row offset, and the first code
// slab) into registers; inside an item each wave loads its next slab's codes
// and terms while it computes the current one.
constexpr int LNW = 4;
constexpr int LT = 33;   // tile row stride (floats)

template <bool X3, int M>
__global__ void __launch_bounds__(64 * LNW) k_scan_list(ListScanArgs la) {
    extern __shared__ __align__(16) unsigned char s_raw[];
    __shared__ int s_next, s_nitems;
    __shared__ float s_qc[LQ];
    __shared__ int s_qid[LQ];
    __shared__ int64_t s_off[LQ];
    constexpr int D = 4 * M, LD = D + 8;   // LD: Q row stride in BF16 (16-byte rows)
    constexpr int QV = M / 8;              // float4 per thread of a 32 x D query tile (8 threads a row)
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int P = la.P;
    bf16x4 *sCBh = reinterpret_cast<bf16x4 *>(s_raw);            // [M][P]
    bf16x4 *sCBl = sCBh + (size_t)M * P;                          // [M][P] (split mode)
    __bf16 *sQh = reinterpret_cast<__bf16 *>(sCBh + (X3 ? 2 : 1) * (size_t)M * P);
    __bf16 *sQl = sQh + LQ * LD;
    float *s_term = reinterpret_cast<float *>(sQh + (X3 ? 2 : 1) * LQ * LD) + wave * 32;
    float *s_tile = reinterpret_cast<float *>(sQh + (X3 ? 2 : 1) * LQ * LD) + LNW * 32 + wave * (32 * LT);
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
    if (tid == 0) {
        s_nitems = *la.n_items;
        s_next = atomicAdd(la.item_counter, 1);
    }
    __syncthreads();
    const int n_items = s_nitems;
    const int col = lane & 31, half = lane >> 5;
    const int qrow = tid >> 3, qsub = tid & 7;   // this thread's slice of the query tile
    float4 *junk = reinterpret_cast<float4 *>(la.junk) + ((int64_t)blockIdx.x * LNW + wave) * 64 + lane;

    // prefetched state of the next item
    int4 nh = make_int4(0, 0, 0, 0);
    int nqid = -1;
    float nqc = 0.f;
    int64_t noff = 0;
    float4 nqv[QV];
    int64_t ne0 = 0, nn = 0;
    uint32_t nw[M / 4];
    float nt = 0.f;
    // codes of entry b + (lane % 32) and the term of entry b + lane (lane < 32)
    // of a list; rows past its end read the list's first entry (items only
    // hold non-empty lists) and are discarded, so the loads are unconditional
    auto load_codes = [&](int64_t e0, int64_t n, int64_t b, uint32_t *w, float &t) {
        const int64_t e = b + col;
        const uint2 *cw = reinterpret_cast<const uint2 *>(codes + (e0 + (e < n ? e : 0)) * M);
#pragma unroll
        for (int j = 0; j < M / 8; ++j) {
            const uint2 v = cw[j];
            w[2 * j] = v.x;
            w[2 * j + 1] = v.y;
        }
        t = la.term[e0 + ((lane < 32 && b + lane < n) ? b + lane : 0)];
    };
    auto prefetch = [&](int item) {
        if (item >= n_items) return;
        nh = la.items[item];
        nqid = -1;
        if (qrow < nh.z) {
            nqid = la.pair_q[nh.y + qrow];
            const float4 *src = reinterpret_cast<const float4 *>(la.queries + (int64_t)nqid * D) + qsub * QV;
#pragma unroll
            for (int v = 0; v < QV; ++v) nqv[v] = src[v];
            if (qsub == 0) {
                nqc = la.pair_qc[nh.y + qrow];
                noff = la.pair_off[nh.y + qrow];
            }
        }
        ne0 = la.list_start[nh.x];
        nn = la.list_len[nh.x];
        load_codes(ne0, nn, (int64_t)wave * 32, nw, nt);
    };
    int item = s_next;
    prefetch(item);

    for (;;) {
        if (item >= n_items) break;
        const int pc = nh.z;
        const int64_t e0 = ne0, n = nn;
        // the item's queries into LDS (rows >= pc zero, so that their
        // discarded columns stay finite)
        {
            __bf16 *dh = sQh + qrow * LD + qsub * (4 * QV);
            __bf16 *dl = sQl + qrow * LD + qsub * (4 * QV);
#pragma unroll
            for (int v = 0; v < QV; ++v) {
                const float4 x = qrow < pc ? nqv[v] : make_float4(0.f, 0.f, 0.f, 0.f);
                bf16x4 h;
                h[0] = (__bf16)x.x;
                h[1] = (__bf16)x.y;
                h[2] = (__bf16)x.z;
                h[3] = (__bf16)x.w;
                *reinterpret_cast<bf16x4 *>(dh + 4 * v) = h;
                if (X3) {
                    bf16x4 g;
                    g[0] = (__bf16)(x.x - (float)h[0]);
                    g[1] = (__bf16)(x.y - (float)h[1]);
                    g[2] = (__bf16)(x.z - (float)h[2]);
                    g[3] = (__bf16)(x.w - (float)h[3]);
                    *reinterpret_cast<bf16x4 *>(dl + 4 * v) = g;
                }
            }
            if (qsub == 0) {
                s_qid[qrow] = qrow < pc ? nqid : -1;
                s_qc[qrow] = qrow < pc ? nqc : 0.f;
                s_off[qrow] = qrow < pc ? noff : 0;
            }
        }
        struct Slab {
            uint32_t w[M / 4];
            float t;
        };
        Slab sa, sb;
#pragma unroll
        for (int j = 0; j < M / 4; ++j) sa.w[j] = nw[j];
        sa.t = nt;
        if (tid == 0) s_next = atomicAdd(la.item_counter, 1);
        __syncthreads();
        const int next_item = s_next;
        prefetch(next_item);   // in flight during this item
        // the B operands (the item's queries) for the whole item
        bf16x8 bq[M / 4], bl[X3 ? M / 4 : 1];
#pragma unroll
        for (int j = 0; j < M / 4; ++j) {
            bq[j] = *reinterpret_cast<const bf16x8 *>(&sQh[col * LD + 16 * j + 8 * half]);
            if (X3) bl[j] = *reinterpret_cast<const bf16x8 *>(&sQl[col * LD + 16 * j + 8 * half]);
        }
        const float qc_r = s_qc[col];
        // this lane's queries in the store phase are 8t + lane / 8: their
        // smallest and largest keys so far (NaN padding is ignored by min/max)
        float lmin[4], lmax[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            lmin[t] = __builtin_inff();
            lmax[t] = -__builtin_inff();
        }
        // one slab: 32 entries of the list against the item's queries. Every
        // slab issues the same memory operations (masked lanes store to the
        // junk words), so each wait covers only what it needs.
        auto slab = [&](const Slab &x, int64_t base) {
            if (lane < 32) s_term[lane] = x.t;
            f32x16 acc;
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[r] = 0.f;
#pragma unroll
            for (int j = 0; j < M / 4; ++j) {
                const int i0 = 4 * j + 2 * half;
                const uint32_t c0 = (x.w[j] >> (16 * half)) & 0xFFu, c1 = (x.w[j] >> (16 * half + 8)) & 0xFFu;
                const bf16x4 x0 = sCBh[i0 * P + c0], x1 = sCBh[(i0 + 1) * P + c1];
                const bf16x8 a = {x0[0], x0[1], x0[2], x0[3], x1[0], x1[1], x1[2], x1[3]};
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bq[j], acc, 0, 0, 0);
                if (X3) {
                    const bf16x4 y0 = sCBl[i0 * P + c0], y1 = sCBl[(i0 + 1) * P + c1];
                    const bf16x8 al = {y0[0], y0[1], y0[2], y0[3], y1[0], y1[1], y1[2], y1[3]};
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bq[j], acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bl[j], acc, 0, 0, 0);
                }
            }
            wave_sync();   // s_term
            // rows of the accumulator: entries; columns: the item's queries
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = 8 * (r >> 2) + 4 * half + (r & 3);
                // fminf maps a NaN to +inf
                s_tile[col * LT + row] = fminf((qc_r + s_term[row]) - 2.f * acc[r], __builtin_inff());
            }
            wave_sync();   // the tile
            // 16-byte runs of the row: lane -> query 8t + lane / 8, entries
            // base + 4 (lane % 8) .. + 3; slots past the list's end get the
            // padding sentinel (last slab only), and masked lanes store to
            // this wave's junk
            const bool full = base + 32 <= n;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int c = 8 * t + (lane >> 3), g = lane & 7;
                const int64_t e4 = base + 4 * g;
                float4 v;
                v.x = s_tile[c * LT + 4 * g + 0];
                v.y = s_tile[c * LT + 4 * g + 1];
                v.z = s_tile[c * LT + 4 * g + 2];
                v.w = s_tile[c * LT + 4 * g + 3];
                if (!full) {
                    v.x = e4 + 0 < n ? v.x : kPadKey;
                    v.y = e4 + 1 < n ? v.y : kPadKey;
                    v.z = e4 + 2 < n ? v.z : kPadKey;
                    v.w = e4 + 3 < n ? v.w : kPadKey;
                }
                lmin[t] = fminf(lmin[t], fminf(fminf(v.x, v.y), fminf(v.z, v.w)));
                lmax[t] = fmaxf(lmax[t], fmaxf(fmaxf(v.x, v.y), fmaxf(v.z, v.w)));
                float4 *dst = (c < pc && e4 < n) ? reinterpret_cast<float4 *>(la.dist + s_off[c] + e4) : junk;
                *dst = v;
            }
            wave_sync();   // s_term and the tile are rewritten by the next slab
        };
        // ping-pong slab buffers: a slab's loads land in the buffer the
        // slab after it reads, with no register copies between them
        int64_t base = (int64_t)wave * 32;
        while (base < n) {
            load_codes(e0, n, base + 32 * LNW, sb.w, sb.t);
            slab(sa, base);
            base += 32 * LNW;
            if (base >= n) break;
            load_codes(e0, n, base + 32 * LNW, sa.w, sa.t);
            slab(sb, base);
            base += 32 * LNW;
        }
        // the row ranges: lanes 8c' .. 8c' + 7 share their queries; a zero
        // minimum or maximum is widened to cover both signed zeros
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            float mn = lmin[t], mx = lmax[t];
            for (int o = 1; o < 8; o <<= 1) {
                mn = fminf(mn, __shfl_xor(mn, o));
                mx = fmaxf(mx, __shfl_xor(mx, o));
            }
            const int c = 8 * t + (lane >> 3);
            if ((lane & 7) == 0 && c < pc && mn <= mx) {
                const int64_t qid = s_qid[c];
                atomicMin(&la.qrange[qid], key_bits(mn == 0.f ? -0.f : mn));
                atomicMax(&la.qrange[la.nq + qid], key_bits(mx == 0.f ? 0.f : mx));
            }
        }
        __syncthreads();   // the next item rewrites the query tile
        item = next_item;
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
__global__ void __launch_bounds__(256) k_select_rows(ScanArgs a, int64_t nq, const float *__restrict__ dist,
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
    int64_t *bpos = reinterpret_cast<int64_t *>(bkey + SEL_ROOM);   // [SEL_ROOM]
    int64_t *poff = bpos + SEL_ROOM;                            // [probe_cap + 1]
    int64_t *pst = poff + a.probe_cap + 1;                      // [probe_cap]
    __shared__ int s_out, s_fill, s_B, s_below, s_nb;
    __shared__ unsigned s_prefix;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t q = blockIdx.x;
    int kq = a.k_q ? a.k_q[q] : a.k_default;
    kq = kq < 1 ? 1 : (kq > a.k_default ? a.k_default : kq);
    const int selk = kq * a.rerank;
    const int64_t Lr = qtot[q], L = qslots[q];   // entries, row slots (a multiple of 4)
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
    const float *row = dist + row_off[q];
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
    // fn(slot, key bits) over the row (16-byte aligned, a multiple of 4
    // slots): four float4 loads in flight per thread
    constexpr int SU = 4;
    const float4 *row4 = reinterpret_cast<const float4 *>(row);
    const int64_t L4 = L / 4;
    auto for_row = [&](auto &&fn) {
        for (int64_t j0 = 0; j0 < L4; j0 += 256 * SU) {
            float4 v[SU];
#pragma unroll
            for (int u = 0; u < SU; ++u) {
                const int64_t j = j0 + u * 256 + tid;
                v[u] = row4[j < L4 ? j : L4 - 1];
            }
#pragma unroll
            for (int u = 0; u < SU; ++u) {
                const int64_t j = j0 + u * 256 + tid;
                if (j < L4) {
                    fn(4 * j + 0, key_bits(v[u].x));
                    fn(4 * j + 1, key_bits(v[u].y));
                    fn(4 * j + 2, key_bits(v[u].z));
                    fn(4 * j + 3, key_bits(v[u].w));
                }
            }
        }
    };
    unsigned lo = qrange[q], hi = qrange[nq + q];
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
        // the bin where the cumulative count reaches need (wave 0: 32 bins a lane)
        if (wave == 0) {
            unsigned own = 0;
            for (int b = 0; b < SEL_BINS / 64; ++b) own += hist[lane * (SEL_BINS / 64) + b];
            unsigned incl = own;
            for (int o = 1; o < 64; o <<= 1) {
                const unsigned v = __shfl_up(incl, o);
                if (lane >= o) incl += v;
            }
            const unsigned excl = incl - own;
            if (excl < (unsigned)need && (unsigned)need <= incl) {
                unsigned run = excl;
                int b = lane * (SEL_BINS / 64);
                while (run + hist[b] < (unsigned)need) run += hist[b++];
                s_B = b;
                s_below = (int)run;
                s_nb = (int)hist[b];
            }
        }
        __syncthreads();
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
                    bpos[slot] = pos_of(i);
                }
            });
        }
        need -= below;
        const unsigned blo = lo + ((unsigned)B << sh);
        const unsigned bspan = (sh >= 32) ? 0xFFFFFFFFu : ((1u << sh) - 1u);
        const unsigned bhi = (hi - blo) < bspan ? hi : blo + bspan;
        if (fits) {
            // sort bin B by (key bits, position); the first need complete the row
            __syncthreads();
            int span = 1;
            while (span < nb) span <<= 1;
            for (int t = nb + tid; t < span; t += 256) {
                bkey[t] = 0xFFFFFFFFu;
                bpos[t] = INT64_MAX;
            }
            __syncthreads();
            for (int size = 2; size <= span; size <<= 1) {
                for (int stride = size >> 1; stride > 0; stride >>= 1) {
                    for (int t = tid; t < (span >> 1); t += 256) {
                        const int i0 = 2 * t - (t & (stride - 1));
                        const int i1 = i0 + stride;
                        const bool up = (i0 & size) == 0;
                        const unsigned ka = bkey[i0], kb = bkey[i1];
                        const int64_t pa = bpos[i0], pb = bpos[i1];
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

void launch_pair_offsets(const int32_t *lcnt, int K, int32_t *lstart, int32_t *ioff, int32_t *lfill,
                         int32_t *totals, hipStream_t st) {
    hipLaunchKernelGGL(dev::k_pair_offsets, dim3(1), dim3(1024), 0, st, lcnt, K, lstart, ioff, lfill, totals);
    HANN_LAUNCH_CHECK();
}

void launch_pair_scatter(const ScanArgs &a, int64_t nq, const double *probe_rank, const int32_t *lstart,
                         const int64_t *row_off, int32_t *lfill, int32_t *pair_q, float *pair_qc, int64_t *pair_off,
                         hipStream_t st) {
    if (nq == 0) return;
    hipLaunchKernelGGL(dev::k_pair_scatter, dim3((unsigned)nq), dim3(64), 0, st, a, probe_rank, lstart, row_off, lfill,
                       pair_q, pair_qc, pair_off);
    HANN_LAUNCH_CHECK();
}

void launch_make_items(const int32_t *lcnt, const int32_t *lstart, const int32_t *ioff, int K, int4 *items,
                       hipStream_t st) {
    hipLaunchKernelGGL(dev::k_make_items, dim3((unsigned)ceil_div(K, 256)), dim3(256), 0, st, lcnt, lstart, ioff, K,
                       items);
    HANN_LAUNCH_CHECK();
}

bool list_scan_supported(int d, int m, int P, int code_bytes) {
    if (code_bytes != 1 || d != 4 * m) return false;
    if (m != 8 && m != 16 && m != 24 && m != 32 && m != 48 && m != 64) return false;
    return list_scan_lds_bytes(1, d, m, P) <= 160 * 1024;
}

size_t list_scan_lds_bytes(int x3, int d, int m, int P) {
    const int nb = x3 ? 2 : 1, ld = d + 8;
    return (size_t)nb * m * P * 8 + (size_t)nb * dev::LQ * ld * 2 + (size_t)dev::LNW * 32 * 4 +
           (size_t)dev::LNW * 32 * dev::LT * 4;
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

void launch_select_rows(const ScanArgs &a, int64_t nq, const float *dist, const int64_t *row_off,
                        const int64_t *qtot, const int64_t *qslots, const unsigned *qrange, int64_t *out_pos,
                        int32_t *out_cnt, int out_stride, int32_t *ovf, int32_t *ovf_cnt, hipStream_t st) {
    if (nq == 0) return;
    if (a.probe_cap > dev::SEL_MAXP) fail(HANN_ERR_UNSUPPORTED, "list-major select: too many probes per query");
    const size_t shm = (size_t)dev::SEL_BINS * 4 + (size_t)dev::SEL_ROOM * 12 + (size_t)(2 * a.probe_cap + 1) * 8;
    HANN_HIP(hipFuncSetAttribute((const void *)dev::k_select_rows, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)shm));
    hipLaunchKernelGGL(dev::k_select_rows, dim3((unsigned)nq), dim3(256), shm, st, a, nq, dist, row_off, qtot, qslots,
                       qrange, out_pos, out_cnt, out_stride, ovf, ovf_cnt);
    HANN_LAUNCH_CHECK();
}

}  // namespace hann
