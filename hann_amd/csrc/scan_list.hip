// scan_list.hip -- list-major scan: decode each probed list once per batch
// and score it against every query that probes it on the matrix cores.
//
// With the entry's reconstruction y_e = concat_i cb_i[code_e,i] (residual
// part) and the per-entry list term t_e,
//   |q - c_l - y_e|^2 = |q - c_l|^2 + t_e - 2 <q, y_e>
// and for a list l probed by queries Q_l the inner products are one product
// of the decoded list [n_l x d] with Q_l^T [d x |Q_l|]. The pairs (query,
// probed list) of a batch are grouped by list; a work item is one list with
// up to 32 of its queries. A workgroup decodes 32 entries per wave into LDS
// (BF16, or BF16 hi + lo for the split mode) and runs
// v_mfma_f32_32x32x16_bf16 over d. Each list's codes are read once per item
// instead of once per (query, list).
//
// Candidates: every entry within the query's threshold tau_q: the selk-th
// table-scan distance over its nearest lists (a pre-pass), widened by a
// bound on |MFMA - table| distance. At least selk entries pass, and the
// selk best by MFMA distance are among them (each is at most the selk-th
// passing distance); queries with more than the candidate room (or, as a
// guard, fewer than selk) are redone by the table scan. The selk best are
// then reranked in exact FP32.
#include "common.h"
#include "scan_list.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

namespace hann {
namespace dev {

constexpr int LQ = 32;       // queries per work item (the MFMA's N)
constexpr int LSLOTS = 32;   // per-query candidate slots of a work item in LDS

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

// one wave per query
__global__ void __launch_bounds__(64) k_pair_count(ScanArgs a, int32_t *__restrict__ lcnt,
                                                   int64_t *__restrict__ qtot) {
    const int64_t q = blockIdx.x;
    const int lane = threadIdx.x;
    const int npq = clamp_nprobe(a, q);
    int64_t tot = 0;
    for (int li = lane; li < npq; li += 64) {
        const int64_t l = a.probes[q * a.probe_stride + li];
        const int64_t len = a.list_len[l];
        tot += len;
        if (len > 0) atomicAdd(&lcnt[l], 1);
    }
    for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o);
    if (lane == 0) qtot[q] = tot;
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
    for (int l = b0; l < b1; ++l) {
        sa += lcnt[l];
        sb += (lcnt[l] + LQ - 1) / LQ;
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
    for (int l = b0; l < b1; ++l) {
        lstart[l] = ra;
        ioff[l] = rb;
        lfill[l] = 0;
        ra += lcnt[l];
        rb += (lcnt[l] + LQ - 1) / LQ;
    }
    if (t == 1023) {
        totals[0] = s_a[1023];
        totals[1] = s_b[1023];
    }
}

__global__ void __launch_bounds__(64) k_pair_scatter(ScanArgs a, const double *__restrict__ probe_rank,
                                                     const int32_t *__restrict__ lstart, int32_t *__restrict__ lfill,
                                                     int32_t *__restrict__ pair_q, float *__restrict__ pair_qc) {
    const int64_t q = blockIdx.x;
    const int lane = threadIdx.x;
    const int npq = clamp_nprobe(a, q);
    for (int li = lane; li < npq; li += 64) {
        const int64_t l = a.probes[q * a.probe_stride + li];
        if (a.list_len[l] == 0) continue;
        const int pos = lstart[l] + atomicAdd(&lfill[l], 1);
        pair_q[pos] = (int32_t)q;
        pair_qc[pos] = probe_rank ? (float)probe_rank[q * a.probe_stride + li]
                                  : sqdiff_exact(PtrLoad{a.queries + q * a.d}, PtrLoad{a.coarse + l * a.d}, a.d);
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

// one thread per query
__global__ void k_list_thresholds(ScanArgs a, int64_t nq, const float *__restrict__ pre_key,
                                  const int32_t *__restrict__ pre_cnt, int pre_stride, int pre_lists,
                                  const int64_t *__restrict__ qtot, const double *__restrict__ probe_rank,
                                  const float *__restrict__ ymax_tmax, float coef, float *__restrict__ tau) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    int kq = a.k_q ? a.k_q[q] : a.k_default;
    kq = kq < 1 ? 1 : (kq > a.k_default ? a.k_default : kq);
    const int selk = kq * a.rerank;
    const int npq = clamp_nprobe(a, q);
    if (pre_cnt[q] < selk || npq == 0) {
        tau[q] = __builtin_inff();
        return;
    }
    // the selk-th key of the nearest lists bounds the selk-th over all
    // probed lists (a superset): at least selk entries pass
    (void)pre_lists;
    (void)qtot;
    const double t0 = pre_key[q * pre_stride + selk - 1];
    double qq = 0.0;
    for (int j = 0; j < a.d; ++j) {
        const double v = a.queries[q * a.d + j];
        qq += v * v;
    }
    const double qn = __builtin_sqrt(qq) * (1.0 + 1e-6);
    const double ym = ymax_tmax[0], tm = ymax_tmax[1];
    const double qcmax = probe_rank[q * a.probe_stride + npq - 1];
    const double u = 1.0 / 16777216.0;
    const double bound = (double)coef * qn * ym + 32.0 * u * (qcmax + tm + 2.0 * qn * ym);
    tau[q] = (float)((t0 + bound) * (1.0 + 1.0 / 1048576.0) + 1e-30);
}

// LNW = 4 waves per workgroup; s == 4 (d = 4 M) and uint8 codes. The
// codebook sits in LDS as BF16 (hi, and lo for the split mode) for the
// kernel's lifetime; the MFMA's A operand of lane l for k-step j is the two
// codewords of subquantizers 4j + 2h and 4j + 2h + 1 (h = l / 32) of entry
// row l % 32, gathered straight from LDS into registers.
//
// Latency: a workgroup takes its next item from the queue while it starts
// the current one, and every thread prefetches its share of the next item
// (its query row slice, the pair's rank and threshold, and the first code
// slab) into registers; inside an item each wave loads its next slab's codes
// and terms while it computes the current one.
constexpr int LNW = 4;

template <bool X3, int M>
__global__ void __launch_bounds__(64 * LNW) k_scan_list(ListScanArgs la) {
    extern __shared__ __align__(16) unsigned char s_raw[];
    __shared__ int s_next, s_nitems;
    __shared__ int s_cnt[LQ];
    __shared__ float s_qc[LQ], s_tau[LQ];
    __shared__ int s_qid[LQ];
    constexpr int D = 4 * M, LD = D + 8;   // LD: Q row stride in BF16 (16-byte rows)
    constexpr int QV = M / 8;              // float4 per thread of a 32 x D query tile (8 threads a row)
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int P = la.P;
    bf16x4 *sCBh = reinterpret_cast<bf16x4 *>(s_raw);            // [M][P]
    bf16x4 *sCBl = sCBh + (size_t)M * P;                          // [M][P] (split mode)
    __bf16 *sQh = reinterpret_cast<__bf16 *>(sCBh + (X3 ? 2 : 1) * (size_t)M * P);
    __bf16 *sQl = sQh + LQ * LD;
    float *s_term = reinterpret_cast<float *>(sQh + (X3 ? 2 : 1) * LQ * LD) + wave * 32;
    float *s_ck = reinterpret_cast<float *>(sQh + (X3 ? 2 : 1) * LQ * LD) + LNW * 32;
    int32_t *s_cp = reinterpret_cast<int32_t *>(s_ck + LQ * LSLOTS);
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

    // prefetched state of the next item
    int4 nh = make_int4(0, 0, 0, 0);
    int nqid = -1;
    float nqc = 0.f, ntau = 0.f;
    float4 nqv[QV];
    int64_t ne0 = 0, nn = 0;
    uint32_t nw[M / 4];
    float nt = 0.f;
    auto load_codes = [&](int64_t e0, int64_t n, int64_t b, uint32_t *w, float &t) {
        const int64_t e = b + col;   // this lane's A row
        if (e < n) {
            const uint2 *cw = reinterpret_cast<const uint2 *>(codes + (e0 + e) * M);
#pragma unroll
            for (int j = 0; j < M / 8; ++j) {
                const uint2 v = cw[j];
                w[2 * j] = v.x;
                w[2 * j + 1] = v.y;
            }
        } else {
#pragma unroll
            for (int j = 0; j < M / 4; ++j) w[j] = 0;
        }
        t = (lane < 32 && b + lane < n) ? la.term[e0 + b + lane] : 0.f;
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
                ntau = la.tau[nqid];
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
                s_tau[qrow] = qrow < pc ? ntau : -__builtin_inff();
                s_cnt[qrow] = 0;
            }
        }
        uint32_t wn[M / 4];
#pragma unroll
        for (int j = 0; j < M / 4; ++j) wn[j] = nw[j];
        float tn = nt;
        if (tid == 0) s_next = atomicAdd(la.item_counter, 1);
        __syncthreads();
        const int next_item = s_next;
        prefetch(next_item);   // in flight during this item
        const float qc_r = s_qc[col], tau_r = s_tau[col];
        for (int64_t base = (int64_t)wave * 32; base < n; base += 32 * LNW) {
            uint32_t w[M / 4];
#pragma unroll
            for (int j = 0; j < M / 4; ++j) w[j] = wn[j];
            const float tc = tn;
            load_codes(e0, n, base + 32 * LNW, wn, tn);
            if (lane < 32) s_term[lane] = tc;
            f32x16 acc;
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[r] = 0.f;
#pragma unroll
            for (int j = 0; j < M / 4; ++j) {
                const int i0 = 4 * j + 2 * half;
                const uint32_t c0 = (w[j] >> (16 * half)) & 0xFFu, c1 = (w[j] >> (16 * half + 8)) & 0xFFu;
                const bf16x4 x0 = sCBh[i0 * P + c0], x1 = sCBh[(i0 + 1) * P + c1];
                const bf16x8 a = {x0[0], x0[1], x0[2], x0[3], x1[0], x1[1], x1[2], x1[3]};
                const bf16x8 b = *reinterpret_cast<const bf16x8 *>(&sQh[col * LD + 16 * j + 8 * half]);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
                if (X3) {
                    const bf16x4 y0 = sCBl[i0 * P + c0], y1 = sCBl[(i0 + 1) * P + c1];
                    const bf16x8 al = {y0[0], y0[1], y0[2], y0[3], y1[0], y1[1], y1[2], y1[3]};
                    const bf16x8 bl = *reinterpret_cast<const bf16x8 *>(&sQl[col * LD + 16 * j + 8 * half]);
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, b, acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bl, acc, 0, 0, 0);
                }
            }
            wave_sync();   // s_term
            // rows of D: entries; columns: the item's queries
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = 8 * (r >> 2) + 4 * half + (r & 3);
                const int64_t er = base + row;
                if (er < n && col < pc) {
                    const float dist = (qc_r + s_term[row]) - 2.f * acc[r];
                    if (dist <= tau_r) {
                        const int slot = atomicAdd(&s_cnt[col], 1);
                        if (slot < LSLOTS) {
                            s_ck[col * LSLOTS + slot] = dist;
                            s_cp[col * LSLOTS + slot] = (int32_t)(e0 + er);
                        } else {
                            const int64_t qid = s_qid[col];
                            const int g = atomicAdd(&la.cnt[qid], 1);
                            if (g < la.cap) {
                                la.cand_key[qid * la.cap + g] = dist;
                                la.cand_pos[qid * la.cap + g] = (int32_t)(e0 + er);
                            }
                        }
                    }
                }
            }
            wave_sync();   // s_term is rewritten by the next slab
        }
        __syncthreads();
        for (int t = wave; t < pc; t += LNW) {
            int ntk = s_cnt[t];
            ntk = ntk < LSLOTS ? ntk : LSLOTS;
            if (ntk == 0) continue;
            const int64_t qid = s_qid[t];
            int g0 = 0;
            if (lane == 0) g0 = atomicAdd(&la.cnt[qid], ntk);
            g0 = __shfl(g0, 0);
            for (int j = lane; j < ntk; j += 64)
                if (g0 + j < la.cap) {
                    la.cand_key[qid * la.cap + g0 + j] = s_ck[t * LSLOTS + j];
                    la.cand_pos[qid * la.cap + g0 + j] = s_cp[t * LSLOTS + j];
                }
        }
        __syncthreads();   // the next item rewrites the query tile and the slots
        item = next_item;
    }
}

__device__ __forceinline__ uint64_t cand_code(float key, int32_t pos) {
    const unsigned u = __float_as_uint(sane_key(key));
    const unsigned k = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    return ((uint64_t)k << 32) | (uint32_t)pos;
}

// one workgroup of 256 per query: the selk smallest (key, position) codes by
// an 8-bit radix select over the 64-bit code
__global__ void __launch_bounds__(256) k_select_candidates(const float *__restrict__ cand_key,
                                                           const int32_t *__restrict__ cand_pos,
                                                           const int32_t *__restrict__ cnt, int cap,
                                                           const int64_t *__restrict__ qtot,
                                                           const int32_t *__restrict__ k_q, int k_default, int R,
                                                           int allow_fallback, int64_t *__restrict__ out_pos,
                                                           int32_t *__restrict__ out_cnt, int out_stride,
                                                           int32_t *__restrict__ ovf, int32_t *__restrict__ ovf_cnt) {
    __shared__ unsigned hist[256];
    __shared__ uint64_t s_prefix;
    __shared__ int s_k, s_out;
    const int tid = threadIdx.x, lane = tid & 63;
    const int64_t q = blockIdx.x;
    int kq = k_q ? k_q[q] : k_default;
    kq = kq < 1 ? 1 : (kq > k_default ? k_default : kq);
    const int selk = kq * R;
    const int c = cnt[q];
    int64_t *orow = out_pos + q * out_stride;
    // overflow, too few candidates for the selection while the probed lists
    // hold enough, or the brute-force fallback: the table scan redoes it
    if (c > cap || (c < selk && qtot[q] > c) || (allow_fallback && qtot[q] < kq)) {
        if (tid == 0) ovf[atomicAdd(ovf_cnt, 1)] = (int32_t)q;
        return;
    }
    const float *ck = cand_key + q * cap;
    const int32_t *cp = cand_pos + q * cap;
    if (c <= selk) {
        for (int t = tid; t < out_stride; t += 256) orow[t] = t < c ? (int64_t)cp[t] : (int64_t)-1;
        if (tid == 0) out_cnt[q] = c;
        return;
    }
    uint64_t prefix = 0, mask = 0;
    int kk = selk;
    for (int pass = 0; pass < 8; ++pass) {
        const int shift = 56 - 8 * pass;
        hist[tid] = 0;
        __syncthreads();
        for (int i = tid; i < c; i += 256) {
            const uint64_t v = cand_code(ck[i], cp[i]);
            if ((v & mask) == prefix) atomicAdd(&hist[(unsigned)(v >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (tid < 64) {
            const unsigned h0 = hist[4 * lane], h1 = hist[4 * lane + 1], h2 = hist[4 * lane + 2],
                           h3 = hist[4 * lane + 3];
            const unsigned own = h0 + h1 + h2 + h3;
            unsigned incl = own;
            for (int o = 1; o < 64; o <<= 1) {
                const unsigned v = __shfl_up(incl, o);
                if (lane >= o) incl += v;
            }
            const unsigned excl = incl - own;
            if (excl < (unsigned)kk && (unsigned)kk <= incl) {
                unsigned run = excl, bin = 4 * lane;
                const unsigned hs[4] = {h0, h1, h2, h3};
                for (int j = 0; j < 4; ++j) {
                    if (run + hs[j] >= (unsigned)kk) {
                        bin = 4 * lane + j;
                        break;
                    }
                    run += hs[j];
                }
                s_prefix = prefix | ((uint64_t)bin << shift);
                s_k = kk - (int)run;
            }
        }
        __syncthreads();
        prefix = s_prefix;
        kk = s_k;
        mask |= (uint64_t)255 << shift;
        __syncthreads();
    }
    // the codes are distinct (positions are), so exactly selk are <= prefix
    if (tid == 0) s_out = 0;
    __syncthreads();
    for (int i = tid; i < c; i += 256)
        if (cand_code(ck[i], cp[i]) <= prefix) {
            const int slot = atomicAdd(&s_out, 1);
            if (slot < selk) orow[slot] = cp[i];
        }
    for (int t = selk + tid; t < out_stride; t += 256) orow[t] = -1;
    if (tid == 0) out_cnt[q] = selk;
}

}  // namespace dev

void launch_fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(dev::k_fill_i32, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, p, n, v);
    HANN_LAUNCH_CHECK();
}

void launch_pair_count(const ScanArgs &a, int64_t nq, int32_t *lcnt, int64_t *qtot, hipStream_t st) {
    if (nq == 0) return;
    hipLaunchKernelGGL(dev::k_pair_count, dim3((unsigned)nq), dim3(64), 0, st, a, lcnt, qtot);
    HANN_LAUNCH_CHECK();
}

void launch_pair_offsets(const int32_t *lcnt, int K, int32_t *lstart, int32_t *ioff, int32_t *lfill,
                         int32_t *totals, hipStream_t st) {
    hipLaunchKernelGGL(dev::k_pair_offsets, dim3(1), dim3(1024), 0, st, lcnt, K, lstart, ioff, lfill, totals);
    HANN_LAUNCH_CHECK();
}

void launch_pair_scatter(const ScanArgs &a, int64_t nq, const double *probe_rank, const int32_t *lstart,
                         int32_t *lfill, int32_t *pair_q, float *pair_qc, hipStream_t st) {
    if (nq == 0) return;
    hipLaunchKernelGGL(dev::k_pair_scatter, dim3((unsigned)nq), dim3(64), 0, st, a, probe_rank, lstart, lfill, pair_q,
                       pair_qc);
    HANN_LAUNCH_CHECK();
}

void launch_make_items(const int32_t *lcnt, const int32_t *lstart, const int32_t *ioff, int K, int4 *items,
                       hipStream_t st) {
    hipLaunchKernelGGL(dev::k_make_items, dim3((unsigned)ceil_div(K, 256)), dim3(256), 0, st, lcnt, lstart, ioff, K,
                       items);
    HANN_LAUNCH_CHECK();
}

void launch_list_thresholds(const ScanArgs &a, int64_t nq, const float *pre_key, const int32_t *pre_cnt,
                            int pre_stride, int pre_lists, const int64_t *qtot, const double *probe_rank,
                            const float *ymax_tmax, int x3, float *tau, hipStream_t st) {
    if (nq == 0) return;
    // 2 |<q, y>_MFMA - <q, y>| <= coef |q| |y|. With eps the BF16 rounding
    // (2^-9 relative to nearest, 2^-8 if truncated): products of rounded
    // operands err by 2 eps + eps^2, split operands (hi + lo, lo*lo dropped)
    // by about 3 eps^2; FP32 accumulation adds (d + 2) 2^-24. The constants
    // take the truncating eps and round up.
    const float coef = x3 ? 2.5e-4f : 1.6e-2f;
    hipLaunchKernelGGL(dev::k_list_thresholds, dim3((unsigned)ceil_div(nq, 256)), dim3(256), 0, st, a, nq, pre_key,
                       pre_cnt, pre_stride, pre_lists, qtot, probe_rank, ymax_tmax, coef, tau);
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
           (size_t)dev::LQ * dev::LSLOTS * 8;
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
    if (x3) HANN_LIST_M(true);
    else HANN_LIST_M(false);
#undef HANN_LIST_M
#undef HANN_LIST
}

void launch_select_candidates(const float *cand_key, const int32_t *cand_pos, const int32_t *cnt, int cap,
                              const int64_t *qtot, const int32_t *k_q, int k_default, int R, int allow_fallback,
                              int64_t nq, int64_t *out_pos, int32_t *out_cnt, int out_stride, int32_t *ovf,
                              int32_t *ovf_cnt, hipStream_t st) {
    if (nq == 0) return;
    hipLaunchKernelGGL(dev::k_select_candidates, dim3((unsigned)nq), dim3(256), 0, st, cand_key, cand_pos, cnt, cap,
                       qtot, k_q, k_default, R, allow_fallback, out_pos, out_cnt, out_stride, ovf, ovf_cnt);
    HANN_LAUNCH_CHECK();
}

}  // namespace hann
