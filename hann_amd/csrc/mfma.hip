// mfma.hip -- BF16 MFMA candidate pass for certified exact top-k (gfx950).
//
// The squared distance ||q||^2 + ||x||^2 - 2 q.x is evaluated on BF16 copies
// of the inputs with v_mfma_f32_32x32x16_bf16 (FP32 accumulation). Every row
// whose approximate distance is within the threshold T_q of a query becomes a
// candidate. T_q comes from an upper bound of the query's true k-th exact
// distance -- the k-th smallest approximate distance over a row sample,
// widened by the error bound below, or the exact rank of a hint row -- and
// is widened again by a rigorous bound on |approx - exact FP32|. The
// candidates are then ranked in the reference's exact FP32 order, so the
// result equals the exact path's.
#include "common.h"
#include "mfma.h"
#include "scan_list.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

namespace hann {
namespace dev {

// FP32 rows -> BF16 rows padded with zeros to dp columns; norm_b = ||x_b||^2
// of the BF16 values (FP32 sum), norm_f = ||x|| of the FP32 row; max_norm
// (float bits, non-negative) and inexact flag gathered with atomics.
// one wave per row, four rows per workgroup; at most one atomic each for
// the maximum and the flag per workgroup
constexpr int TB_ROWS = 4;

__global__ void __launch_bounds__(256) k_to_bf16(const float *__restrict__ x, int64_t n, int d, int dp,
                                                 __bf16 *__restrict__ out, float *__restrict__ norm_b,
                                                 float *__restrict__ norm_f, unsigned *__restrict__ max_norm,
                                                 int *__restrict__ inexact) {
    __shared__ float s_max[4];
    __shared__ int s_bad[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    float wmax = 0.f;
    int wbad = 0;
    for (int rr = w; rr < TB_ROWS; rr += 4) {
        const int64_t i = (int64_t)blockIdx.x * TB_ROWS + rr;
        if (i >= n) break;
        const float *r = x + i * d;
        __bf16 *o = out + i * dp;
        float sb = 0.f, sf = 0.f;
        int bad = 0;
        for (int e = lane; e < dp; e += 64) {
            const float v = e < d ? r[e] : 0.f;
            const __bf16 b = (__bf16)v;
            const float vb = (float)b;
            bad |= vb != v;
            o[e] = b;
            sb = __builtin_fmaf(vb, vb, sb);
            sf = __builtin_fmaf(v, v, sf);
        }
        for (int o2 = 32; o2 > 0; o2 >>= 1) {
            sb += __shfl_xor(sb, o2);
            sf += __shfl_xor(sf, o2);
            bad |= __shfl_xor(bad, o2);
        }
        const float nf = __builtin_sqrtf(sf);
        if (lane == 0) {
            norm_b[i] = sb;
            norm_f[i] = nf;
        }
        wmax = fmaxf(wmax, nf);
        wbad |= bad;
    }
    if (lane == 0) {
        s_max[w] = wmax;
        s_bad[w] = wbad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // atomics only when they change something (a stale read only costs
        // an atomic that changes nothing)
        const float m = fmaxf(fmaxf(s_max[0], s_max[1]), fmaxf(s_max[2], s_max[3]));
        if (__float_as_uint(m) > *(volatile unsigned *)max_norm) atomicMax(max_norm, __float_as_uint(m));
        if ((s_bad[0] | s_bad[1] | s_bad[2] | s_bad[3]) && *(volatile int *)inexact == 0) atomicOr(inexact, 1);
    }
}

// Per-query filter threshold from the sample's exact k-th rank tau.
//   e <= tau (exact FP32, reference order) implies D <= tau' = tau/(1-(d+2)u)
//   BF16 rounding moves sqrt(D) by at most delta = 2^-9 (|q| + max|x|)
//   (zero when every input is a BF16 value), FP32 accumulation of the
//   expansion adds at most (d+2) u (|q|+max|x|)^2 (1+2^-8)^2 for any
//   summation order; E2 takes twice that, since the MFMA unit's internal
//   accumulation order is not specified
//   so approx <= (sqrt(tau') + delta)^2 + E2 =: T, widened by 2^-20.
__global__ void k_thresholds(const double *__restrict__ tau_rank, int k, int nq, const float *__restrict__ qnorm_f,
                             const unsigned *__restrict__ max_norm, const int *__restrict__ inexact, int d,
                             float *__restrict__ thr) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const double t = tau_rank[(int64_t)q * k + (k - 1)];
    if (!(t < 1e300)) {
        thr[q] = __builtin_inff();
        return;
    }
    const double u = 1.0 / 16777216.0;
    const double a = qnorm_f[q], b = (double)__uint_as_float(*max_norm);
    const double tp = t / (1.0 - (d + 2) * u);
    const double delta = *inexact ? (a + b) / 512.0 : 0.0;
    const double e2 = 2.0 * (d + 2) * u * (a + b) * (a + b) * (1.0 + 1.0 / 256.0) * (1.0 + 1.0 / 256.0);
    const double s = __builtin_sqrt(tp) + delta;
    thr[q] = (float)((s * s + e2) * (1.0 + 1.0 / 1048576.0) + 1e-30);
}

// tau_rank[q] = exact reference-order rank of (query q, row hint[q])
__global__ void k_hint_rank(const float *__restrict__ queries, int64_t nq, int d, const float *__restrict__ base,
                            const int32_t *__restrict__ hint, double *__restrict__ tau_rank) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    tau_rank[q] = (double)sane_key(
        sqdiff_exact(PtrLoad{queries + q * d}, PtrLoad{base + (int64_t)hint[q] * d}, d));
}

constexpr int MF_TILE = 128;   // queries and rows per workgroup tile
constexpr int MF_KC = 128;     // K chunk held in LDS (bf16 columns)
constexpr int MF_PAD = 8;      // LDS row padding (bf16) against bank conflicts
constexpr int MF_SLOTS = 32;   // per-query candidate slots of a workgroup
constexpr int MF_OVF = 1024;   // a workgroup's shared slots for queries whose own are full

// One workgroup: 128 queries x (rt_per_wg x 128) rows, 4 waves of 64 x 64
// (2 x 2 tiles of 32 x 32) per row tile. Lane layout of
// v_mfma_f32_32x32x16_bf16: A lane l holds row l%32, k = 8*(l/32) .. +7; B
// lane l holds column l%32, the same k; D register r of lane l is row
// 8*(r/4) + 4*(l/32) + r%4, column l%32.
// DENSE: write every approximate distance to dense[q * ld_dense + row].
// Otherwise rows within thr[q] are gathered per query in LDS and appended
// to the query's global list once per workgroup; rows beyond the LDS slots
// are appended one by one. cnt[q] > cap (more candidates than room) sends
// the query to the exact path.
// the ballot bits of this lane's half of the wave
__device__ __forceinline__ unsigned half_mask(int lane, unsigned mlo, unsigned mhi) { return lane < 32 ? mlo : mhi; }

template <bool DENSE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_mfma_pass(const __bf16 *__restrict__ Q, int64_t nq,
                                                   const __bf16 *__restrict__ X, int64_t nb, int dp,
                                                   const float *__restrict__ qn, const float *__restrict__ xn,
                                                   const float *__restrict__ thr, int rt_per_wg,
                                                   int32_t *__restrict__ cnt, int32_t *__restrict__ cand_row, int cap,
                                                   float *__restrict__ dense, int64_t ld_dense) {
    extern __shared__ __align__(16) unsigned char s_raw[];
    const int kc_max = dp < MF_KC ? dp : MF_KC;
    const int ld = kc_max + MF_PAD;
    __bf16 *sA = reinterpret_cast<__bf16 *>(s_raw);
    __bf16 *sB = sA + MF_TILE * ld;
    float *s_qn = reinterpret_cast<float *>(sB + MF_TILE * ld);   // the tile's query norms
    float *s_thr = s_qn + MF_TILE;                                 // and thresholds (-inf past nq)
    // candidates: each (query, row half) has MF_SLOTS / 2 slots, written by
    // the one wave that tests that row half for that query
    int32_t *s_cc = reinterpret_cast<int32_t *>(s_thr + MF_TILE);   // [MF_TILE][2] counts
    int32_t *s_cr = s_cc + 2 * MF_TILE;                              // [MF_TILE][2][MF_SLOTS / 2]
    int2 *s_ov = reinterpret_cast<int2 *>(s_cr + MF_TILE * MF_SLOTS);   // [MF_OVF] (query, row)
    int32_t *s_on = reinterpret_cast<int32_t *>(s_ov + MF_OVF);
    const int tid = threadIdx.x;
    const int wave = tid / 64, lane = tid % 64;
    const int wm = wave / 2, wn = wave % 2;
    const int64_t q0 = (int64_t)blockIdx.y * MF_TILE;
    const int64_t rbase = (int64_t)blockIdx.x * MF_TILE * rt_per_wg;
    for (int t = tid; t < MF_TILE; t += 256) {
        const int64_t gq = q0 + t;
        s_qn[t] = gq < nq ? qn[gq] : 0.f;
        if (!DENSE) s_thr[t] = gq < nq ? thr[gq] : -__builtin_inff();
    }
    if (!DENSE && tid == 0) *s_on = 0;
    // 128 rows (from row r0, rows past lim zero) x columns k0 .. k0 + kc of
    // src into an LDS tile
    auto load_tile = [&](__bf16 *dst, const __bf16 *src, int64_t r0, int64_t lim, int k0, int kc) {
        const int vecs = kc / 8;
        for (int v = tid; v < MF_TILE * vecs; v += 256) {
            const int row = v / vecs, c8 = v % vecs;
            const int64_t g = r0 + row;
            bf16x8 a = {};
            if (g < lim) a = *reinterpret_cast<const bf16x8 *>(src + g * dp + k0 + c8 * 8);
            *reinterpret_cast<bf16x8 *>(&dst[row * ld + c8 * 8]) = a;
        }
    };
    // with one K chunk the query tile stays in LDS for every row tile
    const bool a_resident = dp <= MF_KC;
    if (a_resident) load_tile(sA, Q, q0, nq, 0, dp);
    // the running candidate count of each of the lane half's 32 queries in
    // this wave's row half (uniform over the half)
    constexpr int SL = MF_SLOTS / 2;
    int cq[2][4][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int g = 0; g < 4; ++g)
#pragma unroll
            for (int u = 0; u < 4; ++u) cq[i][g][u] = 0;

    for (int rt = 0; rt < rt_per_wg; ++rt) {
        const int64_t r0 = rbase + (int64_t)rt * MF_TILE;
        if (r0 >= nb) break;
        f32x16 acc[2][2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

        for (int k0 = 0; k0 < dp; k0 += MF_KC) {
            const int kc = dp - k0 < MF_KC ? dp - k0 : MF_KC;   // multiple of 16
            __syncthreads();   // the previous tile's readers are done
            if (!a_resident) load_tile(sA, Q, q0, nq, k0, kc);
            load_tile(sB, X, r0, nb, k0, kc);
            __syncthreads();
            const int kl = (lane / 32) * 8;
            for (int kk = 0; kk < kc; kk += 16) {
                bf16x8 a[2], b[2];
#pragma unroll
                for (int i = 0; i < 2; ++i)
                    a[i] = *reinterpret_cast<const bf16x8 *>(&sA[(wm * 64 + i * 32 + lane % 32) * ld + kk + kl]);
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    b[j] = *reinterpret_cast<const bf16x8 *>(&sB[(wn * 64 + j * 32 + lane % 32) * ld + kk + kl]);
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
            }
        }
        // epilogue: approximate distance; threshold test or dense store.
        // Registers 4g .. 4g + 3 of a tile are four consecutive queries; a
        // group's norms and thresholds are read once for both row tiles.
        int64_t grj[2];
        bool rvj[2];
        float xrj[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            grj[j] = r0 + wn * 64 + j * 32 + lane % 32;
            rvj[j] = grj[j] < nb;
            xrj[j] = rvj[j] ? xn[grj[j]] : 0.f;
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int qb = wm * 64 + i * 32 + 8 * g + 4 * (lane / 32);
                const float4 qn4 = *reinterpret_cast<const float4 *>(&s_qn[qb]);
                const float qv[4] = {qn4.x, qn4.y, qn4.z, qn4.w};
                float tv[4] = {0.f, 0.f, 0.f, 0.f};
                if (!DENSE) {
                    const float4 t4 = *reinterpret_cast<const float4 *>(&s_thr[qb]);
                    tv[0] = t4.x;
                    tv[1] = t4.y;
                    tv[2] = t4.z;
                    tv[3] = t4.w;
                }
#pragma unroll
                for (int j = 0; j < 2; ++j)
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int ql = qb + u;
                        const float approx = (qv[u] + xrj[j]) - 2.f * acc[i][j][4 * g + u];
                        if (DENSE) {
                            if (rvj[j] && q0 + ql < nq) dense[(q0 + ql) * ld_dense + grj[j]] = approx;
                            continue;
                        }
                        // thresholds past nq are -inf: only real queries pass.
                        // Lanes 0-31 and 32-63 each test one query against 32
                        // rows; a passing row's slot is the half's running
                        // count plus its rank in the half's ballot.
                        const bool pass = rvj[j] && approx <= tv[u];
                        const unsigned long long m = __ballot(pass);
                        if (m == 0) continue;
                        const unsigned mlo = (unsigned)m, mhi = (unsigned)(m >> 32);
                        if (pass) {
                            const int rank = lane < 32 ? (int)__builtin_amdgcn_mbcnt_lo(mlo, 0u)
                                                       : (int)__builtin_amdgcn_mbcnt_hi(mhi, 0u);
                            const int slot = cq[i][g][u] + rank;
                            if (slot < SL) {
                                s_cr[(ql * 2 + wn) * SL + slot] = (int32_t)grj[j];
                            } else {
                                // the query's slots are full: the
                                // workgroup's shared overflow list
                                const int o = atomicAdd(s_on, 1);
                                if (o < MF_OVF) s_ov[o] = make_int2(ql, (int32_t)grj[j]);
                            }
                        }
                        cq[i][g][u] += __popc(half_mask(lane, mlo, mhi));
                    }
            }
    }
    if (DENSE) return;
    // the counts, from the first lane of each half
    if ((lane & 31) == 0) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int g = 0; g < 4; ++g)
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    s_cc[(wm * 64 + i * 32 + 8 * g + 4 * (lane / 32) + u) * 2 + wn] = cq[i][g][u];
    }
    __syncthreads();
    for (int ql = tid; ql < MF_TILE; ql += 256) {
        const int64_t gq = q0 + ql;
        const int n0 = s_cc[ql * 2] < SL ? s_cc[ql * 2] : SL;
        const int n1 = s_cc[ql * 2 + 1] < SL ? s_cc[ql * 2 + 1] : SL;
        if (gq >= nq || n0 + n1 == 0) continue;
        const int base = atomicAdd(&cnt[gq], n0 + n1);
        for (int i = 0; i < n0 && base + i < cap; ++i) cand_row[gq * cap + base + i] = s_cr[(ql * 2) * SL + i];
        for (int i = 0; i < n1 && base + n0 + i < cap; ++i)
            cand_row[gq * cap + base + n0 + i] = s_cr[(ql * 2 + 1) * SL + i];
    }
    const int no = *s_on;
    for (int o = tid; o < (no < MF_OVF ? no : MF_OVF); o += 256) {
        const int2 e = s_ov[o];
        const int64_t gq = q0 + e.x;
        const int g = atomicAdd(&cnt[gq], 1);
        if (g < cap) cand_row[gq * cap + g] = e.y;
    }
    if (no > MF_OVF)
        // rows were lost: every query that overflowed its slots takes the
        // exact path (a count above the room)
        for (int ql = tid; ql < MF_TILE; ql += 256)
            if ((s_cc[ql * 2] > SL || s_cc[ql * 2 + 1] > SL) && q0 + ql < nq) atomicMax(&cnt[q0 + ql], cap + 1);
}

// The 16-bit key of an approximate distance: the top 16 bits of its FP32
// value, a non-decreasing function of it (negative values key 0, NaN the key
// of +inf). approx_key_upper(K) is the largest float whose key is K.
__device__ __forceinline__ uint16_t approx_key16(float a) {
    if (a != a) return 0x7F80u;
    return a > 0.f ? (uint16_t)(__float_as_uint(a) >> 16) : (uint16_t)0u;
}
__device__ __forceinline__ float approx_key_upper(unsigned key) {
    return key >= 0x7F80u ? __builtin_inff() : __uint_as_float((key << 16) | 0xFFFFu);
}

// The same pass for dp = 16 KS <= 128, with each wave's query fragments in
// registers and the next row tile's rows fetched into registers while the
// current tile is multiplied and tested (LDS holds one row tile, so two
// workgroups share a CU). Grid: x = query tiles, y = row groups, so the
// workgroups in flight share their row tiles in L2. Each candidate is
// written with its approximate distance (cand_ap), from which a tighter
// threshold follows; a (query, row half) has MF_SLOTS_R / 2 LDS slots.
constexpr int MF_SLOTS_R = 16;
constexpr int MF_OVF_R = 512;
template <bool DENSE, int KS, bool KEY16 = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_mfma_pass_r(const __bf16 *__restrict__ Q, int64_t nq, const __bf16 *__restrict__ X, int64_t nb,
              const float *__restrict__ qn, const float *__restrict__ xn, const float *__restrict__ thr, int rt_per_wg,
              int32_t *__restrict__ cnt, int32_t *__restrict__ cand_row, float *__restrict__ cand_ap, int cap,
              float *__restrict__ dense, int64_t ld_dense, int64_t rg0) {
    constexpr int DP = KS * 16;
    constexpr int LD = DP + MF_PAD;
    constexpr int VECS = DP / 8;   // 16-byte vectors per row
    constexpr int NV = MF_TILE * VECS / 256;
    static_assert(NV * 256 == MF_TILE * VECS, "whole vectors per thread");
    extern __shared__ __align__(16) unsigned char s_raw[];
    __bf16 *sB = reinterpret_cast<__bf16 *>(s_raw);
    float *s_xn = reinterpret_cast<float *>(sB + MF_TILE * LD);   // the row tile's norms
    float *s_qn = s_xn + MF_TILE;                                  // the query tile's norms
    float *s_thr = s_qn + MF_TILE;                                 // and thresholds (-inf past nq)
    int32_t *s_cc = reinterpret_cast<int32_t *>(s_thr + MF_TILE);  // [MF_TILE][2] counts
    int2 *s_cr = reinterpret_cast<int2 *>(s_cc + 2 * MF_TILE);      // [MF_TILE][2][MF_SLOTS_R / 2] (row, approx)
    int4 *s_ov = reinterpret_cast<int4 *>(s_cr + MF_TILE * MF_SLOTS_R);   // [MF_OVF_R] (query, row, approx)
    int32_t *s_on = reinterpret_cast<int32_t *>(s_ov + MF_OVF_R);
    const int tid = threadIdx.x;
    const int wave = tid / 64, lane = tid % 64;
    const int wm = wave / 2, wn = wave % 2;
    const int kl = (lane / 32) * 8;
    const int64_t q0 = (int64_t)blockIdx.x * MF_TILE;
    const int64_t rbase = ((int64_t)blockIdx.y + rg0) * MF_TILE * rt_per_wg;
    for (int t = tid; t < MF_TILE; t += 256) {
        const int64_t gq = q0 + t;
        s_qn[t] = gq < nq ? qn[gq] : 0.f;
        if (!DENSE) s_thr[t] = gq < nq ? thr[gq] : -__builtin_inff();
    }
    if (!DENSE && tid == 0) *s_on = 0;
    // this wave's query fragments: A lane l holds query row l % 32 of each
    // 32-row tile, columns 16 ks + kl .. + 7
    bf16x8 a[2][KS];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int64_t gq = q0 + wm * 64 + i * 32 + lane % 32;
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            bf16x8 v = {};
            if (gq < nq) v = *reinterpret_cast<const bf16x8 *>(Q + gq * DP + ks * 16 + kl);
            a[i][ks] = v;
        }
    }
    const int64_t left = nb - rbase;   // > 0: the grid covers nb
    const int ntiles = (int)(left < (int64_t)rt_per_wg * MF_TILE ? (left + MF_TILE - 1) / MF_TILE : rt_per_wg);
    // vector v = tid + 256 n of a tile is row v / VECS, columns 8 (v % VECS)
    // .. + 7: a tile's rows are contiguous in X
    bf16x8 breg[NV];
    float xreg = 0.f;
    auto fetch = [&](int64_t r0) {
#pragma unroll
        for (int n = 0; n < NV; ++n) {
            const int v = tid + 256 * n;
            bf16x8 b = {};
            if (r0 + v / VECS < nb) b = *reinterpret_cast<const bf16x8 *>(X + r0 * DP + (int64_t)v * 8);
            breg[n] = b;
        }
        if (tid < MF_TILE) xreg = r0 + tid < nb ? xn[r0 + tid] : 0.f;
    };
    fetch(rbase);
    // KEY16: this lane's least and largest key of each query tile i (its
    // query's valid rows), folded into the query's range at the end
    unsigned kmin[2] = {0xFFFFu, 0xFFFFu}, kmax[2] = {0u, 0u};
    constexpr int SL = MF_SLOTS_R / 2;
    // the running candidate counts of each query in this wave's row half,
    // wave-uniform: lanes 0-31's count in the low 16 bits, lanes 32-63's in
    // the high 16 (at most 64 rt_per_wg rows each)
    unsigned cq[2][4][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int g = 0; g < 4; ++g)
#pragma unroll
            for (int u = 0; u < 4; ++u) cq[i][g][u] = 0;

    for (int rt = 0; rt < ntiles; ++rt) {
        const int64_t r0 = rbase + (int64_t)rt * MF_TILE;
        __syncthreads();   // the previous tile's readers are done
#pragma unroll
        for (int n = 0; n < NV; ++n) {
            const int v = tid + 256 * n;
            *reinterpret_cast<bf16x8 *>(&sB[(v / VECS) * LD + (v % VECS) * 8]) = breg[n];
        }
        if (tid < MF_TILE) s_xn[tid] = xreg;
        __syncthreads();
        if (rt + 1 < ntiles) fetch(r0 + MF_TILE);
        // one 32-row half of the wave's 64 rows at a time: its fragments,
        // the products with both 32-query tiles, then their test
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int rl = wn * 64 + j * 32 + lane % 32;
            const int64_t gr = r0 + rl;
            const bool rv = gr < nb;
            // the lanes whose row exists (a wave-uniform mask, ANDed into each
            // compare's, so no per-lane boolean is kept per element)
            const unsigned long long rvm = __builtin_amdgcn_ballot_w64(rv);
            const float xr = s_xn[rl];
            f32x16 acc[2];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[i][r] = 0.f;
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                const bf16x8 b = *reinterpret_cast<const bf16x8 *>(&sB[rl * LD + ks * 16 + kl]);
#pragma unroll
                for (int i = 0; i < 2; ++i)
                    acc[i] = KEY16 ? __builtin_amdgcn_mfma_f32_32x32x16_bf16(b, a[i][ks], acc[i], 0, 0, 0)
                                   : __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][ks], b, acc[i], 0, 0, 0);
            }
            if (KEY16) {
                // operands swapped: register 4 g + u of lane l holds row
                // 8 g + 4 h + u of this 32-row half for query l % 32 of query
                // tile i, so a lane's four keys of a group are consecutive
                // rows of one query: one 8-byte store (ld_dense a multiple
                // of 4)
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    const int ql = wm * 64 + i * 32 + lane % 32;
                    const float qv = s_qn[ql];
                    const int64_t gq = q0 + ql;
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        const int rl0 = wn * 64 + j * 32 + 8 * g + 4 * (lane / 32);
                        const float4 x4 = *reinterpret_cast<const float4 *>(&s_xn[rl0]);
                        const float xs[4] = {x4.x, x4.y, x4.z, x4.w};
                        unsigned kk[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u) kk[u] = approx_key16((qv + xs[u]) - 2.f * acc[i][4 * g + u]);
                        const int64_t gr0 = r0 + rl0;
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                            if (gr0 + u < nb) {
                                kmin[i] = kk[u] < kmin[i] ? kk[u] : kmin[i];
                                kmax[i] = kk[u] > kmax[i] ? kk[u] : kmax[i];
                            }
                        if (gq < nq) {
                            uint16_t *dst = reinterpret_cast<uint16_t *>(dense) + gq * ld_dense + gr0;
                            if (gr0 + 3 < nb) {
                                *reinterpret_cast<uint2 *>(dst) = make_uint2(kk[0] | (kk[1] << 16), kk[2] | (kk[3] << 16));
                            } else {
#pragma unroll
                                for (int u = 0; u < 4; ++u)
                                    if (gr0 + u < nb) dst[u] = (uint16_t)kk[u];
                            }
                        }
                    }
                }
                continue;
            }
            // the epilogue of k_mfma_pass: approximate distance, then the
            // threshold test (ballot, slot by rank in the lane half) or the
            // dense store
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const int qb = wm * 64 + i * 32 + 8 * g + 4 * (lane / 32);
                    const float4 qn4 = *reinterpret_cast<const float4 *>(&s_qn[qb]);
                    const float qv[4] = {qn4.x, qn4.y, qn4.z, qn4.w};
                    float tv[4] = {0.f, 0.f, 0.f, 0.f};
                    if (!DENSE) {
                        const float4 t4 = *reinterpret_cast<const float4 *>(&s_thr[qb]);
                        tv[0] = t4.x;
                        tv[1] = t4.y;
                        tv[2] = t4.z;
                        tv[3] = t4.w;
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const float approx = (qv[u] + xr) - 2.f * acc[i][4 * g + u];
                        if (DENSE) {
                            int qbo = qb;
                            asm volatile("" : "+v"(qbo));
                            const int ql = qbo + u;
                            if (rv && q0 + ql < nq) dense[(q0 + ql) * ld_dense + gr] = approx;
                            continue;
                        }
                        // the compare goes straight into a lane mask
                        const unsigned long long m = __builtin_amdgcn_ballot_w64(approx <= tv[u]) & rvm;
                        if (m == 0) continue;
                        // the rare path: the query index is computed here
                        // from an opaque copy of the lane's query base, so
                        // that the 32 per-element indices and addresses are
                        // not kept in registers across the loop
                        int qbo = wm * 64 + 4 * (lane / 32);
                        asm volatile("" : "+v"(qbo));
                        const int ql = qbo + i * 32 + 8 * g + u;
                        const unsigned mlo = (unsigned)m, mhi = (unsigned)(m >> 32);
                        if (__builtin_amdgcn_inverse_ballot_w64(m)) {
                            // rank among the passing lanes of the lane's half:
                            // mbcnt over both words counts all of the low
                            // half's for an upper lane, so subtract them
                            const unsigned below = __builtin_amdgcn_mbcnt_hi(mhi, __builtin_amdgcn_mbcnt_lo(mlo, 0u));
                            const int rank = (int)(lane < 32 ? below : below - (unsigned)__popc(mlo));
                            const int slot = (int)(lane < 32 ? cq[i][g][u] & 0xffffu : cq[i][g][u] >> 16) + rank;
                            if (slot < SL) {
                                s_cr[(ql * 2 + wn) * SL + slot] = make_int2((int32_t)gr, __float_as_int(approx));
                            } else {
                                const int o = atomicAdd(s_on, 1);
                                if (o < MF_OVF_R) s_ov[o] = make_int4(ql, (int32_t)gr, __float_as_int(approx), 0);
                            }
                        }
                        cq[i][g][u] += (unsigned)__popc(mlo) + ((unsigned)__popc(mhi) << 16);
                    }
                }
        }
    }
    if (KEY16) {
        // each query's key range: lanes l and l + 32 hold the same query;
        // cnt[q] takes the least key, cand_row[q] the largest
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const unsigned mn = __shfl_xor(kmin[i], 32), mx = __shfl_xor(kmax[i], 32);
            const unsigned lo = mn < kmin[i] ? mn : kmin[i], hi = mx > kmax[i] ? mx : kmax[i];
            const int64_t gq = q0 + wm * 64 + i * 32 + lane % 32;
            if (lane < 32 && gq < nq) {
                atomicMin(reinterpret_cast<unsigned *>(cnt) + gq, lo);
                atomicMax(reinterpret_cast<unsigned *>(cand_row) + gq, hi);
            }
        }
        return;
    }
    if (DENSE) return;
    if ((lane & 31) == 0) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int g = 0; g < 4; ++g)
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    s_cc[(wm * 64 + i * 32 + 8 * g + 4 * (lane / 32) + u) * 2 + wn] =
                        (int)(lane < 32 ? cq[i][g][u] & 0xffffu : cq[i][g][u] >> 16);
    }
    __syncthreads();
    for (int ql = tid; ql < MF_TILE; ql += 256) {
        const int64_t gq = q0 + ql;
        const int n0 = s_cc[ql * 2] < SL ? s_cc[ql * 2] : SL;
        const int n1 = s_cc[ql * 2 + 1] < SL ? s_cc[ql * 2 + 1] : SL;
        if (gq >= nq || n0 + n1 == 0) continue;
        const int base = atomicAdd(&cnt[gq], n0 + n1);
        for (int i = 0; i < n0 + n1 && base + i < cap; ++i) {
            const int2 e = s_cr[(ql * 2 + (i < n0 ? 0 : 1)) * SL + (i < n0 ? i : i - n0)];
            cand_row[gq * cap + base + i] = e.x;
            cand_ap[gq * cap + base + i] = __int_as_float(e.y);
        }
    }
    const int no = *s_on;
    for (int o = tid; o < (no < MF_OVF_R ? no : MF_OVF_R); o += 256) {
        const int4 e = s_ov[o];
        const int64_t gq = q0 + e.x;
        if (gq >= nq) continue;
        const int g = atomicAdd(&cnt[gq], 1);
        if (g < cap) {
            cand_row[gq * cap + g] = e.y;
            cand_ap[gq * cap + g] = __int_as_float(e.z);
        }
    }
    if (no > MF_OVF_R)
        for (int ql = tid; ql < MF_TILE; ql += 256)
            if ((s_cc[ql * 2] > SL || s_cc[ql * 2 + 1] > SL) && q0 + ql < nq) atomicMax(&cnt[q0 + ql], cap + 1);
}

// rows 0, stride, 2 stride, ... of X (BF16, dp columns) with their BF16 norms
__global__ void k_gather_bf16(const __bf16 *__restrict__ X, int dp, const float *__restrict__ xn, int64_t stride,
                              int64_t n, __bf16 *__restrict__ out, float *__restrict__ out_n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * dp) return;
    const int64_t b = t / dp;
    const int c = (int)(t % dp);
    const int64_t src = b * stride;
    out[t] = X[src * dp + c];
    if (c == 0) out_n[b] = xn[src];
}

// order-preserving map of floats (after NaN -> +inf) onto unsigned integers
__device__ __forceinline__ unsigned float_key(float v) {
    const unsigned u = __float_as_uint(sane_key(v));
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_float(unsigned k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

// The filter threshold from A >= the k-th smallest approximate distance of a
// query (qa = its FP32 norm, qb = the base's largest, inexact: some input is
// not a BF16 value): each of the k rows up to the k-th has
//   D <= (sqrt(max(A, 0) + E2) + delta)^2  and exact rank e <= D / (1-(d+2)u)
// (the bounds of k_thresholds, read the other way), which bounds the query's
// true k-th exact rank tau; the threshold is formed from tau as in
// k_thresholds. A = +inf (or NaN) gives +inf.
__device__ __forceinline__ float threshold_from_kth(float A, double qa, double qb, int inexact, int d) {
    if (!(A < __builtin_inff())) return __builtin_inff();
    const double u = 1.0 / 16777216.0;
    const double delta = inexact ? (qa + qb) / 512.0 : 0.0;
    const double e2 = 2.0 * (d + 2) * u * (qa + qb) * (qa + qb) * (1.0 + 1.0 / 256.0) * (1.0 + 1.0 / 256.0);
    const double s0 = __builtin_sqrt((A > 0.f ? (double)A : 0.0) + e2) + delta;
    const double tau = s0 * s0 / (1.0 - (d + 2) * u) * (1.0 + 1.0 / 1048576.0);
    const double tp = tau / (1.0 - (d + 2) * u);
    const double s1 = __builtin_sqrt(tp) + delta;
    return (float)((s1 * s1 + e2) * (1.0 + 1.0 / 1048576.0) + 1e-30);
}

// Per query: A' >= the k-th smallest approximate distance over its ns
// sample rows -- the upper edge of the bin holding the k-th in a histogram
// of ST_BINS bins over the query's [min, max] (in the order-preserving
// unsigned map of the floats). Each of the k rows up to the k-th has
//   D <= (sqrt(max(A', 0) + E2) + delta)^2  and exact rank e <= D / (1-(d+2)u)
// (the bounds of k_thresholds, read the other way), which bounds the query's
// true k-th exact rank tau; the filter threshold is then formed from tau as
// in k_thresholds. One workgroup of 256 threads per query.
constexpr int ST_BINS = 2048;

// cnt non-null: query q's values are its first cnt[q] of ns (its
// candidates); fewer than k of them, or more than ns, give no bound (+inf)
__global__ void __launch_bounds__(256) k_sample_thresholds(const float *__restrict__ approx, int64_t stride,
                                                           const int32_t *__restrict__ cnt, int kmax,
                                                           const int32_t *__restrict__ k_q,
                                                           const float *__restrict__ qnorm_f,
                                                           const unsigned *__restrict__ max_norm,
                                                           const int *__restrict__ inexact, int d,
                                                           float *__restrict__ thr) {
    __shared__ unsigned hist[ST_BINS];
    __shared__ unsigned s_lo[4], s_hi[4];
    __shared__ unsigned s_edge;
    const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
    const int64_t q = blockIdx.x;
    int k = k_q ? k_q[q] : kmax;
    k = k < 1 ? 1 : (k > kmax ? kmax : k);
    const int64_t ns = cnt ? (int64_t)cnt[q] : stride;
    if (k > ns || ns > stride) {
        if (tid == 0) thr[q] = __builtin_inff();
        return;
    }
    const float *a = approx + q * stride;
    // the range of the query's keys
    unsigned lo = 0xFFFFFFFFu, hi = 0u;
    for (int64_t i = tid; i < ns; i += 256) {
        const unsigned u = float_key(a[i]);
        lo = u < lo ? u : lo;
        hi = u > hi ? u : hi;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned xl = __shfl_xor(lo, o), xh = __shfl_xor(hi, o);
        lo = xl < lo ? xl : lo;
        hi = xh > hi ? xh : hi;
    }
    if (lane == 0) {
        s_lo[wave] = lo;
        s_hi[wave] = hi;
    }
    for (int b = tid; b < ST_BINS; b += 256) hist[b] = 0;
    __syncthreads();
    lo = s_lo[0];
    hi = s_hi[0];
    for (int w = 1; w < 4; ++w) {
        lo = s_lo[w] < lo ? s_lo[w] : lo;
        hi = s_hi[w] > hi ? s_hi[w] : hi;
    }
    int sh = 0;
    while (((hi - lo) >> sh) >= (unsigned)ST_BINS) ++sh;
    for (int64_t i = tid; i < ns; i += 256) atomicAdd(&hist[(float_key(a[i]) - lo) >> sh], 1u);
    __syncthreads();
    // the bin where the cumulative count reaches k (wave 0: 32 bins a lane)
    if (wave == 0) {
        unsigned own = 0;
        for (int b = 0; b < ST_BINS / 64; ++b) own += hist[lane * (ST_BINS / 64) + b];
        unsigned incl = own;
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned v = __shfl_up(incl, o);
            if (lane >= o) incl += v;
        }
        const unsigned excl = incl - own;
        if (excl < (unsigned)k && (unsigned)k <= incl) {
            unsigned run = excl;
            int b = lane * (ST_BINS / 64);
            while (run + hist[b] < (unsigned)k) run += hist[b++];
            // the bin's upper edge, at most hi
            const unsigned span = sh >= 32 ? 0xFFFFFFFFu : ((1u << sh) - 1u);
            const unsigned blo = lo + ((unsigned)b << sh);
            s_edge = (hi - blo) < span ? hi : blo + span;
        }
    }
    __syncthreads();
    if (tid != 0) return;
    thr[q] = threshold_from_kth(key_float(s_edge), qnorm_f[q], __uint_as_float(*max_norm), *inexact, d);
}

// Per query, over all its nb rows: the k-th smallest of its approximate-
// distance keys (keys[q * nb + row], approx_key16), A = the largest float
// with that key (at least the k-th smallest approximate distance), the filter
// threshold T from A as in k_sample_thresholds, and the candidates: every
// row whose key is at most T's (a superset of the rows with approximate
// distance <= T, by the keys' monotonicity), written to cand_row[q * cap +
// i] for i < cap; cnt[q] = their number (more than cap: the caller ranks
// every row). The keys of one query span few values: a 2,048-bin histogram
// over their [min, max] (qlo[q], qhi[q], from the key pass) finds the k-th.
// One workgroup of 256 threads per query; its keys load eight at a time (nb
// a multiple of 8), once for the histogram and once for the candidates
// (mostly from the caches).
__global__ void __launch_bounds__(256) k_dense_select(const uint16_t *__restrict__ keys, int64_t nb, int kmax,
                                                      const unsigned *__restrict__ qlo,
                                                      const unsigned *__restrict__ qhi,
                                                      const int32_t *__restrict__ k_q,
                                                      const float *__restrict__ qnorm_f,
                                                      const unsigned *__restrict__ max_norm,
                                                      const int *__restrict__ inexact, int d, int cap,
                                                      int32_t *__restrict__ cnt, int32_t *__restrict__ cand_row) {
    __shared__ unsigned hist[ST_BINS];
    __shared__ unsigned s_kt, s_n;
    const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
    const int64_t q = blockIdx.x;
    int k = k_q ? k_q[q] : kmax;
    k = k < 1 ? 1 : (k > kmax ? kmax : k);
    const int nv = (int)(nb / 8);
    const uint4 *kv = reinterpret_cast<const uint4 *>(keys + q * nb);
    // f(row, key) over the query's keys, each thread its own rows
    auto for_keys = [&](auto f) {
        for (int v = tid; v < nv; v += 256) {
            const uint4 w = kv[v];
            const unsigned ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
            for (int j = 0; j < 8; ++j) f((int64_t)8 * v + j, (ws[j >> 1] >> (16 * (j & 1))) & 0xFFFFu, true);
        }
    };
    // the key range the key pass recorded
    const unsigned lo = qlo[q], hi = qhi[q] > lo ? qhi[q] : lo;
    for (int b = tid; b < ST_BINS; b += 256) hist[b] = 0;
    if (tid == 0) {
        s_n = 0;
        s_kt = 0xFFFFu;   // fewer than k rows: every row is a candidate
    }
    __syncthreads();
    int sh = 0;
    while (((hi - lo) >> sh) >= (unsigned)ST_BINS) ++sh;
    if ((int64_t)k <= nb) {
        for_keys([&](int64_t, unsigned key, bool) { atomicAdd(&hist[(key - lo) >> sh], 1u); });
        __syncthreads();
        // the bin where the cumulative count reaches k (wave 0: 32 bins a
        // lane), its largest key, the float bound A and the threshold's key
        if (wave == 0) {
            unsigned own = 0;
            for (int b = 0; b < ST_BINS / 64; ++b) own += hist[lane * (ST_BINS / 64) + b];
            unsigned incl = own;
            for (int o = 1; o < 64; o <<= 1) {
                const unsigned v = __shfl_up(incl, o);
                if (lane >= o) incl += v;
            }
            const unsigned excl = incl - own;
            if (excl < (unsigned)k && (unsigned)k <= incl) {
                unsigned run = excl;
                int b = lane * (ST_BINS / 64);
                while (run + hist[b] < (unsigned)k) run += hist[b++];
                const unsigned top = lo + (((unsigned)b + 1u) << sh) - 1u;
                const float A = approx_key_upper(top < hi ? top : hi);
                const float T = threshold_from_kth(A, qnorm_f[q], __uint_as_float(*max_norm), *inexact, d);
                s_kt = approx_key16(T);
            }
        }
    }
    __syncthreads();
    const unsigned kt = s_kt;
    // the candidates, compacted per wave with one LDS atomic per wave and key
    // position
    for_keys([&](int64_t r, unsigned key, bool) {
        const bool pass = key <= kt;
        const unsigned long long m = __builtin_amdgcn_ballot_w64(pass);
        if (m == 0) return;
        unsigned base = 0;
        if (lane == (int)__builtin_ctzll(m)) base = atomicAdd(&s_n, (unsigned)__popcll(m));
        base = (unsigned)__shfl((int)base, (int)__builtin_ctzll(m));
        if (pass) {
            const unsigned pos = base + (unsigned)__builtin_amdgcn_mbcnt_hi(
                                            (unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
            if (pos < (unsigned)cap) cand_row[q * cap + pos] = (int32_t)r;
        }
    });
    __syncthreads();
    if (tid == 0) cnt[q] = (int32_t)s_n;
}

// Exact reference-order rank of every candidate, then the k best by
// (rank, row). One wave per query. Queries whose candidate list overflowed
// write nothing (the host reruns them on the exact path). With cand_ap, only
// the candidates whose approximate distance is within thr2[q] are ranked:
// they are packed 64 at a time through LDS, so every lane ranks a row.
// NB > 0 (d = 32 NB, 16-byte aligned rows): each row is ranked by a quad of
// lanes, lane j of the quad holding accumulator slots 8 j .. 8 j + 7 of the
// reference's 32-slot order (elements 8 j + i + 32 b, fused in block order),
// so the quad's loads of a block cover one 128-byte line; the quad then
// combines the slots as the reference does, (acc[l] + acc[8 + l]) +
// (acc[16 + l] + acc[24 + l]) by two quad-permute moves, and lane 0 takes
// the horizontal sum: the same bits as sqdiff_exact, 16 rows a step.
template <typename CT, int NB>
__global__ void __launch_bounds__(64) k_rank_candidates(const float *__restrict__ queries, int d,
                                                        const float *__restrict__ base,
                                                        const int32_t *__restrict__ cnt,
                                                        const CT *__restrict__ cand_row,
                                                        const float *__restrict__ cand_ap,
                                                        const float *__restrict__ thr2, int cap, int64_t nb,
                                                        int k, const int32_t *__restrict__ k_q, int ccap, int sqrt_out,
                                                        int64_t *__restrict__ out_ids, double *__restrict__ out_dist) {
    extern __shared__ __align__(16) unsigned char s_raw[];
    float *s_keys = reinterpret_cast<float *>(s_raw);
    int64_t *s_ids = reinterpret_cast<int64_t *>(s_keys + ccap);
    int64_t *s_buf = s_ids + ccap;   // [2 kWave] rows waiting to be ranked
    float *s_d = reinterpret_cast<float *>(s_buf + 2 * kWave);   // [kWave] distances (NB > 0)
    const int lane = threadIdx.x;
    const int64_t q = blockIdx.x;
    // a query whose candidates overflowed the room ranks every row (nb > 0),
    // or is left to the caller (nb == 0)
    const bool every = cnt[q] > cap;
    if (every && nb == 0) return;
    const int64_t n = every ? nb : cnt[q];
    int kq = k_q ? k_q[q] : k;
    kq = kq < 1 ? 1 : (kq > k ? k : kq);
    WaveTopK sel;
    sel.init(s_keys, s_ids, kq, ccap);
    const float *qv = queries + q * d;
    // NB > 0: this lane's query values, elements 8 j + i + 32 b
    float qr[NB > 0 ? NB : 1][8];
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int i = 0; i < 8; ++i) qr[b][i] = qv[32 * b + 8 * (lane & 3) + i];
    // the valid lanes of every call are a prefix of the wave (lane 0 valid)
    auto rank = [&](int64_t row, bool valid) {
        float key = 0.f;
        if constexpr (NB > 0) {
            const int64_t row0 = (int64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)row) |
                                 ((int64_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)row >> 32)) << 32);
            const int64_t rv = valid ? row : row0;
            const int j = lane & 3, qd = lane >> 2;
            // step s: the quad ranks the row of lane 16 s + qd
            float4 x[4][2 * NB];
#pragma unroll
            for (int st = 0; st < 4; ++st) {
                const int src = 16 * st + qd;
                const int64_t rr = (int64_t)(uint32_t)__builtin_amdgcn_ds_bpermute(4 * src, (int)(uint32_t)rv) |
                                   ((int64_t)__builtin_amdgcn_ds_bpermute(4 * src, (int)(uint32_t)((uint64_t)rv >> 32)) << 32);
                const float4 *xr = reinterpret_cast<const float4 *>(base + rr * d + 8 * j);
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    x[st][2 * b] = xr[8 * b];
                    x[st][2 * b + 1] = xr[8 * b + 1];
                }
            }
#pragma unroll
            for (int st = 0; st < 4; ++st) {
                float acc[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) acc[i] = 0.f;
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    const float4 lo = x[st][2 * b], hi = x[st][2 * b + 1];
                    const float xs[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        const float dd = qr[b][i] - xs[i];
                        acc[i] = __builtin_fmaf(dd, dd, acc[i]);
                    }
                }
                float s8[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    // quad_perm [1,0,3,2], then [2,3,0,1]
                    const float t = acc[i] + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(acc[i]), 0xB1, 0xF, 0xF, false));
                    s8[i] = t + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(t), 0x4E, 0xF, 0xF, false));
                }
                const float dist = hsum8(s8);
                if (j == 0) s_d[16 * st + qd] = dist;
            }
            wave_sync();
            key = s_d[lane];
            wave_sync();   // read before the next call writes
        } else {
            if (valid) key = sqdiff_exact(PtrLoad{qv}, PtrLoad{base + row * d}, d);
        }
        sel.push(key, row, valid);
    };
    if (every || !cand_ap) {
        for (int64_t b = 0; b < n; b += kWave) {
            const int64_t t = b + lane;
            const bool valid = t < n;
            rank(valid ? (every ? t : (int64_t)cand_row[q * cap + t]) : 0, valid);
        }
    } else {
        const float t2 = thr2[q];
        int nbuf = 0;   // wave-uniform
        for (int64_t b = 0; b < n; b += kWave) {
            const int64_t t = b + lane;
            const bool ok = t < n && cand_ap[q * cap + t] <= t2;
            const unsigned long long m = __builtin_amdgcn_ballot_w64(ok);
            if (ok) {
                const int at = nbuf + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                                      __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
                s_buf[at] = (int64_t)cand_row[q * cap + t];
            }
            nbuf += __popcll(m);
            __syncthreads();
            if (nbuf >= kWave) {
                rank(s_buf[lane], true);
                __syncthreads();   // every lane has read its row
                if (lane < nbuf - kWave) s_buf[lane] = s_buf[kWave + lane];
                nbuf -= kWave;
                __syncthreads();
            }
        }
        if (nbuf > 0) rank(lane < nbuf ? s_buf[lane] : 0, lane < nbuf);
    }
    int got = sel.finish();
    got = got < kq ? got : kq;
    for (int t = lane; t < k; t += kWave) {
        double dd = t < got ? (double)sel.keys[t] : __builtin_inf();
        if (sqrt_out && t < got) dd = __builtin_sqrt(dd);
        out_ids[q * k + t] = t < got ? sel.ids[t] : (int64_t)-1;
        out_dist[q * k + t] = dd;
    }
}

}  // namespace dev

void launch_to_bf16(const float *x, int64_t n, int d, int dp, __bf16 *out, float *norm_b, float *norm_f,
                    unsigned *max_norm, int *inexact, hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL(dev::k_to_bf16, dim3((unsigned)ceil_div(n, dev::TB_ROWS)), dim3(256), 0, st, x, n, d, dp, out,
                       norm_b, norm_f, max_norm, inexact);
    HANN_LAUNCH_CHECK();
}

void launch_hint_rank(const float *queries, int64_t nq, int d, const float *base, const int32_t *hint,
                      double *tau_rank, hipStream_t st) {
    if (nq == 0) return;
    hipLaunchKernelGGL(dev::k_hint_rank, dim3((unsigned)ceil_div(nq, 256)), dim3(256), 0, st, queries, nq, d, base,
                       hint, tau_rank);
    HANN_LAUNCH_CHECK();
}

void launch_thresholds(const double *tau_rank, int k, int nq, const float *qnorm_f, const unsigned *max_norm,
                       const int *inexact, int d, float *thr, hipStream_t st) {
    if (nq == 0) return;
    hipLaunchKernelGGL(dev::k_thresholds, dim3((unsigned)ceil_div(nq, 256)), dim3(256), 0, st, tau_rank, k, nq,
                       qnorm_f, max_norm, inexact, d, thr);
    HANN_LAUNCH_CHECK();
}

static size_t mfma_lds(int dp, bool dense) {
    const int ld = (dp < dev::MF_KC ? dp : dev::MF_KC) + dev::MF_PAD;
    return (size_t)2 * dev::MF_TILE * ld * 2 + (size_t)2 * dev::MF_TILE * 4 +
           (dense ? 0 : (size_t)dev::MF_TILE * (2 + dev::MF_SLOTS) * 4 + (size_t)dev::MF_OVF * 8 + 16);
}

static size_t mfma_lds_r(int dp, bool dense) {
    return (size_t)dev::MF_TILE * (dp + dev::MF_PAD) * 2 + (size_t)3 * dev::MF_TILE * 4 +
           (dense ? 0
                  : (size_t)dev::MF_TILE * 2 * 4 + (size_t)dev::MF_TILE * dev::MF_SLOTS_R * 8 +
                        (size_t)dev::MF_OVF_R * 16 + 16);
}

// dp <= 128: the register-resident kernel, grid (query tiles, row groups)
template <bool DENSE, bool KEY16 = false>
static void launch_pass_r(const __bf16 *Q, int64_t nq, const __bf16 *X, int64_t nb, int dp, const float *qn,
                          const float *xn, const float *thr, int rt, int32_t *cnt, int32_t *cand_row,
                          float *cand_ap, int cap, float *dense, int64_t ld_dense, hipStream_t st) {
    const int64_t rg = ceil_div(nb, (int64_t)dev::MF_TILE * rt), qt = ceil_div(nq, dev::MF_TILE);
    if (qt > 0x7fffffff) fail(HANN_ERR_UNSUPPORTED, "BF16 candidate pass: too many queries");
    const size_t shm = mfma_lds_r(dp, DENSE);
    for (int64_t y0 = 0; y0 < rg; y0 += 65535) {
        const int64_t yn = rg - y0 < 65535 ? rg - y0 : 65535;
        const dim3 grid((unsigned)qt, (unsigned)yn);
#define HANN_PASS_R(KS)                                                                                         \
    do {                                                                                                        \
        HANN_HIP(hipFuncSetAttribute((const void *)dev::k_mfma_pass_r<DENSE, KS, KEY16>,                        \
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));                    \
        hipLaunchKernelGGL((dev::k_mfma_pass_r<DENSE, KS, KEY16>), grid, dim3(256), shm, st, Q, nq, X, nb, qn, xn, \
                           thr, rt, cnt, cand_row, cand_ap, cap, dense, ld_dense, y0);                          \
        HANN_LAUNCH_CHECK();                                                                                    \
    } while (0)
        switch (dp) {
        case 32: HANN_PASS_R(2); break;
        case 64: HANN_PASS_R(4); break;
        case 96: HANN_PASS_R(6); break;
        case 128: HANN_PASS_R(8); break;
        default: fail(HANN_ERR_UNSUPPORTED, strf("BF16 candidate pass: padded width %d", dp));
        }
#undef HANN_PASS_R
    }
}

bool mfma_filter_writes_approx(int dp) { return dp <= dev::MF_KC; }

int mfma_filter_slots(int dp) { return dp <= dev::MF_KC ? dev::MF_SLOTS_R : dev::MF_SLOTS; }

void launch_mfma_filter(const __bf16 *Q, int64_t nq, const __bf16 *X, int64_t nb, int dp, const float *qn,
                        const float *xn, const float *thr, int rt_per_wg, int32_t *cnt, int32_t *cand_row,
                        float *cand_ap, int cap, hipStream_t st) {
    if (nq == 0 || nb == 0) return;
    if (rt_per_wg < 1) rt_per_wg = 1;
    if (dp <= dev::MF_KC) {
        if (!cand_ap) fail(HANN_ERR_INVALID_ARG, "BF16 candidate pass: no room for the approximate distances");
        launch_pass_r<false>(Q, nq, X, nb, dp, qn, xn, thr, rt_per_wg, cnt, cand_row, cand_ap, cap, nullptr, 0, st);
        return;
    }
    const int64_t rg = ceil_div(nb, (int64_t)dev::MF_TILE * rt_per_wg), qt = ceil_div(nq, dev::MF_TILE);
    const size_t shm = mfma_lds(dp, false);
    HANN_HIP(hipFuncSetAttribute((const void *)dev::k_mfma_pass<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)shm));
    for (int64_t y0 = 0; y0 < qt; y0 += 65535) {
        const int64_t yn = qt - y0 < 65535 ? qt - y0 : 65535;
        hipLaunchKernelGGL(dev::k_mfma_pass<false>, dim3((unsigned)rg, (unsigned)yn), dim3(256), shm, st,
                           Q + y0 * dev::MF_TILE * dp, nq - y0 * dev::MF_TILE, X, nb, dp, qn + y0 * dev::MF_TILE, xn,
                           thr + y0 * dev::MF_TILE, rt_per_wg, cnt + y0 * dev::MF_TILE,
                           cand_row + y0 * dev::MF_TILE * (int64_t)cap, cap, nullptr, (int64_t)0);
        HANN_LAUNCH_CHECK();
    }
}

void launch_mfma_dense(const __bf16 *Q, int64_t nq, const __bf16 *X, int64_t nb, int dp, const float *qn,
                       const float *xn, float *dense, hipStream_t st) {
    if (nq == 0 || nb == 0) return;
    const int rt = 4;
    if (dp <= dev::MF_KC) {
        launch_pass_r<true>(Q, nq, X, nb, dp, qn, xn, nullptr, rt, nullptr, nullptr, nullptr, 0, dense, nb, st);
        return;
    }
    const int64_t rg = ceil_div(nb, (int64_t)dev::MF_TILE * rt), qt = ceil_div(nq, dev::MF_TILE);
    const size_t shm = mfma_lds(dp, true);
    HANN_HIP(hipFuncSetAttribute((const void *)dev::k_mfma_pass<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)shm));
    for (int64_t y0 = 0; y0 < qt; y0 += 65535) {
        const int64_t yn = qt - y0 < 65535 ? qt - y0 : 65535;
        hipLaunchKernelGGL(dev::k_mfma_pass<true>, dim3((unsigned)rg, (unsigned)yn), dim3(256), shm, st,
                           Q + y0 * dev::MF_TILE * dp, nq - y0 * dev::MF_TILE, X, nb, dp, qn + y0 * dev::MF_TILE, xn,
                           nullptr, rt, nullptr, nullptr, 0, dense + y0 * dev::MF_TILE * nb, nb);
        HANN_LAUNCH_CHECK();
    }
}

void launch_mfma_keys(const __bf16 *Q, int64_t nq, const __bf16 *X, int64_t nb, int dp, const float *qn,
                      const float *xn, uint16_t *keys, unsigned *qlo, unsigned *qhi, hipStream_t st) {
    if (nq == 0 || nb == 0) return;
    if (dp > dev::MF_KC || nb % 4 != 0)
        fail(HANN_ERR_UNSUPPORTED, strf("BF16 key pass: padded width %d, %lld rows", dp, (long long)nb));
    // fill kernels, not memsets: the search may be captured in a graph
    launch_fill_i32(reinterpret_cast<int32_t *>(qlo), nq, 0xFFFF, st);
    launch_fill_i32(reinterpret_cast<int32_t *>(qhi), nq, 0, st);
    launch_pass_r<true, true>(Q, nq, X, nb, dp, qn, xn, nullptr, 4, reinterpret_cast<int32_t *>(qlo),
                              reinterpret_cast<int32_t *>(qhi), nullptr, 0, reinterpret_cast<float *>(keys), nb, st);
}

int64_t dense_select_max_rows() { return 65536; }

void launch_dense_select(const uint16_t *keys, const unsigned *qlo, const unsigned *qhi, int64_t nq, int64_t nb,
                         int k, const int32_t *k_q, const float *qnorm_f, const unsigned *max_norm,
                         const int *inexact, int d, int cap, int32_t *cnt, int32_t *cand_row, hipStream_t st) {
    if (nq == 0) return;
    if (nb % 8 != 0 || nb > dense_select_max_rows())
        fail(HANN_ERR_UNSUPPORTED, strf("dense select: %lld rows (a multiple of 8 up to %lld)", (long long)nb,
                                        (long long)dense_select_max_rows()));
    for (int64_t q0 = 0; q0 < nq; q0 += 65535) {
        const int64_t qn_ = nq - q0 < 65535 ? nq - q0 : 65535;
        hipLaunchKernelGGL(dev::k_dense_select, dim3((unsigned)qn_), dim3(256), 0, st, keys + q0 * nb, nb, k,
                           qlo + q0, qhi + q0, k_q ? k_q + q0 : nullptr, qnorm_f + q0, max_norm, inexact, d, cap, cnt + q0,
                           cand_row + q0 * (int64_t)cap);
        HANN_LAUNCH_CHECK();
    }
}

void launch_gather_bf16(const __bf16 *X, int dp, const float *xn, int64_t stride, int64_t n, __bf16 *out,
                        float *out_n, hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL(dev::k_gather_bf16, dim3((unsigned)ceil_div(n * dp, 256)), dim3(256), 0, st, X, dp, xn, stride,
                       n, out, out_n);
    HANN_LAUNCH_CHECK();
}

void launch_sample_thresholds(const float *approx, int64_t nq, int64_t ns, const int32_t *cnt, int k,
                              const int32_t *k_q, const float *qnorm_f, const unsigned *max_norm, const int *inexact,
                              int d, float *thr, hipStream_t st) {
    if (nq == 0) return;
    hipLaunchKernelGGL(dev::k_sample_thresholds, dim3((unsigned)nq), dim3(256), 0, st, approx, ns, cnt, k, k_q,
                       qnorm_f, max_norm, inexact, d, thr);
    HANN_LAUNCH_CHECK();
}

static size_t rank_lds(int ccap) { return (size_t)ccap * 12 + (size_t)2 * dev::kWave * 8 + (size_t)dev::kWave * 4 + 16; }

// the rank kernel for rows of d floats at base: rows ranked by quads for
// d in {32, 64, 96} with 16-byte aligned rows, one per lane otherwise (at
// d = 128 the quads' row registers halve the waves a SIMD holds, and the
// k-means assignment and the 1M-row top-k measured no faster)
template <typename CT>
static decltype(&dev::k_rank_candidates<CT, 0>) rank_kernel(int d, const float *base) {
    if (d % 32 != 0 || ((uintptr_t)base & 15) != 0) return dev::k_rank_candidates<CT, 0>;
    switch (d / 32) {
    case 1: return dev::k_rank_candidates<CT, 1>;
    case 2: return dev::k_rank_candidates<CT, 2>;
    case 3: return dev::k_rank_candidates<CT, 3>;
    default: return dev::k_rank_candidates<CT, 0>;
    }
}

void launch_rank_candidates(const float *queries, int64_t nq, int d, const float *base, int64_t nb,
                            const int32_t *cnt, const int32_t *cand_row, const float *cand_ap, const float *thr2,
                            int cap, int k, const int32_t *k_q, int sqrt_out, int64_t *out_ids, double *out_dist,
                            hipStream_t st) {
    if (nq == 0) return;
    const int ccap = 2 * kp_for(k);
    const size_t shm = rank_lds(ccap);
    const auto kern = rank_kernel<int32_t>(d, base);
    HANN_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    for (int64_t q0 = 0; q0 < nq; q0 += 65535) {
        const int64_t qn = nq - q0 < 65535 ? nq - q0 : 65535;
        hipLaunchKernelGGL(kern, dim3((unsigned)qn), dim3(64), shm, st, queries + q0 * d, d,
                           base, cnt + q0, cand_row + q0 * (int64_t)cap, cand_ap ? cand_ap + q0 * (int64_t)cap : nullptr,
                           thr2 ? thr2 + q0 : nullptr, cap, nb, k, k_q ? k_q + q0 : nullptr, ccap, sqrt_out,
                           out_ids + q0 * k, out_dist + q0 * k);
        HANN_LAUNCH_CHECK();
    }
}

void launch_rank_rows64(const float *queries, int64_t nq, int d, const float *base, int64_t nb, const int32_t *cnt,
                        const int64_t *rows, int cap, int k, const int32_t *k_q, int sqrt_out, int64_t *out_ids,
                        double *out_dist, hipStream_t st) {
    if (nq == 0) return;
    const int ccap = 2 * kp_for(k);
    const size_t shm = rank_lds(ccap);
    const auto kern = rank_kernel<int64_t>(d, base);
    HANN_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    for (int64_t q0 = 0; q0 < nq; q0 += 65535) {
        const int64_t qn = nq - q0 < 65535 ? nq - q0 : 65535;
        hipLaunchKernelGGL(kern, dim3((unsigned)qn), dim3(64), shm, st, queries + q0 * d,
                           d, base, cnt + q0, rows + q0 * (int64_t)cap, nullptr, nullptr, cap, nb, k,
                           k_q ? k_q + q0 : nullptr, ccap,
                           sqrt_out, out_ids + q0 * k, out_dist + q0 * k);
        HANN_LAUNCH_CHECK();
    }
}

}  // namespace hann
