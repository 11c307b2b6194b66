// distance.hip -- batch distance kernels and exact batch top-k (gfx950).
//
// All FP32 kernels here reproduce the reference's AVX2+FMA arithmetic
// (core/simd_kernels.inc.h) bit for bit; see common.h for the order.
#include <mutex>
#include <vector>

#include "common.h"
#include "kernels.h"
#include "kmeans.h"
#include "scan_list.h"

namespace hann {
namespace dev {

// ---------------------------------------------------------------------------
// One thread per (query, row): out[q*nb + i] = (double) metric(q, row).
// Used by the drop-in simd_* symbols and hann_distance_batch, where the
// host<->device copy dominates; rows are read straight from global memory.
__global__ void __launch_bounds__(256) k_distance_batch(int metric, const float *__restrict__ queries,
                                                        int64_t nq, const float *__restrict__ base,
                                                        int64_t nb, int dim, double *__restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t q = blockIdx.y;
    if (i >= nb || q >= nq) return;
    const float *a = queries + q * dim;
    const float *b = base + i * dim;
    float v = metric_exact(metric, PtrLoad{a}, PtrLoad{b}, dim);
    out[q * nb + i] = (double)v;
}

// ---------------------------------------------------------------------------
// Fused exact distance + top-k.
//
// A workgroup of 4 waves owns a tile of 16 queries (4 per wave) and a range
// of base rows. Rows are staged through LDS 64 at a time, one 32-float block
// of dimensions at a time (the reference's own block structure), with
// coalesced loads; lane l of every wave then owns row l of the tile and
// keeps the 32 block accumulators of each of its 4 (query, row) pairs in
// registers. A wave feeds each of its queries' candidates into that query's
// WaveTopK. The partial top-k of every (query, row range) goes to `part`,
// and k_topk_merge reduces the ranges.
constexpr int TK_WAVES = 4;
constexpr int TK_QPW = 4;                     // queries per wave
constexpr int TK_QT = TK_WAVES * TK_QPW;      // queries per workgroup
constexpr int TK_ROWS = 64;                   // rows per tile
constexpr int TK_STRIDE = 36;                 // LDS row stride in floats (bank-conflict free b128)

template <int KP>
__global__ void __launch_bounds__(256) k_topk_exact(int metric_rank, const float *__restrict__ queries,
                                                    int64_t nq, const float *__restrict__ base,
                                                    int64_t nb, int dim, int k, int64_t rows_per_split,
                                                    float *__restrict__ part_key,
                                                    int64_t *__restrict__ part_id) {
    constexpr int CAP = 2 * KP;
    __shared__ float s_rows[TK_ROWS * TK_STRIDE];
    __shared__ float s_q[TK_QT * 32];
    __shared__ float s_keys[TK_QT * CAP];
    __shared__ int64_t s_ids[TK_QT * CAP];

    const int tid = threadIdx.x;
    const int wave = tid / kWave;
    const int lane = tid % kWave;
    const int64_t q0 = (int64_t)blockIdx.x * TK_QT;
    const int split = blockIdx.y;
    const int64_t r_begin = (int64_t)split * rows_per_split;
    int64_t r_end = r_begin + rows_per_split;
    if (r_end > nb) r_end = nb;

    WaveTopK sel[TK_QPW];
#pragma unroll
    for (int j = 0; j < TK_QPW; ++j) {
        int ql = wave * TK_QPW + j;
        sel[j].init(s_keys + ql * CAP, s_ids + ql * CAP, k, CAP);
    }

    const int nblk = dim / 32;

    for (int64_t t0 = r_begin; t0 < r_end; t0 += TK_ROWS) {
        const int64_t row = t0 + lane;
        const bool valid = row < r_end;
        // acc[j][c]: pair (query j of this wave, row), block slot c
        float acc[TK_QPW][32];
#pragma unroll
        for (int j = 0; j < TK_QPW; ++j)
#pragma unroll
            for (int c = 0; c < 32; ++c) acc[j][c] = 0.f;

        for (int b = 0; b < nblk; ++b) {
            __syncthreads();
            // stage 64 rows x 32 floats: 512 float4, 2 per thread
            for (int v = tid; v < TK_ROWS * 8; v += 256) {
                int r = v / 8, c4 = v % 8;
                int64_t gr = t0 + r;
                float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
                if (gr < r_end) {
                    const float *src = base + gr * dim + b * 32 + c4 * 4;
                    x = make_float4(src[0], src[1], src[2], src[3]);
                }
                *reinterpret_cast<float4 *>(&s_rows[r * TK_STRIDE + c4 * 4]) = x;
            }
            for (int v = tid; v < TK_QT * 32; v += 256) {
                int ql = v / 32, c = v % 32;
                int64_t gq = q0 + ql;
                s_q[v] = gq < nq ? queries[gq * dim + b * 32 + c] : 0.f;
            }
            __syncthreads();
#pragma unroll
            for (int c4 = 0; c4 < 8; ++c4) {
                float4 x = *reinterpret_cast<const float4 *>(&s_rows[lane * TK_STRIDE + c4 * 4]);
#pragma unroll
                for (int j = 0; j < TK_QPW; ++j) {
                    float4 y = *reinterpret_cast<const float4 *>(&s_q[(wave * TK_QPW + j) * 32 + c4 * 4]);
                    float d0 = y.x - x.x, d1 = y.y - x.y, d2 = y.z - x.z, d3 = y.w - x.w;
                    acc[j][c4 * 4 + 0] = __builtin_fmaf(d0, d0, acc[j][c4 * 4 + 0]);
                    acc[j][c4 * 4 + 1] = __builtin_fmaf(d1, d1, acc[j][c4 * 4 + 1]);
                    acc[j][c4 * 4 + 2] = __builtin_fmaf(d2, d2, acc[j][c4 * 4 + 2]);
                    acc[j][c4 * 4 + 3] = __builtin_fmaf(d3, d3, acc[j][c4 * 4 + 3]);
                }
            }
        }
        // cleanup blocks of 8 into slots 0-7, combine, reduce, scalar tail
        float key[TK_QPW];
#pragma unroll
        for (int j = 0; j < TK_QPW; ++j) {
            int64_t gq = q0 + wave * TK_QPW + j;
            const float *qv = queries + (gq < nq ? gq : 0) * dim;
            const float *xv = base + (valid ? row : 0) * dim;
            int i = nblk * 32;
            const int lim = dim - (dim % 8);
            for (; i < lim; i += 8) {
#pragma unroll
                for (int l = 0; l < 8; ++l) {
                    float d = qv[i + l] - xv[i + l];
                    acc[j][l] = __builtin_fmaf(d, d, acc[j][l]);
                }
            }
            float s[8];
#pragma unroll
            for (int l = 0; l < 8; ++l) s[l] = (acc[j][l] + acc[j][8 + l]) + (acc[j][16 + l] + acc[j][24 + l]);
            float sum = hsum8(s);
            for (; i < dim; ++i) {
                float d = qv[i] - xv[i];
                sum = __builtin_fmaf(d, d, sum);
            }
            key[j] = sum;
        }
        (void)metric_rank;
#pragma unroll
        for (int j = 0; j < TK_QPW; ++j) {
            int64_t gq = q0 + wave * TK_QPW + j;
            sel[j].push(key[j], row, valid && gq < nq);
        }
    }
#pragma unroll
    for (int j = 0; j < TK_QPW; ++j) {
        int64_t gq = q0 + wave * TK_QPW + j;
        int got = sel[j].finish();
        if (gq < nq) {
            float *pk = part_key + (gq * gridDim.y + split) * (int64_t)k;
            int64_t *pi = part_id + (gq * gridDim.y + split) * (int64_t)k;
            for (int s = lane; s < k; s += kWave) {
                pk[s] = s < got ? sel[j].keys[s] : __builtin_inff();
                pi[s] = s < got ? sel[j].ids[s] : (int64_t)-1;
            }
        }
    }
}

// Generic (any metric) fused top-k: one wave per query and row range, rows
// read from global memory. Serves Manhattan and cosine, which the tiled
// kernel above does not specialise.
template <int KP>
__global__ void __launch_bounds__(64) k_topk_generic(int metric_rank, const float *__restrict__ queries,
                                                     int64_t nq, const float *__restrict__ base,
                                                     int64_t nb, int dim, int k, int64_t rows_per_split,
                                                     float *__restrict__ part_key,
                                                     int64_t *__restrict__ part_id) {
    constexpr int CAP = 2 * KP;
    __shared__ float s_keys[CAP];
    __shared__ int64_t s_ids[CAP];
    const int lane = threadIdx.x;
    const int64_t q = blockIdx.x;
    const int split = blockIdx.y;
    const int64_t r_begin = (int64_t)split * rows_per_split;
    int64_t r_end = r_begin + rows_per_split;
    if (r_end > nb) r_end = nb;
    WaveTopK sel;
    sel.init(s_keys, s_ids, k, CAP);
    const float *qv = queries + q * dim;
    for (int64_t t0 = r_begin; t0 < r_end; t0 += kWave) {
        int64_t row = t0 + lane;
        bool valid = row < r_end;
        float v = 0.f;
        if (valid) v = metric_exact(metric_rank, PtrLoad{qv}, PtrLoad{base + row * dim}, dim);
        sel.push(v, row, valid);
    }
    int got = sel.finish();
    float *pk = part_key + (q * gridDim.y + split) * (int64_t)k;
    int64_t *pi = part_id + (q * gridDim.y + split) * (int64_t)k;
    for (int s = lane; s < k; s += kWave) {
        pk[s] = s < got ? sel.keys[s] : __builtin_inff();
        pi[s] = s < got ? sel.ids[s] : (int64_t)-1;
    }
}

// Merge nsplit partial lists of k candidates per query into the final k,
// and convert rank to distance in double (sqrt for Euclidean).
template <int KP>
__global__ void __launch_bounds__(64) k_topk_merge(const float *__restrict__ part_key,
                                                   const int64_t *__restrict__ part_id, int nsplit,
                                                   int k, int sqrt_out, int64_t *__restrict__ out_ids,
                                                   double *__restrict__ out_dist) {
    constexpr int CAP = 2 * KP;
    __shared__ float s_keys[CAP];
    __shared__ int64_t s_ids[CAP];
    const int lane = threadIdx.x;
    const int64_t q = blockIdx.x;
    WaveTopK sel;
    sel.init(s_keys, s_ids, k, CAP);
    const int64_t total = (int64_t)nsplit * k;
    const float *pk = part_key + q * total;
    const int64_t *pi = part_id + q * total;
    for (int64_t t0 = 0; t0 < total; t0 += kWave) {
        int64_t s = t0 + lane;
        bool valid = s < total && pi[s] >= 0;
        sel.push(valid ? pk[s] : 0.f, valid ? pi[s] : 0, valid);
    }
    int got = sel.finish();
    for (int s = lane; s < k; s += kWave) {
        int64_t id = s < got ? sel.ids[s] : -1;
        double d = s < got ? (double)sel.keys[s] : __builtin_inf();
        if (sqrt_out && s < got) d = __builtin_sqrt(d);
        out_ids[q * k + s] = id;
        out_dist[q * k + s] = d;
    }
}

}  // namespace dev

// ------------------------------------------------------------ launchers --
// Normalize n rows of dim floats in place, one wave per row, in the order
// of core/simd_kernels.inc.h:267-315 on the AVX2+FMA tier: lane c < 32 keeps
// slot c of the four 8-lane accumulators over the 32-element blocks (slot c
// = accumulator c / 8, lane c % 8), lanes 0-7 add the 8-wide cleanup into
// accumulator 0, the slots combine as (acc0 + acc1) + (acc2 + acc3) per
// lane and the lanes as ((s0+s4)+(s1+s5)) + ((s2+s6)+(s3+s7)), then a fused
// scalar tail; norm = sqrtf(total); a zero norm leaves the row unchanged;
// every element is divided by the norm (correctly rounded).
__global__ void __launch_bounds__(256) k_normalize(float *__restrict__ x, int64_t n, int dim) {
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    float *v = x + row * dim;
    const int lim32 = dim - dim % 32, lim8 = dim - dim % 8;
    float acc = 0.f;
    for (int i = 0; i < lim32; i += 32)
        if (lane < 32) acc = __builtin_fmaf(v[i + lane], v[i + lane], acc);
    for (int i = lim32; i < lim8; i += 8)
        if (lane < 8) acc = __builtin_fmaf(v[i + lane], v[i + lane], acc);
    const float a1 = __shfl(acc, (lane + 8) & 63), a2 = __shfl(acc, (lane + 16) & 63),
                a3 = __shfl(acc, (lane + 24) & 63);
    const float sl = (acc + a1) + (a2 + a3);             // lanes 0-7
    const float t = sl + __shfl(sl, (lane + 4) & 63);    // lanes 0-3: s_l + s_{l+4}
    const float u = t + __shfl(t, (lane + 1) & 63);      // lane 0: t0 + t1, lane 2: t2 + t3
    float total = __shfl(u, 0) + __shfl(u, 2);
    for (int i = lim8; i < dim; ++i) total = __builtin_fmaf(v[i], v[i], total);
    const float norm = __builtin_sqrtf(total);
    if (norm == 0.f) return;
    for (int i = lane; i < dim; i += 64) v[i] = v[i] / norm;
}

void launch_normalize(float *x, int64_t n, int dim, hipStream_t st) {
    if (n == 0 || dim == 0) return;
    hipLaunchKernelGGL(k_normalize, dim3((unsigned)ceil_div(n, 4)), dim3(256), 0, st, x, n, dim);
    HANN_LAUNCH_CHECK();
}

void launch_distance_batch(int metric, const float *queries, int64_t nq, const float *base,
                           int64_t nb, int dim, double *out, hipStream_t st) {
    if (nq == 0 || nb == 0) return;
    for (int64_t q0 = 0; q0 < nq; q0 += 65535) {
        int64_t qn = nq - q0 < 65535 ? nq - q0 : 65535;
        dim3 grid((unsigned)ceil_div(nb, 256), (unsigned)qn);
        hipLaunchKernelGGL(dev::k_distance_batch, grid, dim3(256), 0, st, metric, queries + q0 * dim,
                           qn, base, nb, dim, out + q0 * nb);
        HANN_LAUNCH_CHECK();
    }
}

int kp_for(int k) {
    int kp = 64;
    while (kp < k) kp <<= 1;
    return kp;
}

// the tiled kernel holds 16 queries' selection buffers; it covers the
// (squared) Euclidean metrics up to KP 256
static bool topk_tiled(int metric, int kp) {
    const int rank_metric = metric == HANN_METRIC_EUCLIDEAN ? HANN_METRIC_SQEUCLIDEAN : metric;
    return rank_metric == HANN_METRIC_SQEUCLIDEAN && kp <= 256;
}

// Row ranges of the exact top-k: enough that the grid holds a few thousand
// workgroups, at most max_split (>= 1) ranges of a multiple of 64 rows.
static void topk_plan(int metric, int kp, int64_t nq, int64_t nb, int64_t max_split, int64_t &nsplit,
                      int64_t &rows_per_split) {
    const int64_t qt = topk_tiled(metric, kp) ? dev::TK_QT : 1;
    const int64_t qtiles = ceil_div(nq, qt);
    nsplit = ceil_div(4096, qtiles);
    const int64_t fill = ceil_div(nb, 4 * 64);
    if (nsplit > fill) nsplit = fill;
    if (nsplit > max_split) nsplit = max_split;
    if (nsplit < 1) nsplit = 1;
    if (nsplit > 65535) nsplit = 65535;
    rows_per_split = ceil_div(ceil_div(nb, nsplit), 64) * 64;
    if (rows_per_split < 64) rows_per_split = 64;
    nsplit = ceil_div(nb, rows_per_split);
    if (nsplit < 1) nsplit = 1;
}

// the partial lists of every (query, row range), then their merge; pkey and
// pid hold nq * nsplit * k entries each
template <int KP>
static void topk_run(int metric, const float *queries, int64_t nq, const float *base, int64_t nb, int dim, int k,
                     int64_t nsplit, int64_t rows_per_split, float *pkey, int64_t *pid, int64_t *out_ids,
                     double *out_dist, hipStream_t st) {
    const int rank_metric = metric == HANN_METRIC_EUCLIDEAN ? HANN_METRIC_SQEUCLIDEAN : metric;
    const bool tiled = topk_tiled(metric, KP);
    const int64_t qt = tiled ? dev::TK_QT : 1;
    const int64_t qtiles = ceil_div(nq, qt);
    for (int64_t qt0 = 0; qt0 < qtiles; qt0 += 65535) {
        const int64_t qtn = qtiles - qt0 < 65535 ? qtiles - qt0 : 65535;
        const int64_t qb = qt0 * qt;
        const int64_t qn = nq - qb < qtn * qt ? nq - qb : qtn * qt;
        const dim3 grid((unsigned)qtn, (unsigned)nsplit);
        if constexpr (KP <= 256) {
            if (tiled) {
                hipLaunchKernelGGL(dev::k_topk_exact<KP>, grid, dim3(256), 0, st, rank_metric,
                                   queries + qb * dim, qn, base, nb, dim, k, rows_per_split,
                                   pkey + qb * nsplit * k, pid + qb * nsplit * k);
                HANN_LAUNCH_CHECK();
                continue;
            }
        }
        hipLaunchKernelGGL(dev::k_topk_generic<KP>, grid, dim3(64), 0, st, rank_metric,
                           queries + qb * dim, qn, base, nb, dim, k, rows_per_split,
                           pkey + qb * nsplit * k, pid + qb * nsplit * k);
        HANN_LAUNCH_CHECK();
    }
    for (int64_t q0 = 0; q0 < nq; q0 += 65535) {
        const int64_t qn = nq - q0 < 65535 ? nq - q0 : 65535;
        hipLaunchKernelGGL(dev::k_topk_merge<KP>, dim3((unsigned)qn), dim3(64), 0, st,
                           pkey + q0 * nsplit * k, pid + q0 * nsplit * k, (int)nsplit, k,
                           metric == HANN_METRIC_EUCLIDEAN ? 1 : 0, out_ids + q0 * k, out_dist + q0 * k);
        HANN_LAUNCH_CHECK();
    }
}

static void topk_dispatch(int metric, const float *queries, int64_t nq, const float *base, int64_t nb, int dim,
                          int k, int64_t nsplit, int64_t rows_per_split, float *pkey, int64_t *pid,
                          int64_t *out_ids, double *out_dist, hipStream_t st) {
    switch (kp_for(k)) {
    case 64: topk_run<64>(metric, queries, nq, base, nb, dim, k, nsplit, rows_per_split, pkey, pid, out_ids, out_dist, st); break;
    case 128: topk_run<128>(metric, queries, nq, base, nb, dim, k, nsplit, rows_per_split, pkey, pid, out_ids, out_dist, st); break;
    case 256: topk_run<256>(metric, queries, nq, base, nb, dim, k, nsplit, rows_per_split, pkey, pid, out_ids, out_dist, st); break;
    case 512: topk_run<512>(metric, queries, nq, base, nb, dim, k, nsplit, rows_per_split, pkey, pid, out_ids, out_dist, st); break;
    default:
        fail(HANN_ERR_UNSUPPORTED, strf("top-k with k=%d exceeds the supported maximum of 512", k));
    }
}

size_t topk_exact_scratch_bytes(int metric, int64_t nq, int64_t nb, int k) {
    const int kp = kp_for(k);
    if (nq == 0 || kp > 512) return 0;
    // dim is not known here, and a grow-only workspace keeps the slots of
    // both paths once batches of different shapes have taken each: both
    if (topk_rows_supported(metric, nb, 64, k))
        return topk_rows_scratch_bytes(nq, nb, k) + topk_exact_tiled_scratch_bytes(metric, nq, nb, k);
    return topk_exact_tiled_scratch_bytes(metric, nq, nb, k);
}

size_t topk_exact_tiled_scratch_bytes(int metric, int64_t nq, int64_t nb, int k) {
    const int kp = kp_for(k);
    if (nq == 0 || kp > 512) return 0;
    int64_t nsplit = 0, rows = 0;
    topk_plan(metric, kp, nq, nb, 65535, nsplit, rows);
    return (size_t)(nq * nsplit * k) * 12 + 512;
}

void launch_topk_exact(int metric, const float *queries, int64_t nq, const float *base, int64_t nb,
                       int dim, int k, int64_t *out_ids, double *out_dist, Workspace &ws,
                       hipStream_t st) {
    if (nq == 0) return;
    const int kp = kp_for(k);
    if (kp > 512) fail(HANN_ERR_UNSUPPORTED, strf("top-k with k=%d exceeds the supported maximum of 512", k));
    if (topk_rows_supported(metric, nb, dim, k)) {
        // sample, bound, filter and select over rows held in registers; past
        // 65,536 rows its overflow check waits on the host, which a stream
        // under capture cannot do: the tiled kernel then
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (nb > 65536) HANN_HIP(hipStreamIsCapturing(st, &cs));
        if (cs == hipStreamCaptureStatusNone) {
            launch_topk_rows(metric, queries, nq, base, nb, dim, k, out_ids, out_dist, ws, st);
            return;
        }
    }
    int64_t nsplit = 0, rows = 0;
    topk_plan(metric, kp, nq, nb, 65535, nsplit, rows);
    float *pkey = ws.get<float>(14, (size_t)(nq * nsplit * k));
    int64_t *pid = ws.get<int64_t>(15, (size_t)(nq * nsplit * k));
    topk_dispatch(metric, queries, nq, base, nb, dim, k, nsplit, rows, pkey, pid, out_ids, out_dist, st);
}

void topk_tiled_redo(int metric, const float *queries, const std::vector<int64_t> &redo, const float *base,
                     int64_t nb, int dim, int k, int64_t *out_ids, double *out_dist, Workspace &ws, hipStream_t st) {
    (void)ws;
    const int64_t nr = (int64_t)redo.size();
    if (nr == 0) return;
    const int kp = kp_for(k);
    DevBuf<int64_t> dredo((size_t)nr);
    DevBuf<float> rq((size_t)(nr * dim));
    DevBuf<int64_t> rid((size_t)(nr * k));
    DevBuf<double> rdist((size_t)(nr * k));
    HANN_HIP(hipMemcpyAsync(dredo.get(), redo.data(), (size_t)nr * sizeof(int64_t), hipMemcpyHostToDevice, st));
    launch_gather_rows(queries, dim, dredo.get(), nr, rq.get(), st);
    int64_t nsplit = 0, rows = 0;
    topk_plan(metric, kp, nr, nb, 65535, nsplit, rows);
    DevBuf<float> pkey((size_t)(nr * nsplit * k));
    DevBuf<int64_t> pid((size_t)(nr * nsplit * k));
    topk_dispatch(metric, rq.get(), nr, base, nb, dim, k, nsplit, rows, pkey.get(), pid.get(), rid.get(), rdist.get(),
                  st);
    for (int64_t i = 0; i < nr; ++i) {
        HANN_HIP(hipMemcpyAsync(out_ids + redo[i] * k, rid.get() + i * k, (size_t)k * sizeof(int64_t),
                                hipMemcpyDeviceToDevice, st));
        HANN_HIP(hipMemcpyAsync(out_dist + redo[i] * k, rdist.get() + i * k, (size_t)k * sizeof(double),
                                hipMemcpyDeviceToDevice, st));
    }
    // the local buffers are freed on return
    HANN_HIP(hipStreamSynchronize(st));
}

// ------------------------------------------------- allocation-free top-k --
// The partial lists of hann_batch_topk_noalloc_dev live in this array: one
// copy per GPU, held until the process ends, shared by every call on that
// GPU (calls that overlap on one GPU race on it).
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
static void *g_scratch_dev[kMaxDevices];   // scratch's device address on each GPU

// the device address of `scratch` on the current device, looked up once per
// device with that device current
void *noalloc_scratch_current() {
    int dev = -1;
    HANN_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) fail(HANN_ERR_DEVICE, strf("device %d is out of the supported range", dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            fail(HANN_ERR_DEVICE, strf("hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                                       hipGetErrorString(e)));
        g_scratch_dev[dev] = ptr;
    }
    return g_scratch_dev[dev];
}

void launch_topk_exact_noalloc(int metric, const float *queries, int64_t nq, const float *base, int64_t nb,
                               int dim, int k, int64_t *out_ids, double *out_dist, hipStream_t st) {
    if (nq == 0) return;
    const int kp = kp_for(k);
    if (kp > 512) fail(HANN_ERR_UNSUPPORTED, strf("top-k with k=%d exceeds the supported maximum of 512", k));
    // 12 bytes per partial entry (a float key and an int64 row)
    const int64_t per_split = nq * (int64_t)k * 12;
    const int64_t max_split = (int64_t)kNoAllocScratchBytes / per_split;
    if (max_split < 1)
        fail(HANN_ERR_INVALID_ARG,
             strf("allocation-free top-k: nq * k * 12 = %lld bytes exceeds its scratch of %zu bytes",
                  (long long)per_split, kNoAllocScratchBytes));
    int64_t nsplit = 0, rows = 0;
    topk_plan(metric, kp, nq, nb, max_split, nsplit, rows);
    char *base_ptr = static_cast<char *>(noalloc_scratch_current());
    int64_t *pid = reinterpret_cast<int64_t *>(base_ptr);
    float *pkey = reinterpret_cast<float *>(base_ptr + nq * nsplit * k * 8);
    topk_dispatch(metric, queries, nq, base, nb, dim, k, nsplit, rows, pkey, pid, out_ids, out_dist, st);
}

}  // namespace hann
