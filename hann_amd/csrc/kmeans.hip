// kmeans.hip -- k-means assignment and update, PQ encode (gfx950).
//
// Assignment reproduces assignPoints (pqivf/index.go:738-789): the exact
// AVX2-order squared distance to every centroid and an argmin in which the
// lowest centroid index wins a tie. The update reproduces runKMeans'
// float32 sums in point order (pqivf/index.go:816-844) without atomics on
// the sums: members are grouped per cluster in ascending point order, and one
// thread sums one (cluster, dimension) sequentially.
#include "common.h"
#include "kmeans.h"

namespace hann {
namespace dev {

// ----------------------------------------------------------- assignment --
// Tiled kernel for d >= 32: a workgroup of 4 waves owns 16 points (4 per
// wave); centroids are staged through LDS 64 at a time, one 32-dim block at
// a time, lane l owns centroid l of the tile. Each lane keeps a running
// (distance, index) minimum per point; a wave reduction picks the winner.
constexpr int AS_WAVES = 4;
constexpr int AS_PPW = 4;
constexpr int AS_PT = AS_WAVES * AS_PPW;
constexpr int AS_ROWS = 64;
constexpr int AS_STRIDE = 36;

__device__ __forceinline__ void argmin_merge(float &bd, int &bi, float od, int oi) {
    if (od < bd || (od == bd && oi < bi)) { bd = od; bi = oi; }
}

__global__ void __launch_bounds__(256) k_assign_tiled(const float *__restrict__ data, int64_t n,
                                                      int64_t row_stride, const float *__restrict__ cents,
                                                      int k, int d, int32_t *__restrict__ out,
                                                      float *__restrict__ out_dist) {
    __shared__ float s_c[AS_ROWS * AS_STRIDE];
    __shared__ float s_p[AS_PT * 32];
    const int tid = threadIdx.x;
    const int wave = tid / kWave;
    const int lane = tid % kWave;
    const int64_t p0 = (int64_t)blockIdx.x * AS_PT;
    const int nblk = d / 32;

    float best_d[AS_PPW];
    int best_i[AS_PPW];
#pragma unroll
    for (int j = 0; j < AS_PPW; ++j) { best_d[j] = __builtin_inff(); best_i[j] = INT32_MAX; }

    for (int t0 = 0; t0 < k; t0 += AS_ROWS) {
        const int c = t0 + lane;
        const bool valid = c < k;
        float acc[AS_PPW][32];
#pragma unroll
        for (int j = 0; j < AS_PPW; ++j)
#pragma unroll
            for (int s = 0; s < 32; ++s) acc[j][s] = 0.f;
        for (int b = 0; b < nblk; ++b) {
            __syncthreads();
            for (int v = tid; v < AS_ROWS * 8; v += 256) {
                int r = v / 8, c4 = v % 8;
                int gc = t0 + r;
                float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
                if (gc < k) {
                    const float *src = cents + (int64_t)gc * d + b * 32 + c4 * 4;
                    x = make_float4(src[0], src[1], src[2], src[3]);
                }
                *reinterpret_cast<float4 *>(&s_c[r * AS_STRIDE + c4 * 4]) = x;
            }
            for (int v = tid; v < AS_PT * 32; v += 256) {
                int pl = v / 32, e = v % 32;
                int64_t gp = p0 + pl;
                s_p[v] = gp < n ? data[gp * row_stride + b * 32 + e] : 0.f;
            }
            __syncthreads();
#pragma unroll
            for (int c4 = 0; c4 < 8; ++c4) {
                float4 y = *reinterpret_cast<const float4 *>(&s_c[lane * AS_STRIDE + c4 * 4]);
#pragma unroll
                for (int j = 0; j < AS_PPW; ++j) {
                    float4 x = *reinterpret_cast<const float4 *>(&s_p[(wave * AS_PPW + j) * 32 + c4 * 4]);
                    // d = point - centroid: RankBatch(point, centroids), pqivf/index.go:768
                    float d0 = x.x - y.x, d1 = x.y - y.y, d2 = x.z - y.z, d3 = x.w - y.w;
                    acc[j][c4 * 4 + 0] = __builtin_fmaf(d0, d0, acc[j][c4 * 4 + 0]);
                    acc[j][c4 * 4 + 1] = __builtin_fmaf(d1, d1, acc[j][c4 * 4 + 1]);
                    acc[j][c4 * 4 + 2] = __builtin_fmaf(d2, d2, acc[j][c4 * 4 + 2]);
                    acc[j][c4 * 4 + 3] = __builtin_fmaf(d3, d3, acc[j][c4 * 4 + 3]);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < AS_PPW; ++j) {
            int64_t gp = p0 + wave * AS_PPW + j;
            const float *xv = data + (gp < n ? gp : 0) * row_stride;
            const float *cv = cents + (int64_t)(valid ? c : 0) * d;
            int i = nblk * 32;
            const int lim = d - (d % 8);
            for (; i < lim; i += 8) {
#pragma unroll
                for (int l = 0; l < 8; ++l) {
                    float dd = xv[i + l] - cv[i + l];
                    acc[j][l] = __builtin_fmaf(dd, dd, acc[j][l]);
                }
            }
            float s[8];
#pragma unroll
            for (int l = 0; l < 8; ++l) s[l] = (acc[j][l] + acc[j][8 + l]) + (acc[j][16 + l] + acc[j][24 + l]);
            float sum = hsum8(s);
            for (; i < d; ++i) {
                float dd = xv[i] - cv[i];
                sum = __builtin_fmaf(dd, dd, sum);
            }
            if (valid) argmin_merge(best_d[j], best_i[j], sane_key(sum), c);
        }
    }
#pragma unroll
    for (int j = 0; j < AS_PPW; ++j) {
        float bd = best_d[j];
        int bi = best_i[j];
        for (int off = 32; off > 0; off >>= 1) {
            float od = __shfl_xor(bd, off);
            int oi = __shfl_xor(bi, off);
            argmin_merge(bd, bi, od, oi);
        }
        int64_t gp = p0 + wave * AS_PPW + j;
        if (lane == 0 && gp < n) {
            out[gp] = bi;
            if (out_dist) out_dist[gp] = bd;
        }
    }
}

// Small-dimension kernel (d < 32, the PQ sub-spaces; at d = 32 the
// reference switches to its 4x8 block order, which the tiled kernel covers): one thread per point,
// centroids of group g staged through LDS. Point data lives in registers.
// Batched over `groups` independent problems: group g reads columns
// [g*d, (g+1)*d) of each data row and its own centroid table.
constexpr int SM_MAXD = 32;

__global__ void __launch_bounds__(256) k_assign_small(const float *__restrict__ data, int64_t n,
                                                      int64_t row_stride, const float *__restrict__ cents,
                                                      int64_t cent_group_stride, const int *__restrict__ kg,
                                                      int k_single, int d, int32_t *__restrict__ out,
                                                      int64_t out_group_stride) {
    extern __shared__ float s_cb[];   // [chunk][d]
    const int g = blockIdx.y;
    const int k = kg ? kg[g] : k_single;
    const float *cg = cents + g * cent_group_stride;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float x[SM_MAXD];
#pragma unroll
    for (int e = 0; e < SM_MAXD; ++e) x[e] = 0.f;
    if (p < n) {
        const float *src = data + p * row_stride + (int64_t)g * d;
#pragma unroll
        for (int e = 0; e < SM_MAXD; ++e)
            if (e < d) x[e] = src[e];
    }
    float bd = __builtin_inff();
    int bi = 0;
    const int chunk = 2048 / (d > 0 ? d : 1);
    for (int c0 = 0; c0 < k; c0 += chunk) {
        const int cn = k - c0 < chunk ? k - c0 : chunk;
        __syncthreads();
        for (int v = threadIdx.x; v < cn * d; v += blockDim.x) s_cb[v] = cg[(int64_t)c0 * d + v];
        __syncthreads();
        for (int c = 0; c < cn; ++c) {
            const float *cv = s_cb + c * d;
            // exact AVX2 order for d < 32: 8-wide cleanup blocks, hsum, tail
            float acc[8];
#pragma unroll
            for (int l = 0; l < 8; ++l) acc[l] = 0.f;
            const int lim = d - (d % 8);
#pragma unroll
            for (int i = 0; i < SM_MAXD; i += 8) {
                if (i < lim) {
#pragma unroll
                    for (int l = 0; l < 8; ++l) {
                        float dd = x[i + l] - cv[i + l];
                        acc[l] = __builtin_fmaf(dd, dd, acc[l]);
                    }
                }
            }
            float sum = hsum8(acc);
#pragma unroll
            for (int i = 0; i < SM_MAXD; ++i) {
                if (i >= lim && i < d) {
                    float dd = x[i] - cv[i];
                    sum = __builtin_fmaf(dd, dd, sum);
                }
            }
            sum = sane_key(sum);
            if (sum < bd) { bd = sum; bi = c0 + c; }   // strict <: lowest index wins
        }
    }
    if (p < n) out[g * out_group_stride + p] = bi;
}

// ---------------------------------------------------------------- update --
// Stable grouping of points by cluster, so that the sums below visit each
// cluster's members in ascending point order. Points are cut into blocks of
// PB; hist[(g*k + c)*nb + b] counts block b's members of cluster c; an
// exclusive scan of hist in that (group, cluster, block) order gives every
// (cluster, block) its first slot; inside a block, a bitonic sort of the
// composite keys c*PB + local ranks the members of each cluster in point
// order. members[g*n + slot] = point.
constexpr int PB = 2048;

__global__ void k_block_hist(const int32_t *__restrict__ assign, int64_t n, int64_t assign_group_stride,
                             int k, int64_t nblocks, int32_t *__restrict__ hist) {
    const int g = blockIdx.y;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = assign[g * assign_group_stride + i];
    atomicAdd(&hist[((int64_t)g * k + c) * nblocks + i / PB], 1);
}

// three-phase exclusive scan of int32 values (n < 2^31 in total)
constexpr int SC_TILE = 8192;

__global__ void __launch_bounds__(1024) k_scan_tiles(const int32_t *__restrict__ in, int64_t len,
                                                     int32_t *__restrict__ tile_sums) {
    __shared__ int32_t s_red[1024];
    const int64_t base = (int64_t)blockIdx.x * SC_TILE;
    int32_t local = 0;
    for (int j = 0; j < SC_TILE / 1024; ++j) {
        int64_t i = base + (int64_t)j * 1024 + threadIdx.x;
        if (i < len) local += in[i];
    }
    s_red[threadIdx.x] = local;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) s_red[threadIdx.x] += s_red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = s_red[0];
}

// exclusive scan of up to 1024*64 tile sums in one workgroup
__global__ void __launch_bounds__(1024) k_scan_sums(int32_t *__restrict__ sums, int64_t ntiles) {
    __shared__ int32_t s_part[1024];
    const int t = threadIdx.x;
    const int64_t per = (ntiles + 1023) / 1024;
    const int64_t b = t * per, e = b + per < ntiles ? b + per : ntiles;
    int32_t local = 0;
    for (int64_t i = b; i < e; ++i) local += sums[i];
    s_part[t] = local;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int32_t v = t >= off ? s_part[t - off] : 0;
        __syncthreads();
        s_part[t] += v;
        __syncthreads();
    }
    int32_t run = s_part[t] - local;
    for (int64_t i = b; i < e; ++i) {
        int32_t v = sums[i];
        sums[i] = run;
        run += v;
    }
}

__global__ void __launch_bounds__(1024) k_scan_apply(const int32_t *__restrict__ in, int64_t len,
                                                     const int32_t *__restrict__ tile_pre,
                                                     int32_t *__restrict__ out) {
    __shared__ int32_t s_part[1024];
    const int64_t base = (int64_t)blockIdx.x * SC_TILE;
    constexpr int PER = SC_TILE / 1024;
    int32_t v[PER];
    int32_t local = 0;
    const int t = threadIdx.x;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        int64_t i = base + (int64_t)t * PER + j;
        v[j] = i < len ? in[i] : 0;
        local += v[j];
    }
    s_part[t] = local;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int32_t x = t >= off ? s_part[t - off] : 0;
        __syncthreads();
        s_part[t] += x;
        __syncthreads();
    }
    int32_t run = tile_pre[blockIdx.x] + s_part[t] - local;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        int64_t i = base + (int64_t)t * PER + j;
        if (i < len) out[i] = run;
        run += v[j];
    }
}

__global__ void __launch_bounds__(256) k_block_scatter(const int32_t *__restrict__ assign, int64_t n,
                                                       int64_t assign_group_stride, int k, int64_t nblocks,
                                                       const int32_t *__restrict__ pos,
                                                       int32_t *__restrict__ members) {
    __shared__ uint32_t s_key[PB];
    const int g = blockIdx.y;
    const int64_t b = blockIdx.x;
    const int64_t start = b * PB;
    const int cnt = n - start < PB ? (int)(n - start) : PB;
    const int32_t *ag = assign + g * assign_group_stride + start;
    for (int i = threadIdx.x; i < PB; i += blockDim.x)
        s_key[i] = i < cnt ? (uint32_t)ag[i] * (uint32_t)PB + (uint32_t)i : 0xFFFFFFFFu;
    __syncthreads();
    for (int size = 2; size <= PB; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = threadIdx.x; t < PB / 2; t += blockDim.x) {
                int lo = 2 * t - (t & (stride - 1));
                int hi = lo + stride;
                bool up = (lo & size) == 0;
                uint32_t x = s_key[lo], y = s_key[hi];
                if (up ? (y < x) : (x < y)) { s_key[lo] = y; s_key[hi] = x; }
            }
            __syncthreads();
        }
    for (int p = threadIdx.x; p < cnt; p += blockDim.x) {
        const uint32_t key = s_key[p];
        const uint32_t c = key / PB, local = key % PB;
        // first sorted slot holding cluster c
        int lo = 0, hi = p;
        const uint32_t first = c * (uint32_t)PB;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (s_key[mid] < first) lo = mid + 1; else hi = mid;
        }
        const int rank = p - lo;
        const int32_t slot = pos[((int64_t)g * k + c) * nblocks + b] + rank;
        members[slot] = (int32_t)(start + local);
    }
}

// offs[g*(k+1) + c] = first member slot of cluster c in group g (relative to g*n)
__global__ void k_offsets(const int32_t *__restrict__ pos, int64_t n, int k, int64_t nblocks,
                          int64_t *__restrict__ offs) {
    const int g = blockIdx.y;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > k) return;
    offs[(int64_t)g * (k + 1) + c] =
        c == k ? n : (int64_t)pos[((int64_t)g * k + c) * nblocks] - (int64_t)g * n;
}

// New centroids: thread (g, c, e) sums dimension e of cluster c's members in
// point order in float32, then divides by the count in float32
// (pqivf/index.go:825-827, 839-841). Empty clusters copy the data row the
// host drew for them (reseed_idx >= 0) -- runKMeans' random reseed.
__global__ void k_update(const float *__restrict__ data, int64_t n, int64_t row_stride, int d,
                         const int32_t *__restrict__ members, const int64_t *__restrict__ offs, int k,
                         const int64_t *__restrict__ reseed_idx, float *__restrict__ cents,
                         int64_t cent_group_stride) {
    const int g = blockIdx.y;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)k * d) return;
    const int c = (int)(t / d), e = (int)(t % d);
    const int64_t b = offs[(int64_t)g * (k + 1) + c], en = offs[(int64_t)g * (k + 1) + c + 1];
    const float *col = data + (int64_t)g * d + e;
    float *dst = cents + g * cent_group_stride + (int64_t)c * d + e;
    if (en == b) {
        int64_t r = reseed_idx[(int64_t)g * k + c];
        if (r >= 0) *dst = col[r * row_stride];
        return;
    }
    const int32_t *mem = members + (int64_t)g * n;
    float sum = 0.f;
    for (int64_t i = b; i < en; ++i) sum = sum + col[(int64_t)mem[i] * row_stride];
    *dst = sum / (float)(en - b);
}

// --------------------------------------------------------------- encode --
// codes[i][g] = argmin_j sqdiff(r_g, cb[g][j]) with r = x - c (float32),
// lowest j on a tie (pqivf/index.go:641-666). One thread per (row, group).
__global__ void __launch_bounds__(256) k_encode(const float *__restrict__ vecs, int64_t n, int d,
                                                const float *__restrict__ coarse,
                                                const int32_t *__restrict__ assign, int m,
                                                const float *__restrict__ cb, int pmax,
                                                const int64_t *__restrict__ cb_len,
                                                int32_t *__restrict__ codes) {
    extern __shared__ float s_cb[];   // [pmax][s] of group g, chunked
    const int g = blockIdx.y;
    const int s = d / m;
    const int P = (int)cb_len[g];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float r[SM_MAXD];
#pragma unroll
    for (int e = 0; e < SM_MAXD; ++e) r[e] = 0.f;
    if (i < n) {
        const float *x = vecs + i * d + g * s;
        const float *c = coarse + (int64_t)assign[i] * d + g * s;
#pragma unroll
        for (int e = 0; e < SM_MAXD; ++e)
            if (e < s) r[e] = x[e] - c[e];   // vectorSub, pqivf/index.go:698-700
    }
    float bd = __builtin_inff();
    int bi = 0;
    const int chunk = 2048 / s;
    const float *cbg = cb + (int64_t)g * pmax * s;
    for (int c0 = 0; c0 < P; c0 += chunk) {
        const int cn = P - c0 < chunk ? P - c0 : chunk;
        __syncthreads();
        for (int v = threadIdx.x; v < cn * s; v += blockDim.x) s_cb[v] = cbg[(int64_t)c0 * s + v];
        __syncthreads();
        for (int c = 0; c < cn; ++c) {
            const float *cv = s_cb + c * s;
            float acc[8];
#pragma unroll
            for (int l = 0; l < 8; ++l) acc[l] = 0.f;
            const int lim = s - (s % 8);
#pragma unroll
            for (int q = 0; q < SM_MAXD; q += 8) {
                if (q < lim) {
#pragma unroll
                    for (int l = 0; l < 8; ++l) {
                        float dd = r[q + l] - cv[q + l];
                        acc[l] = __builtin_fmaf(dd, dd, acc[l]);
                    }
                }
            }
            float sum = hsum8(acc);
#pragma unroll
            for (int q = 0; q < SM_MAXD; ++q) {
                if (q >= lim && q < s) {
                    float dd = r[q] - cv[q];
                    sum = __builtin_fmaf(dd, dd, sum);
                }
            }
            sum = sane_key(sum);
            if (sum < bd) { bd = sum; bi = c0 + c; }
        }
    }
    if (i < n) codes[i * m + g] = bi;
}

// Encode for sub-spaces of 32 or more dims: one thread per (row, group),
// exact order through the generic device function.
__global__ void __launch_bounds__(128) k_encode_wide(const float *__restrict__ vecs, int64_t n, int d,
                                                     const float *__restrict__ coarse,
                                                     const int32_t *__restrict__ assign, int m,
                                                     const float *__restrict__ cb, int pmax,
                                                     const int64_t *__restrict__ cb_len,
                                                     float *__restrict__ resid, int32_t *__restrict__ codes) {
    const int g = blockIdx.y;
    const int s = d / m;
    const int P = (int)cb_len[g];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float *r = resid + i * d + g * s;
    const float *x = vecs + i * d + g * s;
    const float *c = coarse + (int64_t)assign[i] * d + g * s;
    for (int e = 0; e < s; ++e) r[e] = x[e] - c[e];
    const float *cbg = cb + (int64_t)g * pmax * s;
    float bd = __builtin_inff();
    int bi = 0;
    for (int j = 0; j < P; ++j) {
        float v = sane_key(sqdiff_exact(PtrLoad{r}, PtrLoad{cbg + (int64_t)j * s}, s));
        if (v < bd) { bd = v; bi = j; }
    }
    codes[i * m + g] = bi;
}

// residual rows for PQ training, in the given order:
// out[r][e] = vecs[order[r]][e] - coarse[assign[order[r]]][e]
__global__ void k_residuals(const float *__restrict__ vecs, int d, const float *__restrict__ coarse,
                            const int32_t *__restrict__ assign, const int32_t *__restrict__ order, int64_t n,
                            float *__restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * d) return;
    const int64_t r = t / d;
    const int e = (int)(t % d);
    const int64_t p = order[r];
    out[t] = vecs[p * d + e] - coarse[(int64_t)assign[p] * d + e];
}

// k-means seeds of `groups` problems: cents[g][c][e] = data[idx[g][c]][g*d + e]
__global__ void k_gather_group(const float *__restrict__ data, int64_t row_stride, int d, int k,
                               const int64_t *__restrict__ idx, float *__restrict__ cents) {
    const int g = blockIdx.y;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)k * d) return;
    const int c = (int)(t / d), e = (int)(t % d);
    cents[((int64_t)g * k + c) * d + e] = data[idx[(int64_t)g * k + c] * row_stride + (int64_t)g * d + e];
}

// gather rows: out[r] = src[idx[r]]
__global__ void k_gather_rows(const float *__restrict__ src, int d, const int64_t *__restrict__ idx,
                              int64_t n, float *__restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * d) return;
    const int64_t r = t / d;
    out[t] = src[idx[r] * d + (t % d)];
}

}  // namespace dev

// ------------------------------------------------------------ launchers --
void launch_assign(const float *data, int64_t n, int64_t row_stride, const float *cents, int k, int d,
                   int32_t *out, hipStream_t st) {
    if (n == 0) return;
    if (d >= 32) {
        int64_t blocks = ceil_div(n, dev::AS_PT);
        for (int64_t b0 = 0; b0 < blocks; b0 += (1u << 30)) {
            int64_t bn = blocks - b0 < (1 << 30) ? blocks - b0 : (1 << 30);
            hipLaunchKernelGGL(dev::k_assign_tiled, dim3((unsigned)bn), dim3(256), 0, st,
                               data + b0 * dev::AS_PT * row_stride, n - b0 * dev::AS_PT, row_stride, cents, k,
                               d, out + b0 * dev::AS_PT, (float *)nullptr);
            HANN_LAUNCH_CHECK();
        }
    } else {
        launch_assign_groups(data, n, row_stride, cents, 0, nullptr, k, d, 1, out, 0, st);
    }
}

void launch_assign_groups(const float *data, int64_t n, int64_t row_stride, const float *cents,
                          int64_t cent_group_stride, const int *kg_dev, int k_single, int d, int groups,
                          int32_t *out, int64_t out_group_stride, hipStream_t st) {
    if (n == 0) return;
    if (d >= 32) fail(HANN_ERR_UNSUPPORTED, strf("grouped small assignment needs d < 32, got %d", d));
    const int chunk = 2048 / (d > 0 ? d : 1);
    size_t shm = (size_t)chunk * d * sizeof(float);
    hipLaunchKernelGGL(dev::k_assign_small, dim3((unsigned)ceil_div(n, 256), (unsigned)groups), dim3(256), shm, st,
                       data, n, row_stride, cents, cent_group_stride, kg_dev, k_single, d, out, out_group_stride);
    HANN_LAUNCH_CHECK();
}

int64_t kmeans_scratch_hist_len(int64_t n, int k, int groups) {
    return (int64_t)groups * k * ceil_div(n > 0 ? n : 1, dev::PB);
}

void launch_group_members(const int32_t *assign, int64_t n, int64_t assign_group_stride, int k, int groups,
                          KMeansScratch &sc, hipStream_t st) {
    const int64_t nblocks = ceil_div(n, dev::PB);
    const int64_t len = (int64_t)groups * k * nblocks;
    const int64_t ntiles = ceil_div(len, dev::SC_TILE);
    if (ntiles > 1024 * 64) fail(HANN_ERR_UNSUPPORTED, "k-means grouping table too large");
    int32_t *hist = sc.hist.get();
    HANN_HIP(hipMemsetAsync(hist, 0, sizeof(int32_t) * (size_t)len, st));
    hipLaunchKernelGGL(dev::k_block_hist, dim3((unsigned)ceil_div(n, 256), (unsigned)groups), dim3(256), 0, st,
                       assign, n, assign_group_stride, k, nblocks, hist);
    HANN_LAUNCH_CHECK();
    hipLaunchKernelGGL(dev::k_scan_tiles, dim3((unsigned)ntiles), dim3(1024), 0, st, hist, len, sc.tile_sums.get());
    HANN_LAUNCH_CHECK();
    hipLaunchKernelGGL(dev::k_scan_sums, dim3(1), dim3(1024), 0, st, sc.tile_sums.get(), ntiles);
    HANN_LAUNCH_CHECK();
    hipLaunchKernelGGL(dev::k_scan_apply, dim3((unsigned)ntiles), dim3(1024), 0, st, hist, len, sc.tile_sums.get(),
                       sc.pos.get());
    HANN_LAUNCH_CHECK();
    hipLaunchKernelGGL(dev::k_block_scatter, dim3((unsigned)nblocks, (unsigned)groups), dim3(256), 0, st, assign,
                       n, assign_group_stride, k, nblocks, sc.pos.get(), sc.members.get());
    HANN_LAUNCH_CHECK();
    hipLaunchKernelGGL(dev::k_offsets, dim3((unsigned)ceil_div(k + 1, 256), (unsigned)groups), dim3(256), 0, st,
                       sc.pos.get(), n, k, nblocks, sc.offs.get());
    HANN_LAUNCH_CHECK();
}

void KMeansScratch::reserve(int64_t n, int k, int groups) {
    const int64_t len = kmeans_scratch_hist_len(n, k, groups);
    hist.resize((size_t)len + 1);
    pos.resize((size_t)len + 1);
    tile_sums.resize((size_t)ceil_div(len, dev::SC_TILE) + 1);
    members.resize((size_t)(n * groups) + 1);
    offs.resize((size_t)groups * (k + 1));
    reseed.resize((size_t)groups * k);
}

void launch_kmeans_sums(const float *data, int64_t n, int64_t row_stride, int d, int k, int groups, float *cents,
                        int64_t cent_group_stride, KMeansScratch &sc, hipStream_t st) {
    hipLaunchKernelGGL(dev::k_update, dim3((unsigned)ceil_div((int64_t)k * d, 256), (unsigned)groups), dim3(256),
                       0, st, data, n, row_stride, d, sc.members.get(), sc.offs.get(), k, sc.reseed.get(), cents,
                       cent_group_stride);
    HANN_LAUNCH_CHECK();
}

void launch_encode(const float *vecs, int64_t n, int d, const float *coarse, const int32_t *assign, int m,
                   const float *cb, int pmax, const int64_t *cb_len_dev, int32_t *codes, float *resid_scratch,
                   hipStream_t st) {
    if (n == 0) return;
    const int s = d / m;
    if (s < 32) {
        const int chunk = 2048 / s;
        size_t shm = (size_t)chunk * s * sizeof(float);
        hipLaunchKernelGGL(dev::k_encode, dim3((unsigned)ceil_div(n, 256), (unsigned)m), dim3(256), shm, st, vecs, n,
                           d, coarse, assign, m, cb, pmax, cb_len_dev, codes);
    } else {
        hipLaunchKernelGGL(dev::k_encode_wide, dim3((unsigned)ceil_div(n, 128), (unsigned)m), dim3(128), 0, st,
                           vecs, n, d, coarse, assign, m, cb, pmax, cb_len_dev, resid_scratch, codes);
    }
    HANN_LAUNCH_CHECK();
}

void launch_gather_group(const float *data, int64_t row_stride, int d, int k, int groups, const int64_t *idx,
                         float *cents, hipStream_t st) {
    if (k == 0) return;
    hipLaunchKernelGGL(dev::k_gather_group, dim3((unsigned)ceil_div((int64_t)k * d, 256), (unsigned)groups), dim3(256),
                       0, st, data, row_stride, d, k, idx, cents);
    HANN_LAUNCH_CHECK();
}

void launch_residuals(const float *vecs, int d, const float *coarse, const int32_t *assign, const int32_t *order,
                      int64_t n, float *out, hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL(dev::k_residuals, dim3((unsigned)ceil_div(n * d, 256)), dim3(256), 0, st, vecs, d, coarse,
                       assign, order, n, out);
    HANN_LAUNCH_CHECK();
}

void launch_gather_rows(const float *src, int d, const int64_t *idx, int64_t n, float *out, hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL(dev::k_gather_rows, dim3((unsigned)ceil_div(n * d, 256)), dim3(256), 0, st, src, d, idx, n,
                       out);
    HANN_LAUNCH_CHECK();
}


// ---------------------------------------------------------- k-means++ --
namespace dev {
constexpr int PP_BLOCK = 256;
constexpr int PP_BS = 4;   // block sums per pick thread held in LDS (32 KB)

// D[i] = min(D[i], e(x_i, c)) in the reference FP32 order, A[i] = the index
// of the center that gave D[i] (first: D[i] = e(x_i, c_0), A[i] = 0); per
// block the FP64 tree sum of D, rewritten only where some D changed. A point
// is skipped when |c_A - c|^2 >= 4 D margin (cd[A], FP64): by the triangle
// inequality its distance to c is then at least sqrt(D) with room for the
// FP32 rounding of both distances, so min(D, e) = D. The skip changes no
// value, only which distances are computed.
// Groups (blockIdx.y; the subspaces of PQ training) run side by side: group
// g's rows start at data + g d, its centers at cents + g k d, and its D, A,
// block sums and center distances at g n, g n, g nblocks and g k.
template <bool FIRST>
__global__ void __launch_bounds__(PP_BLOCK) k_pp_update(const float *__restrict__ data, int64_t n, int64_t row_stride,
                                                        int d, const float *__restrict__ cents, int k, int ci,
                                                        const double *__restrict__ cd_all, double margin,
                                                        float *__restrict__ D_all, int32_t *__restrict__ A_all,
                                                        double *__restrict__ bsum_all) {
    __shared__ double s_sum[PP_BLOCK];
    const int64_t g = blockIdx.y;
    data += g * d;
    const float *c = cents + (g * k + ci) * d;
    const double *cd = cd_all + g * k;
    float *D = D_all + g * n;
    int32_t *A = A_all + g * n;
    double *bsum = bsum_all + g * (int64_t)gridDim.x;
    const int64_t i = (int64_t)blockIdx.x * PP_BLOCK + threadIdx.x;
    double v = 0.0;
    int changed = 0;
    if (i < n) {
        if (FIRST) {
            const float e = sane_key(sqdiff_exact(PtrLoad{data + i * row_stride}, PtrLoad{c}, d));
            D[i] = e;
            A[i] = ci;
            v = e;
            changed = 1;
        } else {
            float dv = D[i];
            if (!(cd[A[i]] >= 4.0 * (double)dv * margin)) {
                const float e = sane_key(sqdiff_exact(PtrLoad{data + i * row_stride}, PtrLoad{c}, d));
                if (e < dv) {
                    dv = e;
                    D[i] = e;
                    A[i] = ci;
                    changed = 1;
                }
            }
            v = dv;
        }
    }
    if (!__syncthreads_or(changed)) return;
    s_sum[threadIdx.x] = v;
    __syncthreads();
    for (int o = PP_BLOCK / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s_sum[threadIdx.x] += s_sum[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = s_sum[0];
}

// cd[i] = |c_i - c_j|^2 in FP64 for the centers i < j (one wave each; group
// blockIdx.y)
__global__ void __launch_bounds__(256) k_pp_center_dist(const float *__restrict__ cents_all, int d, int k, int j,
                                                        double *__restrict__ cd_all) {
    const int w = blockIdx.x * 4 + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (w >= j) return;
    const float *cents = cents_all + (int64_t)blockIdx.y * k * d;
    double *cd = cd_all + (int64_t)blockIdx.y * k;
    const float *a = cents + (int64_t)w * d, *b = cents + (int64_t)j * d;
    double s = 0.0;
    for (int e = lane; e < d; e += 64) {
        const double t = (double)a[e] - (double)b[e];
        s += t * t;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) cd[w] = s;
}

// one workgroup: the row whose D-prefix first exceeds u * total; copy it to
// the next center slot. The chunk that holds the target is walked block by
// block (block sums, then, in the block that crosses, its rows one by one
// in order); the rows of that block are first staged in LDS by the whole
// workgroup, so the walk reads no global memory.
// (one workgroup per group: blockIdx.x; u holds k - 1 draws per group)
__global__ void __launch_bounds__(1024) k_pp_pick(const double *__restrict__ bsum_all, int64_t nblocks,
                                                  const float *__restrict__ D_all, int64_t n,
                                                  const double *__restrict__ u_all, const float *__restrict__ data_all,
                                                  int64_t row_stride, int d, int k, int j,
                                                  float *__restrict__ cents_all) {
    const int64_t g = blockIdx.x;
    const double *bsum = bsum_all + g * nblocks;
    const float *D = D_all + g * n;
    const double *u = u_all + g * (k - 1) + (j - 1);
    const float *data = data_all + g * d;
    float *cent_out = cents_all + (g * k + j) * d;
    __shared__ double s_tot[1024];
    __shared__ double s_bs[1024 * PP_BS];   // the block sums (when at most PP_BS a thread)
    __shared__ float s_d[PP_BLOCK];
    __shared__ double s_run;
    __shared__ int64_t s_pick, s_last, s_blk;
    __shared__ int s_owner;
    const int t = threadIdx.x;
    // per-thread chunk sums in block order, then an inclusive scan over threads
    const int64_t per = (nblocks + 1023) / 1024;
    const int64_t b0 = t * per, b1 = b0 + per < nblocks ? b0 + per : nblocks;
    const bool in_lds = per <= PP_BS;
    double mine = 0.0;
    for (int64_t b = b0; b < b1; ++b) {
        const double v = bsum[b];
        if (in_lds) s_bs[b] = v;
        mine += v;
    }
    s_tot[t] = mine;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const double v = t >= o ? s_tot[t - o] : 0.0;
        __syncthreads();
        s_tot[t] += v;
        __syncthreads();
    }
    const double total = s_tot[1023];
    if (t == 0) {
        s_pick = -1;
        s_last = -1;
        s_owner = -1;
    }
    __syncthreads();
    if (!(total > 0.0)) {
        if (t == 0) {
            int64_t r = (int64_t)(*u * (double)n);
            s_pick = r < n ? r : n - 1;
        }
    } else {
        const double target = *u * total;
        // [previous inclusive sum, own inclusive sum): disjoint intervals,
        // so at most one thread holds the target
        const double excl = t > 0 ? s_tot[t - 1] : 0.0;
        if (excl <= target && target < s_tot[t]) {
            s_owner = t;
            s_run = excl;
            s_blk = b0;
        }
        __syncthreads();
        const int ot = s_owner;
        if (ot >= 0) {
            const int64_t ob1 = ot * per + per < nblocks ? ot * per + per : nblocks;
            for (;;) {
                if (t == ot) {
                    // the next block whose sum carries the running sum past the target
                    double run = s_run;
                    int64_t b = s_blk;
                    if (in_lds)
                        while (b < ob1 && run + s_bs[b] <= target) run += s_bs[b++];
                    else
                        while (b < ob1 && run + bsum[b] <= target) run += bsum[b++];
                    s_run = run;
                    s_blk = b;
                }
                __syncthreads();
                const int64_t b = s_blk;
                if (b >= ob1) break;
                const int64_t r0 = b * PP_BLOCK;
                if (t < PP_BLOCK) s_d[t] = r0 + t < n ? D[r0 + t] : 0.f;
                __syncthreads();
                if (t == ot) {
                    double run = s_run;
                    int64_t pick = -1, last = s_last;
                    const int rows = (int)(n - r0 < PP_BLOCK ? n - r0 : PP_BLOCK);
                    // in row order; eight rows' values read ahead of their
                    // dependent sums (s_d holds PP_BLOCK values)
                    bool done = false;
                    for (int rb = 0; rb < rows && !done; rb += 8) {
                        float dv[8];
#pragma unroll
                        for (int q = 0; q < 8; ++q) dv[q] = s_d[rb + q];
#pragma unroll
                        for (int q = 0; q < 8; ++q) {
                            const int r = rb + q;
                            if (done || r >= rows) continue;
                            const float dr = dv[q];
                            if (dr > 0.f) last = r0 + r;
                            run += (double)dr;
                            if (run > target && dr > 0.f) {
                                pick = r0 + r;
                                done = true;
                            }
                        }
                    }
                    s_run = run;
                    s_last = last;
                    s_pick = pick;
                    s_blk = b + 1;
                }
                __syncthreads();
                if (s_pick >= 0) break;
            }
            if (t == ot && s_pick < 0) s_pick = s_last;
        }
    }
    __syncthreads();
    if (s_pick < 0 && t == 0) {
        // rounding left the crossing unresolved: the last row with D > 0
        for (int64_t r = n - 1; r >= 0; --r)
            if (D[r] > 0.f) {
                s_pick = r;
                break;
            }
        if (s_pick < 0) s_pick = 0;
    }
    __syncthreads();
    const int64_t p = s_pick;
    for (int e = t; e < d; e += 1024) cent_out[e] = data[p * row_stride + e];
}

}  // namespace dev

void launch_kmeanspp(const float *data, int64_t n, int64_t row_stride, int d, int k, int groups,
                     const int64_t *first, const double *u_dev, float *cents, hipStream_t st) {
    using dev::PP_BLOCK;
    if (k <= 0 || n <= 0 || groups <= 0) return;
    const int64_t nb = ceil_div(n, (int64_t)PP_BLOCK);
    if (groups > 65535 || nb > 0x7fffffff) fail(HANN_ERR_UNSUPPORTED, "k-means++: too many groups or rows");
    DevBuf<float> D((size_t)groups * n);
    DevBuf<int32_t> A((size_t)groups * n);
    DevBuf<double> bsum((size_t)groups * nb);
    DevBuf<double> cd((size_t)groups * k);
    // room over the FP32 rounding of the distances (at most (d + 2) u of
    // each), with a wide margin
    const double margin = 1.0 + 8.0 * (double)(d + 2) / 16777216.0 + 1e-5;
    // the first centers
    for (int g = 0; g < groups; ++g)
        HANN_HIP(hipMemcpy2DAsync(cents + (int64_t)g * k * d, d * sizeof(float),
                                  data + first[g] * row_stride + (int64_t)g * d, d * sizeof(float), d * sizeof(float),
                                  1, hipMemcpyDeviceToDevice, st));
    const dim3 ugrid((unsigned)nb, (unsigned)groups);
    for (int j = 1; j < k; ++j) {
        if (j == 1)
            hipLaunchKernelGGL(dev::k_pp_update<true>, ugrid, dim3(PP_BLOCK), 0, st, data, n, row_stride, d, cents, k,
                               0, nullptr, margin, D.get(), A.get(), bsum.get());
        else
            hipLaunchKernelGGL(dev::k_pp_update<false>, ugrid, dim3(PP_BLOCK), 0, st, data, n, row_stride, d, cents,
                               k, j - 1, cd.get(), margin, D.get(), A.get(), bsum.get());
        HANN_LAUNCH_CHECK();
        hipLaunchKernelGGL(dev::k_pp_pick, dim3((unsigned)groups), dim3(1024), 0, st, bsum.get(), nb, D.get(), n, u_dev,
                           data, row_stride, d, k, j, cents);
        HANN_LAUNCH_CHECK();
        if (j + 1 < k) {
            hipLaunchKernelGGL(dev::k_pp_center_dist, dim3((unsigned)ceil_div((int64_t)j, (int64_t)4), (unsigned)groups),
                               dim3(256), 0, st, cents, d, k, j, cd.get());
            HANN_LAUNCH_CHECK();
        }
    }
    HANN_HIP(hipStreamSynchronize(st));
}

}  // namespace hann
