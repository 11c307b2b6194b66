// kmeans.h -- launchers of the k-means / PQ kernels (kmeans.hip).
#pragma once

#include "common.h"

namespace hann {

// Device scratch of the k-means update (stable grouping by cluster).
struct KMeansScratch {
    DevBuf<int32_t> hist, pos, tile_sums, members;
    DevBuf<int64_t> offs, reseed;
    void reserve(int64_t n, int k, int groups);
};

// out[i] = nearest centroid of row i (rows row_stride floats apart)
void launch_assign(const float *data, int64_t n, int64_t row_stride, const float *cents, int k, int d,
                   int32_t *out, hipStream_t st);
// groups of small (d < 32) problems: group g reads columns [g*d, (g+1)*d)
void launch_assign_groups(const float *data, int64_t n, int64_t row_stride, const float *cents,
                          int64_t cent_group_stride, const int *kg_dev, int k_single, int d, int groups,
                          int32_t *out, int64_t out_group_stride, hipStream_t st);
// group members by cluster in point order: sc.members, sc.offs
void launch_group_members(const int32_t *assign, int64_t n, int64_t assign_group_stride, int k, int groups,
                          KMeansScratch &sc, hipStream_t st);
// new centroids from sc.members/sc.offs; empty clusters copy row sc.reseed[g*k+c]
void launch_kmeans_sums(const float *data, int64_t n, int64_t row_stride, int d, int k, int groups, float *cents,
                        int64_t cent_group_stride, KMeansScratch &sc, hipStream_t st);
void launch_encode(const float *vecs, int64_t n, int d, const float *coarse, const int32_t *assign, int m,
                   const float *cb, int pmax, const int64_t *cb_len_dev, int32_t *codes, float *resid_scratch,
                   hipStream_t st);
void launch_gather_group(const float *data, int64_t row_stride, int d, int k, int groups, const int64_t *idx,
                         float *cents, hipStream_t st);
void launch_residuals(const float *vecs, int d, const float *coarse, const int32_t *assign, const int32_t *order,
                      int64_t n, float *out, hipStream_t st);
void launch_gather_rows(const float *src, int d, const int64_t *idx, int64_t n, float *out, hipStream_t st);

// k-means++ seeding of `groups` groups side by side on the device (group g:
// columns [g d, g d + d) of rows row_stride floats apart, its centers at
// cents + g k d, its draws at u_dev + g (k - 1)): cents[0] = row first[g];
// then for j = 1 .. k-1 the row where the running prefix of D^2 (squared
// distance to the nearest chosen center, reference FP32 order; prefix in
// FP64, fixed order) first exceeds u[j-1] times its total (a uniform row
// when the total is 0).
void launch_kmeanspp(const float *data, int64_t n, int64_t row_stride, int d, int k, int groups,
                     const int64_t *first, const double *u_dev, float *cents, hipStream_t st);

}  // namespace hann
