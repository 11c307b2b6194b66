// mfma.h -- launchers of the BF16 MFMA candidate pass (mfma.hip).
#pragma once

#include "common.h"
#include "kernels.h"

namespace hann {

// rows -> BF16 [n][dp] (dp a multiple of 32), ||x_b||^2, ||x||, max ||x||, inexact flag
void launch_to_bf16(const float *x, int64_t n, int d, int dp, __bf16 *out, float *norm_b, float *norm_f,
                    unsigned *max_norm, int *inexact, hipStream_t st);
// tau_rank[q] = exact rank of (query q, row hint[q]) -- an upper bound of its nearest row
void launch_hint_rank(const float *queries, int64_t nq, int d, const float *base, const int32_t *hint,
                      double *tau_rank, hipStream_t st);
// filter thresholds from the sample's exact ranks tau_rank[q*k + k-1]
void launch_thresholds(const double *tau_rank, int k, int nq, const float *qnorm_f, const unsigned *max_norm,
                       const int *inexact, int d, float *thr, hipStream_t st);
// append every row with approx distance <= thr[q] to q's candidate list
// (cnt[q] > cap: more candidates than room); a workgroup covers rt_per_wg
// tiles of 128 rows and holds mfma_filter_slots(dp) LDS slots per query.
// Where mfma_filter_writes_approx(dp) (dp <= 128), each candidate's
// approximate distance goes to cand_ap (same layout as cand_row)
bool mfma_filter_writes_approx(int dp);
int mfma_filter_slots(int dp);
void launch_mfma_filter(const __bf16 *Q, int64_t nq, const __bf16 *X, int64_t nb, int dp, const float *qn,
                        const float *xn, const float *thr, int rt_per_wg, int32_t *cnt, int32_t *cand_row,
                        float *cand_ap, int cap, hipStream_t st);
// every approximate distance: dense[q * nb + row]
void launch_mfma_dense(const __bf16 *Q, int64_t nq, const __bf16 *X, int64_t nb, int dp, const float *qn,
                       const float *xn, float *dense, hipStream_t st);
// every approximate distance's 16-bit key (its top 16 bits; negative
// values 0): keys[q * nb + row], and each query's least and largest key
// (qlo[q], qhi[q]); dp <= 128, nb a multiple of 4
void launch_mfma_keys(const __bf16 *Q, int64_t nq, const __bf16 *X, int64_t nb, int dp, const float *qn,
                      const float *xn, uint16_t *keys, unsigned *qlo, unsigned *qhi, hipStream_t st);
// per query, from its nb keys: the certified threshold of its k-th
// smallest and every row whose key is within it (cand_row[q * cap + i],
// cnt[q] of them; more than cap: the caller ranks every row)
// (nb a multiple of 8, at most dense_select_max_rows())
int64_t dense_select_max_rows();
void launch_dense_select(const uint16_t *keys, const unsigned *qlo, const unsigned *qhi, int64_t nq, int64_t nb,
                         int k, const int32_t *k_q, const float *qnorm_f, const unsigned *max_norm,
                         const int *inexact, int d, int cap, int32_t *cnt, int32_t *cand_row, hipStream_t st);
// out rows = X rows 0, stride, 2 stride, ... (n of them) with their BF16 norms
void launch_gather_bf16(const __bf16 *X, int dp, const float *xn, int64_t stride, int64_t n, __bf16 *out,
                        float *out_n, hipStream_t st);
// filter thresholds from the k-th smallest approximate distance of each
// query's ns sample rows (approx[q * ns + j]), or, with cnt, of its first
// cnt[q] values (its candidates; +inf when fewer than k or more than ns)
// (k_q: each query's own k, clamped to [1, k], or null)
void launch_sample_thresholds(const float *approx, int64_t nq, int64_t ns, const int32_t *cnt, int k,
                              const int32_t *k_q, const float *qnorm_f, const unsigned *max_norm, const int *inexact,
                              int d, float *thr, hipStream_t st);
// exact reference-order ranking of the candidates; a query whose candidates
// overflowed (cnt > cap) ranks all nb rows instead when nb > 0, and is left
// untouched when nb == 0
// cand_ap and thr2 (or null): only candidates with cand_ap <= thr2[q] are
// ranked
void launch_rank_candidates(const float *queries, int64_t nq, int d, const float *base, int64_t nb,
                            const int32_t *cnt, const int32_t *cand_row, const float *cand_ap, const float *thr2,
                            int cap, int k, const int32_t *k_q, int sqrt_out, int64_t *out_ids, double *out_dist,
                            hipStream_t st);
// the same over int64 candidate rows (rows[q * cap + i], i < cnt[q])
void launch_rank_rows64(const float *queries, int64_t nq, int d, const float *base, int64_t nb, const int32_t *cnt,
                        const int64_t *rows, int cap, int k, const int32_t *k_q, int sqrt_out, int64_t *out_ids,
                        double *out_dist, hipStream_t st);

}  // namespace hann
