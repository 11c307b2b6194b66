// ivf.h -- launchers of the inverted-list scan kernels (ivf.hip).
#pragma once

#include <hip/hip_fp16.h>

#include "common.h"
#include "kernels.h"

namespace hann {

enum {
    SCAN_MODE_EXACT = 0,
    SCAN_MODE_ADC_F32 = 1,
    SCAN_MODE_ADC_F16 = 2,
    SCAN_MODE_LIST_BF16 = 3,
    SCAN_MODE_LIST_BF16X3 = 4,
    SCAN_MODE_LIST_FP8 = 5
};

// Arguments of k_scan (passed by value).
struct ScanArgs {
    const float *queries;        // [nq][d]
    int d, m, s, pmax_cb;        // dim, subquantizers, sub-dim, codebook rows (stride)
    int K;                       // coarse lists
    const int64_t *probes;       // [nq][probe_stride] lists in probe order
    int probe_stride;
    int probe_cap;               // probes available per query (nprobe is clamped to it)
    const int32_t *nprobe_q;     // per-query nprobe or null
    int nprobe_default;
    const int32_t *k_q;          // per-query k (clamped to [1, k_default]) or null
    int k_default;               // k of every query without k_q, and the launch's largest k
    int rerank;                  // 0: ids are user ids; R > 0: keep R*k entry positions
    int allow_fallback;
    int cap;                     // selection buffer slots (2 * pow2 >= max selected)
    const int64_t *list_start;   // [K] first slot of list l
    const int64_t *list_len;     // [K] entries of list l (slots start..start+len)
    const int64_t *ids;          // [N]
    const void *codes;           // [N][m] uint8 or uint16
    const uint8_t *has_codes;    // [N] (read only when any_uncoded)
    int any_uncoded;
    const float *vecs;           // [N][d]
    const float *coarse;         // [K][d]
    const float *cb;             // [m][pmax_cb][s]
    const int64_t *cb_len;       // [m]
    float *out_key;              // [nq][out_stride]
    int64_t *out_id;             // [nq][out_stride]
    int out_stride;
    int32_t *out_cnt;            // [nq]
    int32_t *fell_back;          // [nq] or null
    unsigned long long *fb_counter;  // searches that fell back (device counter) or null
    // null, or k_scan's block b ranks query q_order[order_base + b] when
    // order_base + b < *q_order_n (and returns otherwise)
    const int32_t *q_order;
    const int32_t *q_order_n;
    int order_base;
};

size_t scan_lds_bytes(int mode, int d, int m, int pmax_cb, int cap);
void launch_scan(int mode, int code_bytes, const ScanArgs &a, int64_t nq, hipStream_t st);
// k_scan over the queries listed in order[0 .. *order_n) (at most nq_max of
// them; the count is read on the device)
void launch_scan_ordered(int mode, int code_bytes, const ScanArgs &a, int64_t nq_max, const int32_t *order,
                         const int32_t *order_n, hipStream_t st);
// k_q values are clamped to [1, kmax]
void launch_rerank(const float *queries, int d, int64_t nq, const int64_t *cand_pos, const int32_t *cand_cnt,
                   int cand_stride, const int32_t *k_q, int k_default, int kmax, const float *vecs,
                   const int64_t *ids, float *out_key, int64_t *out_id, int32_t *out_cnt, int out_stride,
                   hipStream_t st);
void launch_finalize(const float *key, int64_t n, float *out, hipStream_t st);

}  // namespace hann
