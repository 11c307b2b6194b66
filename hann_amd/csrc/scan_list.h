// scan_list.h -- list-major decode + MFMA scan (scan_list.hip).
#pragma once

#include "ivf.h"

namespace hann {

// One batch's (query, probed list) pairs grouped by list, and the scan's
// candidate lists. Every device buffer is sized by the caller.
struct ListScanArgs {
    const float *queries;        // [nq][d]
    int d, m, s, P, dk;          // dim, subquantizers, sub-dim, codebook stride, d padded to 16
    const int64_t *list_start;   // [K]
    const int64_t *list_len;     // [K]
    const void *codes;           // [N][m] uint8 or uint16
    const float *term;           // [N] per-entry list term
    const float *cb;             // [m][P][s]
    const int32_t *pair_q;       // [pairs] query of each pair, grouped by list
    const float *pair_qc;        // [pairs] |q - c_l|^2 (the coarse rank)
    const int4 *items;           // [n_items] (list, first pair, pairs <= 32, 0)
    const int32_t *n_items;      // device count of items
    int32_t *item_counter;       // zeroed before the launch
    const float *tau;            // [nq] candidate threshold
    int32_t *cnt;                // [nq] candidates appended (zeroed before the launch)
    float *cand_key;             // [nq][cap]
    int32_t *cand_pos;           // [nq][cap]
    int cap;
};

void launch_fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t st);
// per-list pair counts and per-query entry totals (lcnt zeroed by the caller)
void launch_pair_count(const ScanArgs &a, int64_t nq, int32_t *lcnt, int64_t *qtot, hipStream_t st);
// exclusive offsets of pairs and of 32-query items per list; totals[0] =
// pairs, totals[1] = items; lfill is zeroed
void launch_pair_offsets(const int32_t *lcnt, int K, int32_t *lstart, int32_t *ioff, int32_t *lfill,
                         int32_t *totals, hipStream_t st);
void launch_pair_scatter(const ScanArgs &a, int64_t nq, const double *probe_rank, const int32_t *lstart,
                         int32_t *lfill, int32_t *pair_q, float *pair_qc, hipStream_t st);
void launch_make_items(const int32_t *lcnt, const int32_t *lstart, const int32_t *ioff, int K, int4 *items,
                       hipStream_t st);
// tau[q] from the pre-pass over the nearest lists (its keys in ascending
// order): the selk-th key, widened by the bound on |MFMA - table| distance
// (pre_lists and qtot are kept for threshold estimates that use them)
void launch_list_thresholds(const ScanArgs &a, int64_t nq, const float *pre_key, const int32_t *pre_cnt,
                            int pre_stride, int pre_lists, const int64_t *qtot, const double *probe_rank,
                            const float *ymax_tmax, int x3, float *tau, hipStream_t st);
// the list-major scan covers s = d/m = 4, m in {8, 16, 24, 32, 48, 64},
// uint8 codes, and a BF16 codebook that fits in LDS
bool list_scan_supported(int d, int m, int P, int code_bytes);
size_t list_scan_lds_bytes(int x3, int d, int m, int P);
int list_scan_threads();
void launch_scan_list(const ListScanArgs &la, int x3, int grid, hipStream_t st);
// top selk = k_q * R candidates per query by (key, position), as positions
// in out_pos rows; queries that overflowed or need the brute-force
// fallback go to ovf (count in *ovf_cnt, zeroed by the caller)
void launch_select_candidates(const float *cand_key, const int32_t *cand_pos, const int32_t *cnt, int cap,
                              const int64_t *qtot, const int32_t *k_q, int k_default, int R, int allow_fallback,
                              int64_t nq, int64_t *out_pos, int32_t *out_cnt, int out_stride, int32_t *ovf,
                              int32_t *ovf_cnt, hipStream_t st);

}  // namespace hann
