// scan_list.h -- list-major decode + MFMA scan (scan_list.hip).
#pragma once

#include <vector>

#include "ivf.h"
#include "kernels.h"

namespace hann {

// One batch's (query, probed list) pairs grouped by list, and the rows the
// scan writes. Every device buffer is sized by the caller.
struct ListScanArgs {
    const float *queries;        // [nq][d]
    int d, m, s, P, dk;          // dim, subquantizers, sub-dim, codebook stride, d padded to 16
    int64_t nq;
    const int64_t *list_start;   // [K]
    const int64_t *list_len;     // [K]
    const void *codes;           // [N][m] uint8 or uint16
    const float *term;           // [N] per-entry list term
    const float *cb;             // [m][P][s]
    const int32_t *pair_q;       // [pairs] query of each pair, grouped by list
    const float *pair_qc;        // [pairs] |q - c_l|^2 (the coarse rank)
    const float *pair_scale;     // [pairs] the query's key scale (a power of two)
    const int64_t *pair_off;     // [pairs] where the pair's keys start in dist
    const int4 *items;           // [n_items] (list, first pair, pairs <= 32, first entry of the chunk)
    const int32_t *n_items;      // device count of items
    int32_t *item_counter;       // [n_parts] zeroed before the launch
    int n_parts;                 // item partitions (at most the grid)
    uint16_t *dist;              // the sample pass's 16-bit keys in the batch's sample rows: query q's
                                 // sample probes in order, each probe's run padded to a multiple of 8 slots
    uint32_t dist_bytes;         // size of dist in bytes (below 2^32 - 64)
    // the filter pass
    const uint32_t *theta;       // [nq] keys <= theta[q] survive
    uint64_t *rec;               // [nq][capq] each query's records (key << 32 | entry position), in the
                                 // order they were written; the first capq of them are kept
    uint32_t *qcnt;              // [nq] records written per query (zeroed before the launch)
    uint32_t capq;               // record room per query
    const float *coarse;         // [K][d] coarse centroids (the exact scan)
    int chunk;                   // entries per work item (a multiple of 64; at most 2,048)
    float cb_scale;              // FP8 mode: the codebook's scale (a power of two)
};

// work items: up to 2^qg queries of a list times a chunk of up to 2^ch of
// its entries (qg <= 5, 6 <= ch <= 11)
struct ItemShape {
    int qg, ch;
};
// the MFMA scan's items (16 queries x 2,048 entries) and the exact scan's
// (8 x 256) when untuned
constexpr ItemShape kListItems{4, 11};
constexpr ItemShape kExactItems{3, 8};
// a list-major scan's configuration: its work items and item partitions
struct ListTune {
    ItemShape sh;
    int parts;
};
// the configurations the autotuner times (the first is the default); the
// smallest chunk and query group among them size the batch's item buffers
constexpr ListTune kListTunes[] = {{{4, 11}, 256}, {{4, 10}, 256}, {{4, 11}, 128}, {{4, 10}, 128}};
constexpr ListTune kExactTunes[] = {{{3, 8}, 256}, {{2, 8}, 256}, {{3, 7}, 256}, {{3, 9}, 256}, {{4, 8}, 256}};
constexpr ItemShape kListItemsMin{4, 10};
constexpr ItemShape kExactItemsMin{2, 7};
// entries of a list per work item of the list-major scan
int list_scan_chunk();
// work-item partitions (counters) of the list-major scan
constexpr int kListScanParts = 256;
void launch_fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t st);
// per-list pair counts (lcnt) and sample pair counts (lscnt), per-query
// entry totals, and the sample's probes and row slots (its nearest probed
// lists: at least one, and until they hold 4 k R entries; each probed
// list's run is its length rounded up to 8); lcnt, lscnt zeroed by the caller
void launch_pair_count(const ScanArgs &a, int64_t nq, int32_t *lcnt, int32_t *lscnt, int64_t *qtot,
                       int32_t *qsamp_np, int64_t *qsamp_slots, hipStream_t st);
// the sample rows' starts: row_off[q] = qslots[0] + ... + qslots[q - 1]
void launch_row_offsets(const int64_t *qslots, int64_t nq, int64_t *row_off, hipStream_t st);
// exclusive offsets of pairs and of work items per list (ceil(pairs / 32)
// query groups times ceil(len / chunk) entry chunks); totals[0] = pairs,
// totals[1] = items, totals[2] = 0 (then the sample pass's items); lfill and
// lbfill are zeroed
void launch_pair_offsets(const int32_t *lcnt, const int64_t *list_len, int K, int32_t *lstart, int32_t *ioff,
                         int32_t *lfill, int32_t *lbfill, int32_t *totals, ItemShape sh, hipStream_t st);
// the pairs grouped by list (each list's sample pairs first), with the
// coarse rank, the query's key scale and, for a sample pair, where its keys
// start in the sample rows (0 for the others); ymax: bits of an upper bound
// of |y_e| over the index; q_scale (or null): each query's key scale
void launch_pair_scatter(const ScanArgs &a, int64_t nq, const double *probe_rank, const unsigned *ymax,
                         const int32_t *lstart, const int32_t *lcnt, const int32_t *qsamp_np, const int64_t *row_off,
                         int32_t *lfill, int32_t *lbfill, int32_t *pair_q, float *pair_qc, float *pair_scale,
                         int64_t *pair_off, float *q_scale, hipStream_t st);
// the batch's work items and the sample pass's (over each list's sample
// pairs; their count is added to *n_samp), longest first; order:
// list_item_order_ints() ints of scratch (per-bucket counts and fill counters),
// zeroed by the caller
void launch_make_items(const int32_t *lcnt, const int32_t *lscnt, const int32_t *lstart, const int64_t *list_len,
                       int K, int32_t *order, int4 *items, int4 *items_s, int32_t *n_samp, ItemShape sh,
                       hipStream_t st);
int list_item_order_ints();
// the list-major scan covers s = d/m = 4, m in {8, 16, 24, 32, 48, 64},
// uint8 codes, and a BF16 codebook that fits in LDS
bool list_scan_supported(int d, int m, int P, int code_bytes);
// kind 1: the sample pass's keys (every key of the sample pairs, in the
// rows); 2: the filter pass's records. x3: the operands, 0 BF16, 1 split
// BF16, 3 FP8 (e4m3; the codebook scaled by la.cb_scale)
void launch_scan_list_kind(const ListScanArgs &la, int x3, int kind, int grid, hipStream_t st);
size_t list_scan_lds_bytes_kind(int x3, int m, int P, int kind);
// The list-major exact scan: the reference's decode-and-measure distance
// (pqivf/index.go:905-936) of every (query, probed entry) pair in FP32, bit
// for bit, with the same sample / bound / filter / select passes; records
// hold the exact distance. d a multiple of 32, s in {4, 8}, uint8 codes and
// an FP32 codebook that fits in LDS.
bool exact_scan_supported(int d, int m, int P, int code_bytes);
size_t exact_scan_lds_bytes(int d, int m, int P);
void launch_scan_exact_kind(const ListScanArgs &la, int kind, int grid, hipStream_t st);
// The exact FP32 top-k of dense rows by sample, bound, filter and select
// (the same passes, records holding the exact rank), then ranked and
// ordered by (rank, row): identical results to launch_topk_exact's tiled
// kernel. (Squared) Euclidean, d in {64, 96, 128}, k <= 512, nb at least
// 16 k. A query whose records overflow its room is ranked over every row on
// the device (nb <= 65,536) or by the tiled kernel after a host check of the
// counts. Workspace slots 14, 21-23.
bool topk_rows_supported(int metric, int64_t nb, int dim, int k);
size_t topk_rows_scratch_bytes(int64_t nq, int64_t nb, int k);
void launch_topk_rows(int metric, const float *queries, int64_t nq, const float *base, int64_t nb, int dim, int k,
                      int64_t *out_ids, double *out_dist, Workspace &ws, hipStream_t st);
// the tiled kernel over the listed queries (global indices), results
// scattered into their rows of out_ids / out_dist; waits for the stream
void topk_tiled_redo(int metric, const float *queries, const std::vector<int64_t> &redo, const float *base,
                     int64_t nb, int dim, int k, int64_t *out_ids, double *out_dist, Workspace &ws, hipStream_t st);
// per query, the bound theta of the filter pass: the (k R)-th smallest key of
// its sample (qsamp_slots keys from row_off[q] of the sample rows)
void launch_theta(const ScanArgs &a, int64_t nq, const uint16_t *keys, const int64_t *row_off, const int64_t *qtot,
                  const int64_t *qsamp_slots, uint32_t *theta, hipStream_t st);
// per query, the selk = k_q * R best entries by (key, entry position) from the
// filter pass's records, as entry positions (-1 after the last); queries
// that take the brute-force fallback, and those with more than capq
// records, go to ovf (a per-query scan then ranks them). q_scale null: the
// records hold 16-bit keys; else exact FP32 distances, binned by the 16-bit
// key of the distance times q_scale[q]
size_t select_records_lds_bytes();
void launch_select_records(const ScanArgs &a, int64_t nq, const uint64_t *rec, const uint32_t *qcnt, uint32_t capq,
                           const int64_t *qtot, int64_t *out_pos, int32_t *out_cnt, int out_stride, int32_t *ovf,
                           int32_t *ovf_cnt, const float *q_scale, hipStream_t st);
// probes per query the threshold and select kernels take at most
int select_rows_max_probes();

}  // namespace hann
