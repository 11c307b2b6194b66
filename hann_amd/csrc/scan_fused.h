// scan_fused.h -- persistent ADC list scan (scan_fused.hip).
#pragma once

#include "ivf.h"

namespace hann {

struct FusedArgs {
    const float *term;         // [entries] sum_i |cb_i|^2 + 2 <c_l,i, cb_i> of each entry's codes
    const double *probe_rank;  // [nq][probe_stride] exact coarse ranks of the probes, or null
    int32_t *order;            // [nq] work queue (written by the launch)
    int32_t *counter;          // work counter (reset by the launch)
    int nq;
    int lds_probes;            // probes per query kept in LDS (0: read from global memory)
    const int32_t *nq_dev;     // or null: queries order[0 .. *nq_dev) (the caller zeroes counter)
    const float *S;            // [nq][m][P] query tables, or null: built per query in LDS
};

size_t block_lds_bytes(int d, int m, int P, int cap, int lds_probes);
// S[q][i][j] = -2 <q_i, cb_ij> for every query of a batch
void launch_query_tables(const float *queries, int64_t nq, int d, int m, int P, const float *cb, float *S,
                         hipStream_t st);
// top-k room of the scan: a power of two >= max(2 kp, kp + 128 w)
int block_cap(int kp, int w);
// per-entry terms of the scan for lists[0..nlists) (every list when lists is
// null; nlists is then K); max_len bounds the lists' lengths; ymax_tmax (or
// null) gathers max |y_e| and max |t_e| as float bits
void launch_entry_terms(const int64_t *lists, int64_t nlists, int64_t max_len, const int64_t *list_start,
                        const int64_t *list_len, const void *codes, int code_bytes, const uint8_t *has, int d, int m,
                        int P, const float *coarse, const float *cb, float *term, unsigned *ymax_tmax,
                        hipStream_t st);
// the persistent scan (FP32 query table S_q, per-entry terms): w waves
// (2, 4 or 8) per query sharing one top-k of room block_cap(kp, w), with
// kp = kp_for(largest selection size); grid = persistent workgroups
void launch_scan_block(const ScanArgs &a, FusedArgs f, int64_t nq, int w, int kp, int code_bytes, int maxp,
                       int grid, hipStream_t st);

}  // namespace hann
