// scan_fused.h -- persistent ADC list scan (scan_fused.hip).
#pragma once

#include "ivf.h"

namespace hann {

struct FusedArgs {
    const float *T;            // [K][m][P] list tables
    const double *probe_rank;  // [nq][probe_stride] exact coarse ranks of the probes, or null
    int32_t *order;            // [nq] work queue (written by the launch)
    int32_t *counter;          // work counter (reset by the launch)
    int nq;
};

void launch_list_tables(const float *coarse, int K, int d, int m, int P, const float *cb, const int64_t *cb_len,
                        float *T, hipStream_t st);
size_t fused_lds_bytes(int wpq, int lut_bytes, int d, int m, int P, int cap);
// wpq waves per query (1, 2 or 4); grid = persistent workgroups
void launch_scan_fused(const ScanArgs &a, FusedArgs f, int64_t nq, int wpq, int lut16, int code_bytes, int maxp,
                       int grid, hipStream_t st);

}  // namespace hann
