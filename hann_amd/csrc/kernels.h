// kernels.h -- host-side launchers of the hann_amd device kernels.
#pragma once

#include <memory>

#include "common.h"

namespace hann {

// Grow-only device scratch, in numbered slots, owned by one stream's user.
// Slot map: 0-4 and 10-11 host-facing staging of the entry points, 5-9 the
// search pipeline, 12-13 the persistent scan's queue, 14-15
// launch_topk_exact's partial lists (14 also its row path's selection), 16
// device_min, 17-19 the list-major scan, 20 the batch's query tables, 21-23
// launch_topk_exact's row path.
struct Workspace {
    DevBuf<char> slots[24];
    template <typename T>
    T *get(int slot, size_t count) {
        slots[slot].resize(count * sizeof(T) + 256);
        return reinterpret_cast<T *>(slots[slot].ptr);
    }
    // a second workspace owned by this one (for a pass nested in a search)
    Workspace &aux() {
        if (!aux_) aux_.reset(new Workspace);
        return *aux_;
    }
    // device bytes held (this workspace's slots and its nested one's)
    size_t bytes() const {
        size_t b = 0;
        for (const auto &s : slots) b += s.n;
        return b + (aux_ ? aux_->bytes() : 0);
    }

  private:
    std::unique_ptr<Workspace> aux_;
};

int kp_for(int k);
// an upper bound of the workspace bytes launch_topk_exact takes for this shape
size_t topk_exact_scratch_bytes(int metric, int64_t nq, int64_t nb, int k);
// the same for its tiled kernel alone
size_t topk_exact_tiled_scratch_bytes(int metric, int64_t nq, int64_t nb, int k);

// distance.hip
// normalize n rows of dim floats in place (core/simd_kernels.inc.h:267-315)
void launch_normalize(float *x, int64_t n, int dim, hipStream_t st);
void launch_distance_batch(int metric, const float *queries, int64_t nq, const float *base,
                           int64_t nb, int dim, double *out, hipStream_t st);
void launch_topk_exact(int metric, const float *queries, int64_t nq, const float *base, int64_t nb,
                       int dim, int k, int64_t *out_ids, double *out_dist, Workspace &ws,
                       hipStream_t st);
// The same without allocating: the partial lists live in a static device
// array of kNoAllocScratchBytes per GPU (nq * k * 12 bytes must fit).
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
void launch_topk_exact_noalloc(int metric, const float *queries, int64_t nq, const float *base, int64_t nb,
                               int dim, int k, int64_t *out_ids, double *out_dist, hipStream_t st);
// the static array's device address on the current device (looked up once
// per device)
void *noalloc_scratch_current();

}  // namespace hann
