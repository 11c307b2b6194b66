// runtime.h -- device context shared by the C-ABI entry points.
#pragma once

#include <mutex>

#include "common.h"
#include "kernels.h"

namespace hann {

// One per process: the library's own stream and scratch for the stateless
// entry points. `mu` serialises their use of the scratch.
struct Context {
    hipStream_t stream = nullptr;
    Workspace ws;
    Workspace ws_bf16;   // the BF16 candidate pass
    std::mutex mu;
};

Context &global_ctx();
void h2d(void *dst, const void *src, size_t bytes, hipStream_t st);
void d2h(void *dst, const void *src, size_t bytes, hipStream_t st);
void sync(hipStream_t st);

}  // namespace hann

// C-ABI bodies: turn hann::Error into a status and the thread's last error.
#define HANN_API_BEGIN try {
#define HANN_API_END                                                    \
    return HANN_OK;                                                     \
    }                                                                   \
    catch (const hann::Error &e) {                                      \
        hann::set_last_error(e.what());                                 \
        return e.code;                                                  \
    }                                                                   \
    catch (const std::bad_alloc &) {                                    \
        hann::set_last_error("host allocation failed");                 \
        return HANN_ERR_OOM;                                            \
    }                                                                   \
    catch (const std::exception &e) {                                   \
        hann::set_last_error(e.what());                                 \
        return HANN_ERR_DEVICE;                                         \
    }
