// Internal helpers shared by the HIP sources of libatomx_slater.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "atomx_slater.h"

namespace ax {

void set_error(const char *fmt, ...);

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

inline int grid_for(int64_t work, int threads, int cap = 65535) {
    int64_t g = (work + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return static_cast<int>(g);
}

// Gauss-Legendre nodes and weights on [-1, 1] (ascending), computed on the
// host in long double by Newton iteration on P_n.
void gauss_legendre(int n, double *x, double *w);

// Caching device allocator (pool.hip): blocks are rounded up to a size class
// and returned to a per-device free list on pool_free.
hipError_t pool_alloc(void **ptr, size_t bytes);
void pool_free(void *ptr);
size_t pool_class(size_t bytes);
// Non-blocking stream of the calling thread on the current device, for
// set-up copies (never destroyed).
hipStream_t setup_stream();
// Pinned host staging buffer of at least `bytes` for work queued on
// `stream`; staging_release records the slot's event on that stream after
// the copies out of the buffer, and the reads of its device mirror, have been
// queued there, and the slot is handed out again (to users of the same
// stream only) once that event has completed.  A slot is bound to one stream
// because an event last recorded on a stream that has since been destroyed
// can no longer be queried.  With `dev` non-null the slot also hands out a
// device buffer of the same size on the current device (kept with the slot),
// so a whole staging buffer moves with one copy.
hipError_t staging_acquire(size_t bytes, hipStream_t stream, void **ptr, int *slot,
                           void **dev = nullptr);
hipError_t staging_release(int slot);
// Frees the idle staging slots bound to `stream`, whose queued work must have
// completed; called before a library-owned stream is destroyed.
void staging_forget_stream(hipStream_t stream);

}  // namespace ax

#define AX_CHECK_HIP(expr)                                                        \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess) {                                                   \
            ax::set_error("%s failed: %s", #expr, hipGetErrorString(e_));         \
            return AX_ERR_HIP;                                                    \
        }                                                                         \
    } while (0)

#define AX_CHECK_LAUNCH()                                                         \
    do {                                                                          \
        hipError_t e_ = hipGetLastError();                                        \
        if (e_ != hipSuccess) {                                                   \
            ax::set_error("kernel launch failed: %s", hipGetErrorString(e_));     \
            return AX_ERR_HIP;                                                    \
        }                                                                         \
    } while (0)

#define AX_REQUIRE(cond, ...)                                                     \
    do {                                                                          \
        if (!(cond)) {                                                            \
            ax::set_error(__VA_ARGS__);                                           \
            return AX_ERR_VALUE;                                                  \
        }                                                                         \
    } while (0)

// AX_REQUIRE with an explicit status code (AX_ERR_MEMORY for size limits)
#define AX_REQUIRE_CODE(cond, code, ...)                                          \
    do {                                                                          \
        if (!(cond)) {                                                            \
            ax::set_error(__VA_ARGS__);                                           \
            return (code);                                                        \
        }                                                                         \
    } while (0)
