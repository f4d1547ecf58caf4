// abd_internal.h — host-side internals shared by the libabd_hip.so translation
// units: error handling, the per-device scratch arena, and launchers of the
// generic scan / reduction kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/abd_hip.h"

namespace abd {

#define ABD_CHECK(expr)                                   \
    do {                                                  \
        hipError_t _e = (expr);                           \
        if (_e != hipSuccess) return -static_cast<int>(_e); \
    } while (0)
#define ABD_CHECK_LAUNCH() ABD_CHECK(hipGetLastError())

inline hipStream_t to_stream(abd_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// Per-device scratch arena: `slots` independent growable device buffers plus a
// small pinned host buffer for count read-back.  Growing synchronizes the
// device first, so a buffer still in use by queued work is never freed early.
constexpr int kArenaSlots = 19;
enum ArenaSlot {
    kSlotScanA = 0,
    kSlotScanB,
    kSlotReduce,
    kSlotFlags,
    kSlotList,
    kSlotHull1,
    kSlotHull2,
    kSlotFilter,
    kSlotFilterCount,
    kSlotChunk,
    kSlotChunkCount,
    kSlotContact,
    kSlotPair,
    kSlotMisc,
    kSlotSolver,
    kSlotSpare,
    kSlotSolveBytes,
    kSlotPlan,
    kSlotPrecond,
};
// Launch shapes that change no result (abd_set_tuning); process-wide.
AbdTuning tuning();
// Swept boxes dilated at both ends (kind ABD_KIND_*, or -1 for body boxes): the union of
// the dilated boxes at x0 and at x1.
int swept_boxes(hipStream_t s, const AbdScene* sc, int kind, const double* x0, const double* x1, double dil,
                float* lo, float* hi);
int arena_get(int slot, size_t bytes, void** ptr);
int arena_host(size_t bytes, void** ptr);  // pinned host memory
int arena_release_current();

// Exclusive scan of n int64 counts (device) in place into `out` (may alias),
// writing the total to *total (device).  Fixed-order, single workgroup passes.
int scan_exclusive_i64(hipStream_t s, const int64_t* in, int64_t* out, int64_t n, int64_t* total);
// Deterministic sum of n doubles: *out (device) = sum, fixed order.
// Copies one int64 from device to host (synchronizes `s`).
int read_i64(hipStream_t s, const int64_t* dev, int64_t* host);

constexpr int kBlock = 256;
inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline unsigned grid_for(int64_t n, int block = kBlock, int64_t cap = 65535) {
    int64_t g = ceil_div(n, block);
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return static_cast<unsigned>(g);
}

}  // namespace abd
