// abd_profile.h — optional per-kernel timing with HIP events recorded on the
// launch stream, plus the algorithmic bytes each launch moves.  Off unless
// abd_profile_begin() was called; then every instrumented launch records a
// start/stop event pair that abd_profile_query() resolves.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace abd {

enum ProfKernel {
    kProfClassify = 0,  // gh stage A: candidate classification (HBM stream)
    kProfContact,       // gh stage C: per active contact terms
    kProfPairReduce,    // gh stage D
    kProfBodyFinalize,  // gh stage E
    kProfEnergy,        // line-search energy
    kProfToi,           // CCD
    kProfFilter,        // culling: hull filter
    kProfPrimCount,     // culling: chunk count pass
    kProfPrimFill,      // culling: chunk write pass
    kProfBodyPairs,     // culling: body pairs (all passes)
    kProfVg,            // culling: vertex-ground
    kProfSolve,         // Krylov solve (one workgroup)
    kProfProject,       // vertex projection
    kProfBoxes,         // primitive boxes
    kProfPrecond,       // block-Jacobi: diagonal-block inverses
    kProfCount
};

bool prof_enabled();
// Records the start of an instrumented launch; returns a handle (or -1 when off).
int prof_start(int kernel, hipStream_t s, double bytes);
// `dev_scale` (nullable): device int32 whose value multiplies `bytes` at query time
// (iteration counts); `dev_bytes` (nullable): device double, likewise (byte counts
// computed on the device, e.g. from per-system iteration counts).  Both are copied
// on the launch stream, so recording them costs no host synchronization.
void prof_stop(int handle, hipStream_t s, const int32_t* dev_scale = nullptr, const double* dev_bytes = nullptr);

}  // namespace abd
