// abd_profile.h — optional per-kernel timing with HIP events recorded on the
// launch stream, plus the algorithmic bytes each launch moves.  Off unless
// abd_profile_begin() was called; then every instrumented launch records a
// start/stop event pair that abd_profile_query() resolves.
#pragma once
#include <hip/hip_ext.h>
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
    kProfToiSolve,      // CCD: root searches of the queued candidates
    kProfKeep,          // superset filter: count, scans and write passes
    kProfCount
};

bool prof_enabled();
// Records the start of an instrumented launch; returns a handle (or -1 when off).
int prof_start(int kernel, hipStream_t s, double bytes);
// Ends the timed region of `handle` (the first call on a handle records its end event).
// `dev_scale` (nullable): device int32 whose value multiplies `bytes` at query time
// (iteration counts); `dev_bytes` (nullable): device double, likewise (byte counts
// computed on the device, e.g. from per-system iteration counts).  Both are copied on
// the launch stream after the end event, so recording them costs no host
// synchronization and no time inside the region; a second call after the kernels that
// compute such a value attaches it without moving the end.
void prof_stop(int handle, hipStream_t s, const int32_t* dev_scale = nullptr, const double* dev_bytes = nullptr);

// Algorithmic bytes of a launch whose list sizes are on the device (dn in the layout
// [nij, nij, npt, nee, nvg]): base + c_pt dn[2] + c_ee dn[3] + c_vg dn[4] + c_nij dn[0], or
// with `single` base + c_pt dn[0] (dn one count), computed on the stream and attached to
// `handle` (after its end event: call it after prof_stop or on a prof_start_ext record),
// whose own byte count (the multiplier of device values) must be 1.  Nothing when the
// handle is -1 or dn is null.
int prof_count_bytes(int handle, hipStream_t s, const int64_t* dn, double c_pt, double c_ee, double c_vg,
                     double c_nij, bool single = false, double base = 0.0);

// A device double (a slot of a ring reused in stream order) for a byte count that a caller's
// own kernel computes; attach it with prof_stop(handle, s, nullptr, *out) after that kernel.
// *out is null when the handle is -1.
int prof_byte_slot(int handle, double** out);

// A record whose events the kernel dispatch itself takes (prof_launch): it times the kernel
// alone, without the gaps around it (e.g. a launch behind a host synchronization).
int prof_start_ext(int kernel, double bytes);
void prof_ext_events(int handle, hipEvent_t* a, hipEvent_t* b);
void prof_ext_done(int handle);

// Launches `kernel`; with a live record (prof_start_ext) the dispatch records its start and
// stop events.  prof_stop may follow to attach device values.
template <typename F, typename... Args>
inline void prof_launch(int handle, F kernel, const dim3& grid, const dim3& block, uint32_t shm, hipStream_t s,
                        Args... args) {
    if (handle < 0) {  // an ordinary launch (also the form a recorded graph takes)
        hipLaunchKernelGGL(kernel, grid, block, shm, s, args...);
        return;
    }
    hipEvent_t a = nullptr, b = nullptr;
    prof_ext_events(handle, &a, &b);
    hipExtLaunchKernelGGL(kernel, grid, block, shm, s, a, b, 0, args...);
    prof_ext_done(handle);
}

}  // namespace abd
