// abd_batch.hip — per-scene reductions and updates for many independent scenes
// advanced together (batching has no reference counterpart; each scene follows
// AffineBodySimulator.step, abd.py:212-276, on its own).  Bodies of a scene are
// contiguous, so every candidate list (sorted by body pair) and every vector
// indexed by body is a concatenation of per-scene segments; one workgroup
// reduces one scene in a fixed order.
#include <hip/hip_runtime.h>

#include "abd_block.h"
#include "abd_device.h"
#include "abd_inertia.h"
#include "abd_internal.h"

namespace abd {

constexpr int kSegThreads = 256;

__device__ __forceinline__ SceneView view_of_scene(const AbdScene& sc) {
    SceneView v;
    v.voff = sc.voff;
    v.foff = sc.foff;
    v.eoff = sc.eoff;
    v.tris = sc.tris;
    v.edges = sc.edges;
    v.x0 = sc.x0;
    return v;
}

// first k in [0, n) whose tuple's first body belongs to a scene >= s
__device__ __forceinline__ int64_t seg_lb(const int32_t* __restrict__ list, int w, int64_t n,
                                          const int32_t* __restrict__ body_scene, int s) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (body_scene[list[w * mid]] < s)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(kSegThreads) void scene_energy_kernel(AbdScene sc, AbdParams P,
                                                                    const double* __restrict__ xk,
                                                                    const int32_t* __restrict__ pt, int64_t npt,
                                                                    const int32_t* __restrict__ ee, int64_t nee,
                                                                    const int32_t* __restrict__ vg, int64_t nvg,
                                                                    double* __restrict__ E) {
    __shared__ double sh[17];
    const int s = blockIdx.x;
    const SceneView S = view_of_scene(sc);
    const int32_t* bs = sc.body_scene;
    double acc = 0.0;
    {
        const int64_t a = seg_lb(ee, 5, nee, bs, s), b = seg_lb(ee, 5, nee, bs, s + 1);
        for (int64_t k = a + threadIdx.x; k < b; k += blockDim.x) acc += cand_energy_ee(S, xk, ee + 5 * k, P.d2hat, P.kappa);
    }
    {
        const int64_t a = seg_lb(pt, 5, npt, bs, s), b = seg_lb(pt, 5, npt, bs, s + 1);
        for (int64_t k = a + threadIdx.x; k < b; k += blockDim.x) acc += cand_energy_pt(S, xk, pt + 5 * k, P.d2hat, P.kappa);
    }
    {
        const int64_t a = seg_lb(vg, 2, nvg, bs, s), b = seg_lb(vg, 2, nvg, bs, s + 1);
        for (int64_t k = a + threadIdx.x; k < b; k += blockDim.x)
            acc += cand_energy_vg(sc.voff, xk, vg + 2 * k, P.ground, P.d2hat, P.kappa);
    }
    const double r = block_sum_fast(acc, sh);
    if (threadIdx.x == 0) E[s] = r;
}

__global__ __launch_bounds__(kSegThreads) void scene_inertia_kernel(AbdScene sc, AbdParams P,
                                                                     const double* __restrict__ Ak,
                                                                     const double* __restrict__ pk,
                                                                     const double* __restrict__ A0,
                                                                     const double* __restrict__ p0,
                                                                     const double* __restrict__ Adot,
                                                                     const double* __restrict__ pdot,
                                                                     double* __restrict__ E) {
    __shared__ double sh[17];
    const int s = blockIdx.x;
    double acc = 0.0;
    for (int b = sc.scene_boff[s] + threadIdx.x; b < sc.scene_boff[s + 1]; b += blockDim.x)
        acc += inertia_energy_body(P, sc.mass, sc.I0, b, Ak, pk, A0, p0, Adot, pdot);
    const double r = block_sum_fast(acc, sh);
    if (threadIdx.x == 0) E[s] = r;
}

__global__ __launch_bounds__(kSegThreads) void scene_toi_kernel(AbdScene sc, AbdParams P,
                                                                 const double* __restrict__ x0,
                                                                 const double* __restrict__ x1,
                                                                 const int32_t* __restrict__ pt, int64_t npt,
                                                                 const int32_t* __restrict__ ee, int64_t nee,
                                                                 const int32_t* __restrict__ vg, int64_t nvg,
                                                                 double* __restrict__ toi) {
    __shared__ double sh[kSegThreads];
    const int s = blockIdx.x;
    const SceneView S = view_of_scene(sc);
    const int32_t* bs = sc.body_scene;
    double acc = 1.0;
    {
        const int64_t a = seg_lb(vg, 2, nvg, bs, s), b = seg_lb(vg, 2, nvg, bs, s + 1);
        for (int64_t k = a + threadIdx.x; k < b; k += blockDim.x) {
            const double t = cand_toi_vg(sc.voff, x0, x1, vg + 2 * k, P.ground);
            if (t < 1.0 && t < acc) acc = t;
        }
    }
    {
        const int64_t a = seg_lb(ee, 5, nee, bs, s), b = seg_lb(ee, 5, nee, bs, s + 1);
        for (int64_t k = a + threadIdx.x; k < b; k += blockDim.x) {
            const double t = cand_toi(1, S, x0, x1, ee + 5 * k);
            if (t < 1.0 && t < acc) acc = t;
        }
    }
    {
        const int64_t a = seg_lb(pt, 5, npt, bs, s), b = seg_lb(pt, 5, npt, bs, s + 1);
        for (int64_t k = a + threadIdx.x; k < b; k += blockDim.x) {
            const double t = cand_toi(0, S, x0, x1, pt + 5 * k);
            if (t < 1.0 && t < acc) acc = t;
        }
    }
    const double r = block_min(acc, sh);
    if (threadIdx.x == 0) toi[s] = r;
}

__global__ __launch_bounds__(kSegThreads) void scene_dot_kernel(const int32_t* __restrict__ sboff,
                                                                 const double* __restrict__ a,
                                                                 const double* __restrict__ b,
                                                                 double* __restrict__ out) {
    __shared__ double sh[17];
    const int s = blockIdx.x;
    double acc = 0.0;
    for (int64_t i = 12 * (int64_t)sboff[s] + threadIdx.x; i < 12 * (int64_t)sboff[s + 1]; i += blockDim.x)
        acc += a[i] * b[i];
    const double r = block_sum_fast(acc, sh);
    if (threadIdx.x == 0) out[s] = r;
}

// abd.py:337-353 with the step length of the body's scene
__global__ void update_q_scenes_kernel(int nb, const int32_t* __restrict__ body_scene, const double* __restrict__ A,
                                       const double* __restrict__ p, const double* __restrict__ dq,
                                       const double* __restrict__ alpha, double* A_out, double* p_out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 12 * nb) return;
    const int b = i / 12, r = i % 12;
    const double al = alpha[body_scene[b]];
    if (al == 0.0) {  // a scene that does not move keeps its state bit for bit, whatever dq holds
        if (r < 3)
            p_out[3 * b + r] = p[3 * b + r];
        else
            A_out[9 * b + r - 3] = A[9 * b + r - 3];
        return;
    }
    if (r < 3) {
        p_out[3 * b + r] = p[3 * b + r] - dq[3 * (4 * b) + r] * al;
    } else {
        const int row = (r - 3) / 3, k = (r - 3) % 3;
        A_out[9 * b + 3 * row + k] = A[9 * b + 3 * row + k] - dq[3 * (4 * b + 1 + row) + k] * al;
    }
}

}  // namespace abd

using namespace abd;

extern "C" {

static bool scenes_ok(const AbdScene* sc) {
    return sc && sc->n_scenes > 0 && sc->scene_boff && sc->body_scene;
}

int abd_scene_contact_energy(abd_stream_t stream, const AbdScene* sc, const AbdParams* P, const double* xk,
                             const int32_t* pt, int64_t npt, const int32_t* ee, int64_t nee, const int32_t* vg,
                             int64_t nvg, double* E) {
    if (!scenes_ok(sc) || !P || !xk || !E || npt < 0 || nee < 0 || nvg < 0) return ABD_ERR_ARG;
    if ((npt && !pt) || (nee && !ee) || (nvg && !vg)) return ABD_ERR_ARG;
    if (!P->pt) npt = 0;
    if (!P->ee) nee = 0;
    if (!P->vg) nvg = 0;
    hipLaunchKernelGGL(scene_energy_kernel, dim3(sc->n_scenes), dim3(kSegThreads), 0, to_stream(stream), *sc, *P, xk,
                       pt, npt, ee, nee, vg, nvg, E);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

int abd_scene_inertia_energy(abd_stream_t stream, const AbdScene* sc, const AbdParams* P, const double* Ak,
                             const double* pk, const double* A0, const double* p0, const double* Adot,
                             const double* pdot, double* E) {
    if (!scenes_ok(sc) || !P || !Ak || !pk || !A0 || !p0 || !Adot || !pdot || !E) return ABD_ERR_ARG;
    hipLaunchKernelGGL(scene_inertia_kernel, dim3(sc->n_scenes), dim3(kSegThreads), 0, to_stream(stream), *sc, *P,
                       Ak, pk, A0, p0, Adot, pdot, E);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

int abd_scene_toi(abd_stream_t stream, const AbdScene* sc, const AbdParams* P, const double* x_t0,
                  const double* x_t1, const int32_t* pt, int64_t npt, const int32_t* ee, int64_t nee,
                  const int32_t* vg, int64_t nvg, double* toi) {
    if (!scenes_ok(sc) || !P || !x_t0 || !x_t1 || !toi || npt < 0 || nee < 0 || nvg < 0) return ABD_ERR_ARG;
    if ((npt && !pt) || (nee && !ee) || (nvg && !vg)) return ABD_ERR_ARG;
    if (!P->pt) npt = 0;
    if (!P->ee) nee = 0;
    if (!P->vg) nvg = 0;
    hipLaunchKernelGGL(scene_toi_kernel, dim3(sc->n_scenes), dim3(kSegThreads), 0, to_stream(stream), *sc, *P, x_t0,
                       x_t1, pt, npt, ee, nee, vg, nvg, toi);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

int abd_scene_dot(abd_stream_t stream, const AbdScene* sc, const double* a, const double* b, double* out) {
    if (!scenes_ok(sc) || !a || !b || !out) return ABD_ERR_ARG;
    hipLaunchKernelGGL(scene_dot_kernel, dim3(sc->n_scenes), dim3(kSegThreads), 0, to_stream(stream), sc->scene_boff,
                       a, b, out);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

int abd_update_q_scenes(abd_stream_t stream, const AbdScene* sc, const double* A, const double* p, const double* dq,
                        const double* alpha, double* A_out, double* p_out) {
    if (!scenes_ok(sc) || !A || !p || !dq || !alpha || !A_out || !p_out) return ABD_ERR_ARG;
    const int nb = sc->n_bodies;
    hipLaunchKernelGGL(update_q_scenes_kernel, dim3(grid_for(12 * (int64_t)nb)), dim3(kBlock), 0, to_stream(stream),
                       nb, sc->body_scene, A, p, dq, alpha, A_out, p_out);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

}  // extern "C"
