// abd_ccd.hip — continuous collision detection (ccd.py, cubic_roots.py):
// coplanarity cubic per candidate, bracketed Newton/bisection roots on [0, 1],
// first verified root, and an order-independent global minimum.
#include <hip/hip_runtime.h>

#include "abd_block.h"
#include "abd_device.h"
#include "abd_internal.h"
#include "abd_profile.h"

namespace abd {

// ccd.py:9-40 toi_vg / toi_ee / toi_pt over [VG | EE | PT]; ccd.py:188-195 vg_collison_time.
__global__ __launch_bounds__(kBlock) void toi_partial_kernel(AbdScene sc, AbdParams P, const double* __restrict__ x0,
                                                              const double* __restrict__ x1,
                                                              const int32_t* __restrict__ pt, int64_t npt,
                                                              const int32_t* __restrict__ ee, int64_t nee,
                                                              const int32_t* __restrict__ vg, int64_t nvg,
                                                              double* __restrict__ part) {
    __shared__ double sh[kBlock];
    SceneView S;
    S.voff = sc.voff;
    S.foff = sc.foff;
    S.eoff = sc.eoff;
    S.tris = sc.tris;
    S.edges = sc.edges;
    S.x0 = sc.x0;
    const int64_t n = nvg + nee + npt;
    double acc = 1.0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const double t = i < nvg ? cand_toi_vg(sc.voff, x0, x1, vg + 2 * i, P.ground)
                         : (i < nvg + nee ? cand_toi(1, S, x0, x1, ee + 5 * (i - nvg))
                                          : cand_toi(0, S, x0, x1, pt + 5 * (i - nvg - nee)));
        if (t < 1.0 && t < acc) acc = t;  // `if t < 1: atomic_min`
    }
    double r = block_min(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}

__global__ __launch_bounds__(kBlock) void min_final_kernel(const double* __restrict__ part, int n,
                                                            double* __restrict__ out) {
    __shared__ double sh[kBlock];
    double acc = 1.0;
    for (int i = threadIdx.x; i < n; i += kBlock) acc = part[i] < acc ? part[i] : acc;
    double r = block_min(acc, sh);
    if (threadIdx.x == 0) *out = r;
}

__global__ void cubic_roots_kernel(const double* __restrict__ coef, int64_t n, double x0, double x1,
                                   int32_t* __restrict__ nroots, double* __restrict__ roots) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double c[4] = {coef[4 * i], coef[4 * i + 1], coef[4 * i + 2], coef[4 * i + 3]};
    double r[3];
    nroots[i] = cubic_roots(c, x0, x1, r);
    roots[3 * i] = r[0];
    roots[3 * i + 1] = r[1];
    roots[3 * i + 2] = r[2];
}

__global__ void collision_time_kernel(int kind, const double* __restrict__ p0, const double* __restrict__ p1,
                                      int64_t n, double* __restrict__ t) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    d3 a[4], b[4];
    for (int q = 0; q < 4; ++q) {
        a[q] = ld3(p0 + 12 * i + 3 * q);
        b[q] = ld3(p1 + 12 * i + 3 * q);
    }
    t[i] = collision_time(kind, a, b);
}

}  // namespace abd

using namespace abd;

extern "C" {

int abd_toi(abd_stream_t stream, const AbdScene* sc, const AbdParams* P, const double* x_t0, const double* x_t1,
            const int32_t* pt, int64_t npt, const int32_t* ee, int64_t nee, const int32_t* vg, int64_t nvg,
            double* toi) {
    if (!sc || !P || !x_t0 || !x_t1 || !toi || npt < 0 || nee < 0 || nvg < 0) return ABD_ERR_ARG;
    if ((npt && !pt) || (nee && !ee) || (nvg && !vg)) return ABD_ERR_ARG;
    if (!P->pt) npt = 0;  // const_params.py:42-44, abd.py:288-299
    if (!P->ee) nee = 0;
    if (!P->vg) nvg = 0;
    hipStream_t s = to_stream(stream);
    const int64_t n = npt + nee + nvg;
    const int nblk = (int)grid_for(n, kBlock, 2048);
    double* part;
    int rc = arena_get(kSlotReduce, sizeof(double) * 2048, (void**)&part);
    if (rc) return rc;
    int ph = prof_start(kProfToi, s, 224.0 * npt + 228.0 * nee + 56.0 * nvg);
    hipLaunchKernelGGL(toi_partial_kernel, dim3(nblk), dim3(kBlock), 0, s, *sc, *P, x_t0, x_t1, pt, npt, ee, nee, vg,
                       nvg, part);
    ABD_CHECK_LAUNCH();
    prof_stop(ph, s);
    hipLaunchKernelGGL(min_final_kernel, dim3(1), dim3(kBlock), 0, s, part, nblk, toi);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

int abd_cubic_roots(abd_stream_t stream, const double* coef, int64_t n, double x0, double x1, int32_t* nroots,
                    double* roots) {
    if (n < 0 || (n && (!coef || !nroots || !roots))) return ABD_ERR_ARG;
    if (n == 0) return ABD_OK;
    hipLaunchKernelGGL(cubic_roots_kernel, dim3(grid_for(n, kBlock, INT32_MAX)), dim3(kBlock), 0, to_stream(stream),
                       coef, n, x0, x1, nroots, roots);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

int abd_collision_time(abd_stream_t stream, int32_t kind, const double* p_t0, const double* p_t1, int64_t n,
                       double* t) {
    if (n < 0 || (kind != 0 && kind != 1) || (n && (!p_t0 || !p_t1 || !t))) return ABD_ERR_ARG;
    if (n == 0) return ABD_OK;
    hipLaunchKernelGGL(collision_time_kernel, dim3(grid_for(n, kBlock, INT32_MAX)), dim3(kBlock), 0,
                       to_stream(stream), kind, p_t0, p_t1, n, t);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

}  // extern "C"
