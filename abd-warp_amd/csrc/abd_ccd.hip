// abd_ccd.hip — continuous collision detection (ccd.py, cubic_roots.py):
// coplanarity cubic per candidate, bracketed Newton/bisection roots on [0, 1],
// first verified root, and an order-independent global minimum.
#include <hip/hip_runtime.h>

#include "abd_block.h"
#include "abd_device.h"
#include "abd_internal.h"
#include "abd_profile.h"

namespace abd {

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
