// abd_solve.hip — Newton linear solve over the body-block sparse matrix:
// unpreconditioned BiCGSTAB (abd.py:244, warp.optim.linear.bicgstab) and the
// Jacobi-preconditioned CG of bsr_utils.py:126-235.
//
// One workgroup runs the whole Krylov loop for one system: the vectors live in
// the scratch arena (L2-resident for the scene sizes this path sees), every dot
// product is a fixed-order LDS reduction, and the stopping test is evaluated
// inside the kernel every `check_every` iterations, so a solve is one launch
// with no host round trips.
#include <hip/hip_runtime.h>

#include "abd_block.h"
#include "abd_internal.h"
#include "abd_profile.h"

namespace abd {

constexpr int kSolveThreads = 1024;
// Systems with 6 n doubles up to this size keep every Krylov vector in LDS.
constexpr size_t kLdsVecBytes = 96 * 1024;

struct Bsr {
    int nb;
    const int32_t* row_ptr;
    const int32_t* cols;
    const int32_t* slots;
    const double* blocks;
};

// out = H in; row idx of body b = idx / 12, sub-row (ii, k) = ((idx % 12) / 3, idx % 3)
__device__ __forceinline__ void spmv(const Bsr& M, const double* __restrict__ in, double* __restrict__ out, int n) {
    for (int idx = threadIdx.x; idx < n; idx += blockDim.x) {
        const int b = idx / 12, lr = idx % 12, ii = lr / 3, k = lr % 3;
        double acc = 0.0;
        for (int e = M.row_ptr[b]; e < M.row_ptr[b + 1]; ++e) {
            const int64_t s = M.slots[e];
            const double* xin = in + 12 * (int64_t)M.cols[e];
            const double* base = M.blocks + 9 * 16 * s + 3 * k;
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                const double* B = base + 9 * (ii + 4 * jj);
                acc += B[0] * xin[3 * jj] + B[1] * xin[3 * jj + 1] + B[2] * xin[3 * jj + 2];
            }
        }
        out[idx] = acc;
    }
    __syncthreads();
}

__device__ __forceinline__ double vdot(const double* __restrict__ a, const double* __restrict__ b, int n,
                                       double* sh) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += a[i] * b[i];
    return block_sum_fast(acc, sh);
}

// Vector storage: LDS (dynamic shared memory) or the scratch arena.
template <bool LDS>
__device__ __forceinline__ double* vec_base(double* ws) {
    extern __shared__ double lds_vec[];
    return LDS ? lds_vec : ws;
}

template <bool LDS>
__global__ __launch_bounds__(kSolveThreads) void bicgstab_kernel(Bsr M, const double* __restrict__ b,
                                                                  double* __restrict__ xg, double tol, int maxiter,
                                                                  int check_every, double* __restrict__ ws,
                                                                  int32_t* __restrict__ iters,
                                                                  double* __restrict__ resid) {
    __shared__ double sh[17];
    const int n = 12 * M.nb;
    double* V = vec_base<LDS>(ws);
    double* x = V;
    double* r = V + n;
    double* r0 = V + 2 * n;
    double* p = V + 3 * n;
    double* v = V + 4 * n;
    double* t = V + 5 * n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) x[i] = xg[i];
    __syncthreads();
    spmv(M, x, t, n);
    for (int i = threadIdx.x; i < n; i += blockDim.x) r[i] = b[i] - t[i];
    __syncthreads();
    double rr = vdot(r, r, n, sh);
    const double bb = vdot(b, b, n, sh);
    const double atol = tol * sqrt(bb);
    const double atol_sq = atol * atol;
    int it = 0;
    if (rr > atol_sq) {
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            r0[i] = r[i];
            p[i] = r[i];
        }
        __syncthreads();
        double rho = rr;
        for (;;) {
            for (int c = 0; c < check_every; ++c) {
                spmv(M, p, v, n);
                const double r0v = vdot(r0, v, n, sh);
                const double alpha = r0v != 0.0 ? rho / r0v : 0.0;
                for (int i = threadIdx.x; i < n; i += blockDim.x) {
                    x[i] += alpha * p[i];
                    r[i] -= alpha * v[i];
                }
                __syncthreads();
                spmv(M, r, t, n);
                const double tr = vdot(t, r, n, sh);
                const double tt = vdot(t, t, n, sh);
                const double omega = tt != 0.0 ? tr / tt : 0.0;
                for (int i = threadIdx.x; i < n; i += blockDim.x) {
                    x[i] += omega * r[i];
                    r[i] -= omega * t[i];
                }
                __syncthreads();
                rr = vdot(r, r, n, sh);
                const double rho_old = rho;
                rho = vdot(r0, r, n, sh);
                const double beta = (rho_old != 0.0 && omega != 0.0) ? (rho / rho_old) * (alpha / omega) : 0.0;
                for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = r[i] + beta * (p[i] - omega * v[i]);
                __syncthreads();
            }
            it += check_every;
            if (rr <= atol_sq || it >= maxiter) break;
        }
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) xg[i] = x[i];
    if (threadIdx.x == 0) {
        if (iters) *iters = it;
        if (resid) *resid = rr;
    }
}

__device__ __forceinline__ void jacobi(const double* __restrict__ dg, const double* __restrict__ r,
                                       double* __restrict__ z, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int i3 = i - i % 3;
        const double d0 = dg[i3], d1 = dg[i3 + 1], d2 = dg[i3 + 2];
        z[i] = (d0 * d0 + d1 * d1 + d2 * d2 == 0.0) ? r[i] : r[i] / fabs(dg[i]);
    }
    __syncthreads();
}

template <bool LDS>
__global__ __launch_bounds__(kSolveThreads) void cg_kernel(Bsr M, const double* __restrict__ b,
                                                            double* __restrict__ xg, double tol, int maxiter,
                                                            int check_every, double* __restrict__ ws,
                                                            int32_t* __restrict__ iters,
                                                            double* __restrict__ resid) {
    __shared__ double sh[17];
    const int n = 12 * M.nb;
    double* V = vec_base<LDS>(ws);
    double* x = V;
    double* r = V + n;
    double* p = V + 2 * n;
    double* Ap = V + 3 * n;
    double* z = V + 4 * n;
    double* dg = V + 5 * n;
    // bsr_get_diag: diagonal of the diagonal 3x3 blocks (sub-block (ii, ii) of slot b)
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int bb = i / 12, ii = (i % 12) / 3, a = i % 3;
        dg[i] = M.blocks[9 * (16 * (int64_t)bb + ii + 4 * ii) + 4 * a];
        x[i] = xg[i];
    }
    __syncthreads();
    spmv(M, x, Ap, n);
    for (int i = threadIdx.x; i < n; i += blockDim.x) r[i] = b[i] - Ap[i];
    __syncthreads();
    jacobi(dg, r, z, n);
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = z[i];
    __syncthreads();
    double rz_old = vdot(r, z, n, sh), rz_new = 0.0;
    const double tol_sq = tol * tol * (double)n;
    double err = rz_old;
    int end_iter = 0;
    if (err > tol_sq) {
        end_iter = maxiter;
        for (int i = 0; i < maxiter; ++i) {
            spmv(M, p, Ap, n);
            const double pAp = vdot(p, Ap, n, sh);
            const double alpha = rz_old / pAp;
            for (int k = threadIdx.x; k < n; k += blockDim.x) {
                x[k] += alpha * p[k];
                r[k] -= alpha * Ap[k];
            }
            __syncthreads();
            jacobi(dg, r, z, n);
            rz_new = vdot(r, z, n, sh);
            if (((i + 1) % check_every) == 0 && rz_new <= tol_sq) {
                end_iter = i;
                err = rz_new;
                break;
            }
            const double beta = rz_new > tol_sq ? rz_new / rz_old : 0.0;
            for (int k = threadIdx.x; k < n; k += blockDim.x) p[k] = z[k] + beta * p[k];
            __syncthreads();
            const double tmp = rz_old;
            rz_old = rz_new;
            rz_new = tmp;
            err = rz_old;
        }
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) xg[i] = x[i];
    if (threadIdx.x == 0) {
        if (iters) *iters = end_iter;
        if (resid) *resid = err;
    }
}

}  // namespace abd

using namespace abd;

extern "C" {

int abd_bsr_solve(abd_stream_t stream, int32_t method, int32_t nb, const int32_t* row_ptr, const int32_t* cols,
                  const int32_t* slots, const double* blocks, const double* b, double* x, double tol, int32_t maxiter,
                  int32_t check_every, int32_t* iters, double* resid) {
    if (nb <= 0 || !row_ptr || !cols || !slots || !blocks || !b || !x || check_every <= 0) return ABD_ERR_ARG;
    if (method != ABD_SOLVER_BICGSTAB && method != ABD_SOLVER_CG_JACOBI) return ABD_ERR_ARG;
    const int n = 12 * nb;
    if (maxiter <= 0) maxiter = n;
    double* ws;
    int rc = arena_get(kSlotSolver, sizeof(double) * 6 * (size_t)n, (void**)&ws);
    if (rc) return rc;
    Bsr M{nb, row_ptr, cols, slots, blocks};
    hipStream_t s = to_stream(stream);
    // per iteration: two block SpMVs (nnzb 12x12 blocks + column ids + row pointers + in/out
    // vectors) and ~10 vector passes; scaled by the iteration count read back at query time
    int32_t nnzb = 0;
    if (prof_enabled()) ABD_CHECK(hipMemcpy(&nnzb, row_ptr + nb, sizeof(int32_t), hipMemcpyDeviceToHost));
    const double per_it = 2.0 * (1156.0 * nnzb + 4.0 * (nb + 1) + 2.0 * 96 * nb) + 10.0 * 96 * nb;
    int ph = prof_start(kProfSolve, s, per_it);
    const size_t vec_bytes = sizeof(double) * 6 * (size_t)n;
    const bool lds = vec_bytes <= kLdsVecBytes;
    if (lds) {
        static bool attr_set = false;
        if (!attr_set) {
            hipFuncSetAttribute((const void*)bicgstab_kernel<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)kLdsVecBytes);
            hipFuncSetAttribute((const void*)cg_kernel<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)kLdsVecBytes);
            attr_set = true;
        }
        if (method == ABD_SOLVER_BICGSTAB)
            hipLaunchKernelGGL(bicgstab_kernel<true>, dim3(1), dim3(kSolveThreads), vec_bytes, s, M, b, x, tol,
                               maxiter, check_every, ws, iters, resid);
        else
            hipLaunchKernelGGL(cg_kernel<true>, dim3(1), dim3(kSolveThreads), vec_bytes, s, M, b, x, tol, maxiter,
                               check_every, ws, iters, resid);
    } else {
        if (method == ABD_SOLVER_BICGSTAB)
            hipLaunchKernelGGL(bicgstab_kernel<false>, dim3(1), dim3(kSolveThreads), 0, s, M, b, x, tol, maxiter,
                               check_every, ws, iters, resid);
        else
            hipLaunchKernelGGL(cg_kernel<false>, dim3(1), dim3(kSolveThreads), 0, s, M, b, x, tol, maxiter,
                               check_every, ws, iters, resid);
    }
    ABD_CHECK_LAUNCH();
    prof_stop(ph, s, iters);
    return ABD_OK;
}

}  // extern "C"
