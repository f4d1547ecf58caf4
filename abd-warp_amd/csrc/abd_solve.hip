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

#include <algorithm>
#include <mutex>

#include "abd_block.h"
#include "abd_internal.h"
#include "abd_profile.h"

namespace abd {

constexpr int kSolveThreads = 1024;
// The direct block LDL^T (ldlt_kernel) takes systems of up to kLdltMaxBodies bodies; it
// marks a system it declines with kLdltFailed in the iteration count, and the block-Jacobi
// CG that follows it solves exactly those.
constexpr int kLdltMaxBodies = 64;
constexpr int kLdltFailed = -2;
// LDS for a system's staged block structure (row offsets and (column, slot) entries).
constexpr size_t kLdsStructBytes = 24 * 1024;

struct Bsr {
    int nb;
    const int32_t* row_ptr;
    const int32_t* cols;
    const int32_t* slots;
    const double* blocks;
};

// One 12x12 block's row (ii, k) against 12 entries of a vector: the three-term sums of
// sub-blocks (ii, jj), jj = 0..3, added to acc in jj order.
__device__ __forceinline__ double block_row_add(double acc, const double (&B)[12], const double* __restrict__ xin) {
#pragma unroll
    for (int jj = 0; jj < 4; ++jj)
        acc += B[3 * jj] * xin[3 * jj] + B[3 * jj + 1] * xin[3 * jj + 1] + B[3 * jj + 2] * xin[3 * jj + 2];
    return acc;
}

__device__ __forceinline__ void load_block_row(const double* __restrict__ blocks, int64_t slot, int ii, int k,
                                               double (&B)[12]) {
    const double* base = blocks + 9 * 16 * slot + 3 * k;
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
        const double* Bj = base + 9 * (ii + 4 * jj);
        B[3 * jj] = Bj[0];
        B[3 * jj + 1] = Bj[1];
        B[3 * jj + 2] = Bj[2];
    }
}

// A system's block structure as its SpMV reads it: from the global CSR (columns outside
// the system filtered on the fly) or from an LDS copy made once per solve by stage_rows
// (local row offsets lrp, entries ent = (local column or -1, slot)), which takes two
// dependent global loads off every block the SpMV visits.
struct Rows {
    int row0, nbs;
    const int* lrp;
    const int2* ent;
    double* part;  // LDS partial products of the entry-parallel SpMV (12 per block)
};

// SpMV modes: rows over the global structure, rows over the staged structure, and
// entry-parallel over the staged structure (systems that fit the LDS budget, entries_fit).
// The first two sum each row over (block, column triple) in one sequence; the third sums
// each block's row first and the block results in column order.  The mode of a system
// depends on its own size only, so a scene gives the same bits alone or in a batch.
enum { kSpmvRows = 0, kSpmvRowsStaged = 1, kSpmvEntries = 2 };

// lrp capacity of the staged structure (bodies + 1)
constexpr int kLrpCap = 2048;
constexpr size_t kLrpBytes = sizeof(int) * (size_t)kLrpCap;
constexpr int kEntCap = (int)((kLdsStructBytes - kLrpBytes) / sizeof(int2));

// Does a system of n rows and nnz blocks take the entry-parallel SpMV?  Its six vectors
// and its partial products must fit the LDS budget beside the structure, counting the
// vectors as if in LDS wherever they live, so that the answer depends on the system
// alone and on the device, never on the other systems of a launch.
__device__ __forceinline__ bool entries_fit(int n, int nnz, int budget) {
    return budget >= 0 && 8 * (6 * (int64_t)n + 12 * (int64_t)nnz) + (int64_t)kLdsStructBytes <= budget;
}

template <int MODE>
__device__ __forceinline__ void row_entry(const Bsr& M, const Rows& R, int e, int& lc, int64_t& slot) {
    if (MODE != kSpmvRows) {
        const int2 v = R.ent[e];
        lc = v.x;
        slot = v.y;
    } else {
        lc = M.cols[e] - R.row0;
        if ((unsigned)lc >= (unsigned)R.nbs) lc = -1;  // a block coupling two scenes is part of neither
        slot = M.slots[e];
    }
}

// (H in)[row (lb, ii, k)] of the system (local body lb), summed over the row's blocks in
// column order.  Two blocks are loaded per step so that their memory latencies overlap;
// the sum order is that of a one-block loop.
template <int MODE>
__device__ __forceinline__ double row_product(const Bsr& M, const Rows& R, int lb, int ii, int k,
                                              const double* __restrict__ in) {
    double acc = 0.0;
    int e = MODE != kSpmvRows ? R.lrp[lb] : M.row_ptr[R.row0 + lb];
    const int e1 = MODE != kSpmvRows ? R.lrp[lb + 1] : M.row_ptr[R.row0 + lb + 1];
    for (; e + 1 < e1; e += 2) {
        int lc0, lc1;
        int64_t s0, s1;
        row_entry<MODE>(M, R, e, lc0, s0);
        row_entry<MODE>(M, R, e + 1, lc1, s1);
        double B0[12], B1[12];
        load_block_row(M.blocks, s0, ii, k, B0);
        load_block_row(M.blocks, s1, ii, k, B1);
        if (lc0 >= 0) acc = block_row_add(acc, B0, in + 12 * (int64_t)lc0);
        if (lc1 >= 0) acc = block_row_add(acc, B1, in + 12 * (int64_t)lc1);
    }
    if (e < e1) {
        int lc;
        int64_t sl;
        row_entry<MODE>(M, R, e, lc, sl);
        if (lc >= 0) {
            double B[12];
            load_block_row(M.blocks, sl, ii, k, B);
            acc = block_row_add(acc, B, in + 12 * (int64_t)lc);
        }
    }
    return acc;
}

// Entry-parallel products: work item (block e, block row ii) multiplies rows 3 ii .. 3 ii + 2
// of block e by its column's 12 entries of `in` (each row's four sub-block sums in
// order) into R.part[12 e + 3 ii + r]; every load of an item is independent of the others.
__device__ __forceinline__ void entry_products(const Bsr& M, const Rows& R, const double* __restrict__ in) {
    const int nnz = R.lrp[R.nbs];
    for (int wi = threadIdx.x; wi < 4 * nnz; wi += blockDim.x) {
        const int e = wi >> 2, ii = wi & 3;
        const int2 en = R.ent[e];
        if (en.x < 0) continue;  // outside the system: skipped by the row sums too
        const double* B = M.blocks + 144 * (int64_t)en.y + 9 * ii;
        const double* xin = in + 12 * en.x;
        double y0 = 0.0, y1 = 0.0, y2 = 0.0;
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
            const double* Bj = B + 36 * jj;  // sub-block (ii, jj), row-major 3x3
            const double x0 = xin[3 * jj], x1 = xin[3 * jj + 1], x2 = xin[3 * jj + 2];
            y0 = y0 + (Bj[0] * x0 + Bj[1] * x1 + Bj[2] * x2);
            y1 = y1 + (Bj[3] * x0 + Bj[4] * x1 + Bj[5] * x2);
            y2 = y2 + (Bj[6] * x0 + Bj[7] * x1 + Bj[8] * x2);
        }
        double* pe = R.part + 12 * e + 3 * ii;
        pe[0] = y0;
        pe[1] = y1;
        pe[2] = y2;
    }
    __syncthreads();
}

// row idx of the entry-parallel SpMV: its blocks' partial products in column order
__device__ __forceinline__ double entry_row_sum(const Rows& R, int idx) {
    const int lb = idx / 12, lr = idx % 12;
    double acc = 0.0;
    for (int e = R.lrp[lb]; e < R.lrp[lb + 1]; ++e)
        if (R.ent[e].x >= 0) acc += R.part[12 * e + lr];
    return acc;
}

// Copies the structure of the system's rows into LDS (lrp: kLrpCap ints, ent: kEntCap
// entries) when the device holds it (budget >= 0) and it fits; returns the SpMV mode of
// the system.  Uniform over the workgroup; ends with a barrier when it stages.
__device__ __forceinline__ int stage_rows(const Bsr& M, Rows& R, int* lrp, int2* ent, int budget) {
    const int e0 = M.row_ptr[R.row0], nnz = M.row_ptr[R.row0 + R.nbs] - e0;
    if (budget < 0 || nnz > kEntCap || R.nbs + 1 > kLrpCap) return kSpmvRows;
    for (int r = threadIdx.x; r <= R.nbs; r += blockDim.x) lrp[r] = M.row_ptr[R.row0 + r] - e0;
    for (int e = threadIdx.x; e < nnz; e += blockDim.x) {
        const int lc = M.cols[e0 + e] - R.row0;
        ent[e] = make_int2((unsigned)lc < (unsigned)R.nbs ? lc : -1, M.slots[e0 + e]);
    }
    __syncthreads();
    R.lrp = lrp;
    R.ent = ent;
    return entries_fit(12 * R.nbs, nnz, budget) ? kSpmvEntries : kSpmvRowsStaged;
}

// out = H in over the system's rows (in / out indexed from its first row); row idx of
// local body idx / 12, sub-row (ii, k) = ((idx % 12) / 3, idx % 3)
template <int MODE>
__device__ __forceinline__ void spmv(const Bsr& M, const Rows& R, const double* __restrict__ in,
                                     double* __restrict__ out, int n) {
    if (MODE == kSpmvEntries) entry_products(M, R, in);
    for (int idx = threadIdx.x; idx < n; idx += blockDim.x) {
        const int lr = idx % 12;
        out[idx] = MODE == kSpmvEntries ? entry_row_sum(R, idx) : row_product<MODE>(M, R, idx / 12, lr / 3, lr % 3, in);
    }
    __syncthreads();
}

// out = H in over this thread's rows (the rows spmv() gives it), returning the
// thread's partial sum of w[idx] * out[idx] in row order; no barrier at the end: the
// caller reduces the partials (which also orders the writes before any cross-thread read).
template <int MODE>
__device__ __forceinline__ double spmv_rows_dot(const Bsr& M, const Rows& R, const double* __restrict__ in,
                                                double* __restrict__ out, int n, const double* __restrict__ w) {
    if (MODE == kSpmvEntries) entry_products(M, R, in);
    double part = 0.0;
    for (int idx = threadIdx.x; idx < n; idx += blockDim.x) {
        const int lr = idx % 12;
        const double acc =
            MODE == kSpmvEntries ? entry_row_sum(R, idx) : row_product<MODE>(M, R, idx / 12, lr / 3, lr % 3, in);
        out[idx] = acc;
        part += w[idx] * acc;
    }
    return part;
}

__device__ __forceinline__ double vdot(const double* __restrict__ a, const double* __restrict__ b, int n,
                                       double* sh) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += a[i] * b[i];
    return block_sum_fast(acc, sh);
}

// Dynamic LDS of a solve workgroup: the staged structure (lrp, then ent), the six Krylov
// vectors of the system's n rows when LDS (else in the scratch arena), then the partial
// products of the entry-parallel SpMV.
template <bool LDS>
struct SolveLds {
    double* V;
    double* part;
    int* lrp;
    int2* ent;
    __device__ SolveLds(double* ws, int n) {
        extern __shared__ double lds_dyn[];
        char* base = (char*)lds_dyn;
        lrp = (int*)base;
        ent = (int2*)(base + kLrpBytes);
        double* after = (double*)(base + kLdsStructBytes);
        V = LDS ? after : ws;
        part = LDS ? after + 6 * n : after;
    }
};

template <int MODE>
__device__ __forceinline__ void bicgstab_body(const Bsr& M, const Rows& R, int n, const double* __restrict__ b,
                                              double* __restrict__ xg, double tol, int maxiter, int check_every,
                                              double* V, int32_t* __restrict__ iters, double* __restrict__ resid,
                                              double* sh, double* sh2) {
    double* x = V;
    double* r = V + n;
    double* r0 = V + 2 * n;
    double* p = V + 3 * n;
    double* v = V + 4 * n;
    double* t = V + 5 * n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) x[i] = xg[i];
    __syncthreads();
    spmv<MODE>(M, R, x, t, n);
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
            // fused passes: each dot is accumulated where its operand is produced, in
            // the row order of vdot(), so the iterates match the unfused sequence bit for bit
            for (int c = 0; c < check_every; ++c) {
                const double r0v = block_sum_fast(spmv_rows_dot<MODE>(M, R, p, v, n, r0), sh);
                const double alpha = r0v != 0.0 ? rho / r0v : 0.0;
                for (int i = threadIdx.x; i < n; i += blockDim.x) {
                    x[i] += alpha * p[i];
                    r[i] -= alpha * v[i];
                }
                __syncthreads();
                // t = H r, then t.r and t.t over this thread's rows in row order
                spmv<MODE>(M, R, r, t, n);
                double tr = 0.0, tt = 0.0;
                for (int idx = threadIdx.x; idx < n; idx += blockDim.x) {
                    const double acc = t[idx];
                    tr += acc * r[idx];
                    tt += acc * acc;
                }
                block_sum2_fast(tr, tt, sh2);
                const double omega = tt != 0.0 ? tr / tt : 0.0;
                double rr_p = 0.0, rho_p = 0.0;
                for (int i = threadIdx.x; i < n; i += blockDim.x) {
                    x[i] += omega * r[i];
                    r[i] -= omega * t[i];
                    rr_p += r[i] * r[i];
                    rho_p += r0[i] * r[i];
                }
                block_sum2_fast(rr_p, rho_p, sh2);
                rr = rr_p;
                const double rho_old = rho;
                rho = rho_p;
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

// budget: bytes of dynamic LDS (SolveLds; -1: the device holds no staged structure).
template <bool LDS>
__global__ __launch_bounds__(kSolveThreads) void bicgstab_kernel(Bsr M, const int32_t* __restrict__ sboff,
                                                                  const uint8_t* __restrict__ active,
                                                                  const double* __restrict__ b,
                                                                  double* __restrict__ xg, double tol, int maxiter,
                                                                  int check_every, int cap_rows, int budget,
                                                                  double* __restrict__ ws,
                                                                  int32_t* __restrict__ iters,
                                                                  double* __restrict__ resid) {
    __shared__ double sh[17];
    __shared__ double sh2[34];
    // one system per workgroup: the bodies of scene blockIdx.x (all bodies without scenes)
    const int sidx = blockIdx.x;
    if (active && !active[sidx]) return;
    const int row0 = sboff ? sboff[sidx] : 0;
    const int n = 12 * ((sboff ? sboff[sidx + 1] : M.nb) - row0);
    b += 12 * (int64_t)row0;
    xg += 12 * (int64_t)row0;
    if (iters) iters += sidx;
    if (resid) resid += sidx;
    if (n > cap_rows) {  // larger than the vector storage sized by the caller: report, touch nothing
        if (threadIdx.x == 0 && iters) *iters = -1;
        return;
    }
    const SolveLds<LDS> S(ws + (LDS ? 0 : 6 * 12 * (int64_t)row0), n);
    if (maxiter <= 0) maxiter = n;  // the system size of this workgroup's system
    Rows R{row0, n / 12, nullptr, nullptr, S.part};
    const int mode = stage_rows(M, R, S.lrp, S.ent, budget);
    if (mode == kSpmvEntries)
        bicgstab_body<kSpmvEntries>(M, R, n, b, xg, tol, maxiter, check_every, S.V, iters, resid, sh, sh2);
    else if (mode == kSpmvRowsStaged)
        bicgstab_body<kSpmvRowsStaged>(M, R, n, b, xg, tol, maxiter, check_every, S.V, iters, resid, sh, sh2);
    else
        bicgstab_body<kSpmvRows>(M, R, n, b, xg, tol, maxiter, check_every, S.V, iters, resid, sh, sh2);
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

// Inverse of each body's 12x12 diagonal block (slot b) by Gauss-Jordan elimination with
// partial pivoting, written to dinv + 144 b; a singular block gives the identity (that
// body stays unpreconditioned).  One 64-thread workgroup per body: the augmented 12x24
// matrix in LDS, the 288 entries of each elimination step updated in parallel, each with
// the arithmetic of the sequential elimination (row c scaled by 1 / pivot, then
// a[r][j] -= a[r][c] a[c][j] for every other row with a[r][c] != 0).
// pending (nullable): per-system status of a direct solve; then only the bodies of the
// systems it declined (kLdltFailed) are inverted, the system of body b found in sboff [nsys + 1].
constexpr int kInvThreads = 64;
__global__ __launch_bounds__(kInvThreads) void invert_blocks_kernel(Bsr M, double* __restrict__ dinv,
                                                                    const int32_t* __restrict__ pending,
                                                                    const int32_t* __restrict__ sboff, int nsys) {
    __shared__ double a[12][24];
    __shared__ double srow[24];
    __shared__ double f[12];
    const int b = blockIdx.x, t = threadIdx.x;
    if (pending) {
        int sys = 0;
        if (sboff) {
            int lo = 0, hi = nsys - 1;  // the last system whose first body is <= b
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (sboff[mid] <= b)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            sys = lo;
        }
        if (pending[sys] != kLdltFailed) return;  // uniform over the workgroup
    }
    for (int e = t; e < 288; e += kInvThreads) {
        const int i = e / 24, j = e % 24;
        a[i][j] = j < 12 ? M.blocks[9 * (16 * (int64_t)b + i / 3 + 4 * (j / 3)) + 3 * (i % 3) + j % 3]
                         : (i == j - 12 ? 1.0 : 0.0);
    }
    __syncthreads();
    bool ok = true;
    for (int c = 0; c < 12; ++c) {
        int piv = c;  // the first row of largest |a[r][c]|, r >= c (every thread alike)
        for (int r = c + 1; r < 12; ++r)
            if (fabs(a[r][c]) > fabs(a[piv][c])) piv = r;
        if (a[piv][c] == 0.0) {
            ok = false;
            break;
        }
        __syncthreads();
        if (piv != c && t < 24) {
            const double u = a[c][t];
            a[c][t] = a[piv][t];
            a[piv][t] = u;
        }
        __syncthreads();
        const double inv = 1.0 / a[c][c];
        if (t < 24) srow[t] = a[c][t] * inv;
        if (t < 12) f[t] = a[t][c];
        __syncthreads();
        for (int e = t; e < 288; e += kInvThreads) {
            const int i = e / 24, j = e % 24;
            if (i == c)
                a[i][j] = srow[j];
            else if (f[i] != 0.0)
                a[i][j] -= f[i] * srow[j];
        }
        __syncthreads();
    }
    double* out = dinv + 144 * (int64_t)b;
    for (int e = t; e < 144; e += kInvThreads) {
        const int i = e / 12, j = e % 12;
        out[e] = ok ? a[i][12 + j] : (i == j ? 1.0 : 0.0);
    }
}

// z = D_b^-1 r for row k of a system whose body-diagonal inverses are in dinv (global body ids)
__device__ __forceinline__ double block_precond(const double* __restrict__ dinv, const double* __restrict__ r,
                                                int row0, int k) {
    const int lb = k / 12, l = k % 12;
    const double* Di = dinv + 144 * (int64_t)(row0 + lb) + 12 * l;
    const double* rb = r + 12 * lb;
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < 12; ++q) acc += Di[q] * rb[q];
    return acc;
}

template <bool BJ, int MODE>
__device__ __forceinline__ void cg_body(const Bsr& M, const Rows& R, int n, const double* __restrict__ b,
                                        double* __restrict__ xg, double tol, int maxiter, int check_every,
                                        double* V, const double* __restrict__ dinv, int32_t* __restrict__ iters,
                                        double* __restrict__ resid, double* sh) {
    const int row0 = R.row0;
    double* x = V;
    double* r = V + n;
    double* p = V + 2 * n;
    double* Ap = V + 3 * n;
    double* z = V + 4 * n;
    double* dg = V + 5 * n;
    // bsr_get_diag: diagonal of the diagonal 3x3 blocks (sub-block (ii, ii) of slot b)
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int bb = row0 + i / 12, ii = (i % 12) / 3, a = i % 3;
        dg[i] = M.blocks[9 * (16 * (int64_t)bb + ii + 4 * ii) + 4 * a];
        x[i] = xg[i];
    }
    __syncthreads();
    spmv<MODE>(M, R, x, Ap, n);
    for (int i = threadIdx.x; i < n; i += blockDim.x) r[i] = b[i] - Ap[i];
    __syncthreads();
    if (BJ) {
        for (int i = threadIdx.x; i < n; i += blockDim.x) z[i] = block_precond(dinv, r, row0, i);
        __syncthreads();
    } else {
        jacobi(dg, r, z, n);
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = z[i];
    __syncthreads();
    double rz_old = vdot(r, z, n, sh), rz_new = 0.0;
    const double tol_sq = tol * tol * (double)n;
    double err = rz_old;
    int end_iter = 0;
    if (err > tol_sq) {
        end_iter = maxiter;
        // fused passes (dots accumulated where their operands are produced, in vdot()'s
        // row order: the iterates match the unfused sequence bit for bit)
        for (int i = 0; i < maxiter; ++i) {
            const double pAp = block_sum_fast(spmv_rows_dot<MODE>(M, R, p, Ap, n, p), sh);
            // unguarded in the reference: p = 0 once the residual vanishes exactly between checks
            const double alpha = pAp != 0.0 ? rz_old / pAp : 0.0;
            double rz_p = 0.0;
            if (BJ) {
                // the body's whole residual enters z: update r, then precondition
                for (int k = threadIdx.x; k < n; k += blockDim.x) {
                    x[k] += alpha * p[k];
                    r[k] -= alpha * Ap[k];
                }
                __syncthreads();
                for (int k = threadIdx.x; k < n; k += blockDim.x) {
                    z[k] = block_precond(dinv, r, row0, k);
                    rz_p += r[k] * z[k];
                }
            } else {
                for (int k = threadIdx.x; k < n; k += blockDim.x) {
                    x[k] += alpha * p[k];
                    r[k] -= alpha * Ap[k];
                    const int k3 = k - k % 3;
                    const double d0 = dg[k3], d1 = dg[k3 + 1], d2 = dg[k3 + 2];
                    z[k] = (d0 * d0 + d1 * d1 + d2 * d2 == 0.0) ? r[k] : r[k] / fabs(dg[k]);
                    rz_p += r[k] * z[k];
                }
            }
            rz_new = block_sum_fast(rz_p, sh);
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

// BJ: the per-body 12x12 block-Jacobi preconditioner (beyond the reference) in place of
// the diagonal one; dinv: [nb][144] workspace of the inverses.  budget as bicgstab_kernel.
template <bool LDS, bool BJ>
__global__ __launch_bounds__(kSolveThreads) void cg_kernel(Bsr M, const int32_t* __restrict__ sboff,
                                                           const uint8_t* __restrict__ active,
                                                           const double* __restrict__ b, double* __restrict__ xg,
                                                           double tol, int maxiter, int check_every, int cap_rows,
                                                           int budget, double* __restrict__ ws,
                                                           double* __restrict__ dinv, int32_t* __restrict__ iters,
                                                           double* __restrict__ resid,
                                                           const int32_t* __restrict__ pending) {
    __shared__ double sh[17];
    // one system per workgroup: the bodies of scene blockIdx.x (all bodies without scenes);
    // pending (nullable): only the systems the direct solve declined
    const int sidx = blockIdx.x;
    if (active && !active[sidx]) return;
    if (pending && pending[sidx] != kLdltFailed) return;
    const int row0 = sboff ? sboff[sidx] : 0;
    const int n = 12 * ((sboff ? sboff[sidx + 1] : M.nb) - row0);
    b += 12 * (int64_t)row0;
    xg += 12 * (int64_t)row0;
    if (iters) iters += sidx;
    if (resid) resid += sidx;
    if (n > cap_rows) {  // larger than the vector storage sized by the caller: report, touch nothing
        if (threadIdx.x == 0 && iters) *iters = -1;
        return;
    }
    const SolveLds<LDS> S(ws + (LDS ? 0 : 6 * 12 * (int64_t)row0), n);
    if (maxiter <= 0) maxiter = n;  // the system size of this workgroup's system
    Rows R{row0, n / 12, nullptr, nullptr, S.part};
    const int mode = stage_rows(M, R, S.lrp, S.ent, budget);
    if (mode == kSpmvEntries)
        cg_body<BJ, kSpmvEntries>(M, R, n, b, xg, tol, maxiter, check_every, S.V, dinv, iters, resid, sh);
    else if (mode == kSpmvRowsStaged)
        cg_body<BJ, kSpmvRowsStaged>(M, R, n, b, xg, tol, maxiter, check_every, S.V, dinv, iters, resid, sh);
    else
        cg_body<BJ, kSpmvRows>(M, R, n, b, xg, tol, maxiter, check_every, S.V, dinv, iters, resid, sh);
}

// ------------------------------------------------------------------ direct solve
// Beyond the reference: the system solved directly by a block LDL^T factorization of its
// 12x12 body blocks (no exchanges between bodies), fill-in kept; one workgroup per system
// of at most kLdltMaxBodies bodies.  Structure: one 64-bit upper row mask per body, its
// fill (eliminating the bodies in their own order), the transposed masks, and the levels
// of that order's elimination tree (parent of k: the first body above k coupled to it;
// level: 0 for a leaf, else one more than the highest child's) -- bodies of one level are
// never coupled to each other and are eliminated together, level after level, in chunks
// of at most kLdltCols augmented columns.  The factor's upper blocks live in the system's
// workspace region (row i's blocks from pos[i], in column order), the chunk's W blocks
// after them.  Elimination of body k: Gauss-Jordan elimination without row exchanges on
// [P | A'_k,j1 .. A'_k,jd | I], P = A'_kk, one thread per column (12 doubles in
// registers; the chunk's pivot columns broadcast through LDS, one barrier per pivot
// column), leaving W_j = P^-1 A'_kj and P^-1 (which replaces P in the factor).  Each
// Schur target A'_ij (and each forward-solve target y_i) is updated by its chunk's bodies
// in ascending order, A'_ij -= A'_ki^T W_j, y_i -= W_i^T y_k (12-term sums), by the
// thread of the lowest of them, so a system gives the same bits whatever else a launch
// holds; z_k = P^-1 y_k; W then replaces row k.  Backward solve, levels from the top:
// x_k = z_k - sum_i W_i x_i, each W_i x_i a 12-term sum, added in ascending i.  This is
// the arithmetic, in the same order, of eliminating body by body in (level, body) order.
// A pivot that is not positive (the system is not positive definite as factored) or a
// system above the size limit leaves x untouched and status kLdltFailed; else status 1
// and resid (nullable) = |b - H x|^2.  The factor and the W blocks live in the workgroup's dynamic
// LDS (dyn_blocks 12x12 blocks) when they fit there -- the factor's upper blocks with fill
// and at most min(kLdltWBlocks, off-diagonal blocks) W blocks -- else in the system's
// workspace region; the arithmetic is the same in both places.
constexpr int kLdltColsPerThread = 2;  // adjacent columns 2t, 2t + 1 of one body
constexpr int kLdltCols = kLdltColsPerThread * kSolveThreads;
constexpr int kLdltWBlocks = kLdltCols / 12;  // bound on a chunk's W blocks
__host__ __device__ inline int64_t ldlt_region(int m) { return 144 * ((int64_t)m * (m + 1) / 2 + kLdltWBlocks); }

// the largest index r in [0, n) with v[r] <= key (v ascending, v[0] <= key)
__device__ __forceinline__ int last_le(const int* v, int n, int key) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (v[mid] <= key) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ unsigned long long bits_above(unsigned long long row, int k) {
    return row & ~((2ull << k) - 1ull);
}
__device__ __forceinline__ unsigned long long bits_below(int j) { return (1ull << j) - 1ull; }

__global__ __launch_bounds__(kSolveThreads) void ldlt_kernel(Bsr M, const int32_t* __restrict__ sboff,
                                                             const uint8_t* __restrict__ active,
                                                             const double* __restrict__ b, double* __restrict__ xg,
                                                             int64_t region, double* __restrict__ ws,
                                                             int32_t* __restrict__ status, double* __restrict__ resid,
                                                             int dyn_blocks) {
    constexpr int MB = kLdltMaxBodies;
    extern __shared__ double ldlt_dyn[];
    __shared__ unsigned long long adj[MB];   // upper row pattern with fill (bits >= row)
    __shared__ unsigned long long colm[MB];  // colm[i]: the rows k < i whose pattern holds i
    __shared__ unsigned long long levm[MB];  // the bodies of each level
    __shared__ int lvl[MB];                  // the level of each body
    __shared__ unsigned char mlist[MB][MB];  // mlist[k]: the columns above k in row k, ascending
    __shared__ int lrp[MB + 1], pos[MB + 1];
    // the system's CSR entries (slot, local column: 0xff outside the system; at most m^2), read
    // only while the factor is filled, share their LDS with the elimination's pivot columns
    // (two per step, double-buffered) and reciprocal pivots
    constexpr int kEntryBytes = MB * MB * 5, kPivotBytes = MB * 2 * 2 * 13 * 8;
    __shared__ __align__(16) unsigned char tables[kEntryBytes > kPivotBytes ? kEntryBytes : kPivotBytes];
    int* eslot = reinterpret_cast<int*>(tables);
    unsigned char* ecol = tables + 4 * MB * MB;
    double(*colc)[2][2][12] = reinterpret_cast<double(*)[2][2][12]>(tables);
    double(*colinv)[2][2] = reinterpret_cast<double(*)[2][2]>(tables + MB * 2 * 2 * 12 * 8);
    __shared__ int cnode[MB], cstart[MB + 1], sstart[MB + 1], woff[MB];  // the chunk
    __shared__ int nlev, fail, nchunk;
    __shared__ unsigned long long chunkmask, restm;  // the chunk; the level's bodies not yet eliminated
    __shared__ double y[12 * MB];  // b, then the forward solve, then x
    __shared__ double z[12 * MB];
    __shared__ double sh[17];
    const int sidx = blockIdx.x, t = threadIdx.x, nt = blockDim.x;
    if (active && !active[sidx]) return;
    const int row0 = sboff ? sboff[sidx] : 0;
    const int m = (sboff ? sboff[sidx + 1] : M.nb) - row0;
    if (m > MB) {
        if (t == 0) status[sidx] = kLdltFailed;
        return;
    }
    const int n = 12 * m;
    b += 12 * (int64_t)row0;
    xg += 12 * (int64_t)row0;
    double* U = ws + region * sidx;  // in LDS instead when the factor fits there (below)
    double* W = U + 144 * ((int64_t)m * (m + 1) / 2);
    const int e0 = M.row_ptr[row0];
    if (t <= m) lrp[t] = M.row_ptr[row0 + t] - e0;
    if (t < MB) {
        adj[t] = t < m ? 1ull << t : 0ull;
        levm[t] = 0ull;
    }
    if (t == 0) {
        nlev = 0;
        fail = 0;
    }
    for (int i = t; i < n; i += nt) y[i] = b[i];
    __syncthreads();
    const int nnz = lrp[m];
    if (nnz > MB * MB) {  // more entries than m^2 (blocks coupling other systems): declined
        if (t == 0) status[sidx] = kLdltFailed;
        return;
    }
    for (int e = t; e < nnz; e += nt) {
        const int j = M.cols[e0 + e] - row0;
        const int i = last_le(lrp, m, e);
        eslot[e] = M.slots[e0 + e];
        ecol[e] = (unsigned)j < (unsigned)m ? (unsigned char)j : 0xff;
        if (j > i && j < m) atomicOr(&adj[i], 1ull << j);
    }
    __syncthreads();
    if (t < 64) {  // wave 0, lane = row: fill, transposed pattern and tree levels, row by row
        unsigned long long mine = t < m ? adj[t] : 0ull, cm = 0ull;
        int h = 0;
        // only rows with entries above the diagonal change anything (fill, colm, a parent's
        // level); fill reaches only later rows, so the rows left are found again after each
        unsigned long long todo = __ballot(t < m && bits_above(mine, t) != 0ull);
        while (todo) {  // uniform
            const int k = __builtin_ctzll(todo);
            // row k is final once k is reached, and so is its level (lane k's values,
            // read by the uniform lane index)
            const unsigned lo32 = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)mine, k);
            const unsigned hi32 = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(mine >> 32), k);
            const unsigned long long ak = ((unsigned long long)hi32 << 32) | lo32;
            const int hk = __builtin_amdgcn_readlane(h, k);
            const unsigned long long above = bits_above(ak, k);
            if ((above >> t) & 1ull) {
                mine |= ak & (~0ull << t);
                cm |= 1ull << k;
            }
            if (above && t == __builtin_ctzll(above) && h < hk + 1) h = hk + 1;
            todo = __ballot(t < m && t > k && bits_above(mine, t) != 0ull);
        }
        if (t < m) {
            adj[t] = mine;
            colm[t] = cm;
            lvl[t] = h;
            atomicOr(&levm[h], 1ull << t);
            atomicMax(&nlev, h + 1);
        }
    }
    __syncthreads();
    if (t < m) {
        int idx = 0;
        for (unsigned long long bt = bits_above(adj[t], t); bt; bt &= bt - 1ull) mlist[t][idx++] = __builtin_ctzll(bt);
    }
    if (t < kWave) {  // wave 0: pos = the exclusive prefix sum of the rows' block counts
        const int cnt = t < m ? __popcll(adj[t]) : 0;
        int incl = cnt;
        for (int off = 1; off < kWave; off <<= 1) {
            const int v = __shfl_up(incl, off, kWave);
            if (t >= off) incl += v;
        }
        if (t < m) pos[t] = incl - cnt;
        const int acc = __shfl(incl, kWave - 1, kWave);
        if (t == 0) {
            pos[m] = acc;
            const int offd = acc - m, wmax = offd < kLdltWBlocks ? offd : kLdltWBlocks;
            fail = acc + wmax <= dyn_blocks ? 2 : 0;  // 2: the factor in LDS (reset below)
        }
    }
    __syncthreads();
    const bool in_lds = fail == 2;  // uniform
    __syncthreads();
    if (t == 0) fail = 0;
    // The factorization and both solves, instantiated for a factor in LDS (its accesses then
    // LDS instructions) and for one in the global workspace; false: a pivot was not positive.
    auto factor_solve = [&](auto U, auto W) __attribute__((always_inline)) -> bool {
        for (int64_t e = t; e < 144 * (int64_t)pos[m]; e += nt) U[e] = 0.0;
        __syncthreads();
        // the matrix's upper blocks into the factor (row-major 12x12): one wave per block, its
        // lanes reading the slot's 144 doubles in storage order (16 3x3 sub-blocks, sub-block
        // (R/3, C/3) at 9 (R/3 + 4 C/3), row-major inside)
        {
            const int lane = t & (kWave - 1), wv = t / kWave, nwv = nt / kWave;
            for (int e = wv; e < nnz; e += nwv) {  // uniform over the wave
                const int j = ecol[e];
                const int i = last_le(lrp, m, e);
                if (j < i || j >= m) continue;  // lower, or outside the system (0xff)
                const double* src = M.blocks + 144 * (int64_t)eslot[e];
                auto dst = U + 144 * (int64_t)(pos[i] + __popcll(adj[i] & bits_below(j)));
                for (int q = lane; q < 144; q += kWave) {
                    const int sb = q / 9, w9 = q % 9;
                    dst[12 * (3 * (sb % 4) + w9 / 3) + 3 * (sb / 4) + w9 % 3] = src[q];
                }
            }
        }
        const int levels = nlev;
        for (int L = 0; L < levels; ++L) {
            if (t == 0) restm = levm[L];
            for (;;) {
                __syncthreads();  // the previous chunk is written back, its tables are free
                if (t < kWave) {
                    // wave 0, lane = body: the chunk is the lowest bodies of the level not yet
                    // eliminated whose augmented columns fit kLdltCols (prefix sums in body order;
                    // a single body always fits)
                    const unsigned long long rest = restm;
                    const bool member = t < m && ((rest >> t) & 1ull);
                    const int d = member ? __popcll(bits_above(adj[t], t)) : 0;
                    const int c0 = member ? 12 * (d + 2) : 0, s0 = member ? 144 * (d * (d + 1) / 2) : 0, w0 = d;
                    int ic = c0, is = s0, iw = w0;
                    for (int off = 1; off < kWave; off <<= 1) {
                        const int vc = __shfl_up(ic, off, kWave), vs = __shfl_up(is, off, kWave),
                                  vw = __shfl_up(iw, off, kWave);
                        if (t >= off) {
                            ic += vc;
                            is += vs;
                            iw += vw;
                        }
                    }
                    const bool take = member && ic <= kLdltCols;
                    const unsigned long long cmask = __ballot(take);
                    const int nc = __popcll(cmask);
                    if (take) {
                        const int q = __popcll(cmask & bits_below(t));
                        cnode[q] = t;
                        cstart[q] = ic - c0;
                        sstart[q] = is - s0;
                        woff[t] = iw - w0;
                        if (q == nc - 1) {
                            cstart[nc] = ic;
                            sstart[nc] = is;
                        }
                    }
                    if (t == 0) {
                        if (nc == 0) cstart[0] = sstart[0] = 0;
                        nchunk = nc;
                        chunkmask = cmask;
                        restm = rest & ~cmask;
                    }
                }
                __syncthreads();
                const int nc = nchunk;
                if (nc == 0) break;  // uniform: the level is done
                const unsigned long long cmask = chunkmask;
                const int C = cstart[nc];
                double a[kLdltColsPerThread][12];
                int cln[kLdltColsPerThread], cg[kLdltColsPerThread], cc_[kLdltColsPerThread], cd[kLdltColsPerThread];
    #pragma unroll
                for (int s = 0; s < kLdltColsPerThread; ++s) {
                    const int col = kLdltColsPerThread * t + s;
                    cln[s] = -1;
                    cg[s] = cc_[s] = cd[s] = 0;
                    if (col < C) {
                        const int ln = last_le(cstart, nc, col);
                        const int k = cnode[ln], within = col - cstart[ln];
                        const int g = within / 12, c = within % 12, d = (cstart[ln + 1] - cstart[ln]) / 12 - 2;
                        cln[s] = ln;
                        cg[s] = g;
                        cc_[s] = c;
                        cd[s] = d;
                        if (g <= d) {
                            const auto src = U + 144 * (int64_t)(pos[k] + g) + c;
    #pragma unroll
                            for (int r = 0; r < 12; ++r) a[s][r] = src[12 * r];
                        } else {
    #pragma unroll
                            for (int r = 0; r < 12; ++r) a[s][r] = r == c ? 1.0 : 0.0;
                        }
                    }
                }
                // two pivot columns per barrier.  A thread's columns are adjacent and a body's
                // column count is a multiple of 12, so both belong to one body and the pivot pair
                // (cp, cp + 1) of a body is one thread's: it publishes column cp as it stands and
                // column cp + 1 as pivot cp leaves it (the update every column gets), with both
                // reciprocal pivots; every column then takes pivot cp, then pivot cp + 1 -- the
                // arithmetic of one pivot per step
    #pragma unroll
                for (int cp = 0; cp < 12; cp += 2) {
                    const int bf = (cp >> 1) & 1;
                    if (cln[0] >= 0 && cg[0] == 0 && cc_[0] == cp) {
                        const double inv0 = 1.0 / a[0][cp];
                        if (!(a[0][cp] > 0.0) || !isfinite(a[0][cp])) fail = 1;
                        const double sc1 = a[1][cp] * inv0;
                        double piv1 = 0.0;
    #pragma unroll
                        for (int r = 0; r < 12; ++r) {
                            const double v = r == cp ? sc1 : a[1][r] - a[0][r] * sc1;
                            colc[cln[0]][bf][0][r] = a[0][r];
                            colc[cln[0]][bf][1][r] = v;
                            if (r == cp + 1) piv1 = v;
                        }
                        if (!(piv1 > 0.0) || !isfinite(piv1)) fail = 1;
                        colinv[cln[0]][bf][0] = inv0;
                        colinv[cln[0]][bf][1] = 1.0 / piv1;
                    }
                    __syncthreads();
                    if (cln[0] >= 0) {
    #pragma unroll
                        for (int h = 0; h < 2; ++h) {  // pivot cp + h
                            const int pc = cp + h;
                            double f[12];
    #pragma unroll
                            for (int r = 0; r < 12; ++r) f[r] = colc[cln[0]][bf][h][r];
                            const double inv = colinv[cln[0]][bf][h];
    #pragma unroll
                            for (int s = 0; s < kLdltColsPerThread; ++s)
                                if (cln[s] >= 0) {
                                    const double sc = a[s][pc] * inv;
    #pragma unroll
                                    for (int r = 0; r < 12; ++r)
                                        if (r != pc) a[s][r] -= f[r] * sc;
                                    a[s][pc] = sc;
                                }
                        }
                    }
                }
                if (fail) {  // uniform: written before the last pivot column's barrier
                    if (t == 0) status[sidx] = kLdltFailed;
                    return false;
                }
                // W (g = 1..d) to the chunk's area, P^-1 (g = d + 1) over P
    #pragma unroll
                for (int s = 0; s < kLdltColsPerThread; ++s)
                    if (cln[s] >= 0 && cg[s] >= 1) {
                        const int k = cnode[cln[s]];
                        auto dst = cg[s] <= cd[s] ? W + 144 * (int64_t)(woff[k] + cg[s] - 1) + cc_[s]
                                                     : U + 144 * (int64_t)pos[k] + cc_[s];
    #pragma unroll
                        for (int r = 0; r < 12; ++r) dst[12 * r] = a[s][r];
                    }
                __syncthreads();
                // Schur complement: target (i, j), i <= j, updated by the thread of its lowest
                // contributor in the chunk, the contributors in ascending order
                for (int w = t; w < sstart[nc]; w += nt) {
                    const int ln = last_le(sstart, nc, w);
                    const int k = cnode[ln], within = w - sstart[ln];
                    const int pidx = within / 144, q = within % 144;
                    int v = (int)((sqrt(8.0 * pidx + 1.0) - 1.0) * 0.5);
                    while (v * (v + 1) / 2 > pidx) --v;
                    while ((v + 1) * (v + 2) / 2 <= pidx) ++v;
                    const int u = pidx - v * (v + 1) / 2;
                    const int i = mlist[k][u], j = mlist[k][v];
                    const unsigned long long contrib = cmask & colm[i] & colm[j];
                    if (__builtin_ctzll(contrib) != k) continue;
                    auto tgt = U + 144 * (int64_t)(pos[i] + __popcll(adj[i] & bits_below(j))) + q;
                    double val = *tgt;
                    for (unsigned long long bt = contrib; bt; bt &= bt - 1ull) {
                        const int k2 = __builtin_ctzll(bt);
                        const unsigned long long ab = bits_above(adj[k2], k2);
                        const auto Aki = U + 144 * (int64_t)(pos[k2] + 1 + __popcll(ab & bits_below(i))) + q / 12;
                        const auto Wj = W + 144 * (int64_t)(woff[k2] + __popcll(ab & bits_below(j))) + q % 12;
                        double acc = 0.0;
    #pragma unroll
                        for (int r = 0; r < 12; ++r) acc += Aki[12 * r] * Wj[12 * r];
                        val -= acc;
                    }
                    *tgt = val;
                }
                // forward solve: y_i -= W_i^T y_k over the chunk's bodies k coupled to i, ascending
                for (int w = t; w < n; w += nt) {
                    const int i = w / 12, a2 = w % 12;
                    unsigned long long bt = cmask & colm[i];
                    if (!bt) continue;
                    double val = y[w];
                    for (; bt; bt &= bt - 1ull) {
                        const int k2 = __builtin_ctzll(bt);
                        const unsigned long long ab = bits_above(adj[k2], k2);
                        const auto Wi = W + 144 * (int64_t)(woff[k2] + __popcll(ab & bits_below(i))) + a2;
                        double acc = 0.0;
    #pragma unroll
                        for (int r = 0; r < 12; ++r) acc += Wi[12 * r] * y[12 * k2 + r];
                        val -= acc;
                    }
                    y[w] = val;
                }
                // z_k = P^-1 y_k for the chunk's bodies
                for (int w = t; w < 12 * nc; w += nt) {
                    const int k = cnode[w / 12], a2 = w % 12;
                    const auto Di = U + 144 * (int64_t)pos[k] + 12 * a2;
                    double acc = 0.0;
    #pragma unroll
                    for (int r = 0; r < 12; ++r) acc += Di[r] * y[12 * k + r];
                    z[12 * k + a2] = acc;
                }
                __syncthreads();
                // rows of the factor: W
    #pragma unroll
                for (int s = 0; s < kLdltColsPerThread; ++s)
                    if (cln[s] >= 0 && cg[s] >= 1 && cg[s] <= cd[s]) {
                        auto dst = U + 144 * (int64_t)(pos[cnode[cln[s]]] + cg[s]) + cc_[s];
    #pragma unroll
                        for (int r = 0; r < 12; ++r) dst[12 * r] = a[s][r];
                    }
            }
        }
        // backward solve, levels from the top
        for (int L = levels - 1; L >= 0; --L) {
            __syncthreads();
            for (int w = t; w < n; w += nt) {
                const int k = w / 12, a2 = w % 12;
                if (lvl[k] != L) continue;
                double acc = 0.0;
                int u = 0;
                for (unsigned long long bt = bits_above(adj[k], k); bt; bt &= bt - 1ull, ++u) {
                    const int i = __builtin_ctzll(bt);
                    const auto Wu = U + 144 * (int64_t)(pos[k] + 1 + u) + 12 * a2;
                    double p = 0.0;
    #pragma unroll
                    for (int c = 0; c < 12; ++c) p += Wu[c] * y[12 * i + c];
                    acc += p;
                }
                y[w] = z[w] - acc;
            }
        }
        return true;
    };
    bool ok;
    if (in_lds) {
        typedef __attribute__((address_space(3))) double lds_double;
        lds_double* Ul = (lds_double*)ldlt_dyn;
        ok = factor_solve(Ul, Ul + 144 * (int64_t)pos[m]);
    } else {
        ok = factor_solve(U, W);
    }
    if (!ok) return;
    __syncthreads();
    for (int i = t; i < n; i += nt) xg[i] = y[i];
    double rr = 0.0;
    if (resid) {  // uniform: the residual's product with H only when asked for
        Rows R{row0, m, nullptr, nullptr, nullptr};
        for (int i = t; i < n; i += nt) {
            const int lr = i % 12;
            const double r = b[i] - row_product<kSpmvRows>(M, R, i / 12, lr / 3, lr % 3, y);
            rr += r * r;
        }
        rr = block_sum_fast(rr, sh);
    }
    if (t == 0) {
        status[sidx] = 1;
        if (resid) resid[sidx] = rr;
    }
}

// Algorithmic bytes of one solve launch, summed over its systems: per iteration
// each SpMV streams the system's blocks (144 doubles), column and slot ids, and
// row pointers; with vectors outside LDS, ~10 vector passes of 12 n doubles too.
__global__ void solve_bytes_kernel(const int32_t* __restrict__ row_ptr, int nb, const int32_t* __restrict__ sboff,
                                   const uint8_t* __restrict__ active, int nsys, const int32_t* __restrict__ iters,
                                   int spmv_per_it, int lds, double precond_per_body, double* __restrict__ out) {
    __shared__ double sh[17];
    double acc = 0.0;
    for (int s = threadIdx.x; s < nsys; s += blockDim.x) {
        if (active && !active[s]) continue;  // not solved in this launch
        const int r0 = sboff ? sboff[s] : 0, r1 = sboff ? sboff[s + 1] : nb;
        const double nnzb = (double)(row_ptr[r1] - row_ptr[r0]);
        const double nbs = (double)(r1 - r0);
        const double per_it = spmv_per_it * (1160.0 * nnzb + 4.0 * (nbs + 1)) + (lds ? 0.0 : 10.0 * 96.0 * nbs) +
                              precond_per_body * nbs;
        const int it = iters[s] > 0 ? iters[s] : 0;
        acc += per_it * it;
    }
    const double r = block_sum_fast(acc, sh);
    if (threadIdx.x == 0) *out = r;
}

// ---------------------------------------------------------------------------
// Jacobi CG of bsr_utils.py:126-235 (or its 12x12 block-Jacobi form) for one system
// too large for one workgroup's LDS (more than kLargeBodies bodies): the rows are split
// over workgroups of kMwThreads (kMwBjThreads, whole bodies, for the block form; one row
// per thread), and each iteration is three launches —
// [beta, p update], [SpMV + partial p.Ap], [alpha, x / r / z update + partial r.z].
// Every workgroup reduces the per-workgroup partials itself, in workgroup order,
// so all of them hold the same scalars and the sequence is deterministic.  The
// stopping test (every check_every iterations) sets a device flag that ends the
// remaining launches of the batch; the host reads it once per batch.
constexpr int kMwThreads = 256;
constexpr int kMwBjThreads = 192;  // 16 bodies of 12 rows
static_assert(kMwBjThreads % 12 == 0 && kMwBjThreads <= kMwThreads, "whole bodies per block-Jacobi workgroup");
constexpr int kLargeBodies = 170;  // from here one workgroup's SpMV is the bottleneck

struct MwState {
    double* x;
    double* r;
    double* p;
    double* Ap;
    double* z;
    double* dg;
    double* part;   // [2][nblk]: partials of p.Ap and r.z
    double* rz;     // [2]: r.z of the current iteration (slot i % 2) and the next
    double* err;    // final r.z
    int32_t* done;  // 1 once stopped
    int32_t* it;    // end_iter
};

__device__ __forceinline__ double mw_sum(const double* __restrict__ part, int nblk, double* sh) {
    double acc = 0.0;
    for (int k = threadIdx.x; k < nblk; k += blockDim.x) acc += part[k];
    return block_sum_fast(acc, sh);
}

__device__ __forceinline__ double mw_row(const Bsr& M, const double* __restrict__ in, int idx) {
    const int b = idx / 12, lr = idx % 12, ii = lr / 3, k = lr % 3;
    double acc = 0.0;
    for (int e = M.row_ptr[b]; e < M.row_ptr[b + 1]; ++e) {
        const double* xin = in + 12 * (int64_t)M.cols[e];
        const double* base = M.blocks + 9 * 16 * (int64_t)M.slots[e] + 3 * k;
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
            const double* B = base + 9 * (ii + 4 * jj);
            acc += B[0] * xin[3 * jj] + B[1] * xin[3 * jj + 1] + B[2] * xin[3 * jj + 2];
        }
    }
    return acc;
}

__device__ __forceinline__ double mw_jacobi(const double* __restrict__ dg, const double r, int i) {
    const int i3 = i - i % 3;
    const double d0 = dg[i3], d1 = dg[i3 + 1], d2 = dg[i3 + 2];
    return (d0 * d0 + d1 * d1 + d2 * d2 == 0.0) ? r : r / fabs(dg[i]);
}

// Block-Jacobi preconditioner of row i: z_i = (D_b^-1 r_b)_i with the 12 residuals of body b
// in rs (LDS, the workgroup's rows; a workgroup of kMwBjThreads rows holds whole bodies).
__device__ __forceinline__ double mw_block_precond(const double* __restrict__ dinv, const double* rs, int i) {
    const int lr = i % 12;
    const double* Di = dinv + 144 * (int64_t)(i / 12) + 12 * lr;
    const double* rb = rs + (threadIdx.x - lr);
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < 12; ++q) acc += Di[q] * rb[q];
    return acc;
}

// setup: diagonal, x = x0, r = b - H x, z, p = z, partial r.z
template <bool BJ>
__global__ __launch_bounds__(kMwThreads) void mw_init_kernel(Bsr M, const double* __restrict__ b,
                                                             const double* __restrict__ xg, MwState W,
                                                             const double* __restrict__ dinv) {
    __shared__ double sh[17];
    __shared__ double rs[BJ ? kMwThreads : 1];
    const int n = 12 * M.nb;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double part = 0.0, r = 0.0, z = 0.0;
    if (i < n) {
        // bsr_get_diag: the three diagonal entries of this row's 3x3 diagonal sub-block (a
        // 3-row group may straddle two workgroups, so each thread reads all three itself)
        const int bb = i / 12, ii = (i % 12) / 3, a = i % 3;
        const double* D = M.blocks + 9 * (16 * (int64_t)bb + ii + 4 * ii);
        const double d[3] = {D[0], D[4], D[8]};
        W.dg[i] = d[a];
        W.x[i] = xg[i];
        r = b[i] - mw_row(M, xg, i);
        W.r[i] = r;
        if (!BJ) z = (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] == 0.0) ? r : r / fabs(d[a]);
    }
    if (BJ) {
        rs[threadIdx.x] = r;
        __syncthreads();
        if (i < n) z = mw_block_precond(dinv, rs, i);
    }
    if (i < n) {
        W.z[i] = z;
        W.p[i] = z;
        part = r * z;
    }
    const double t = block_sum_fast(part, sh);
    if (threadIdx.x == 0) W.part[gridDim.x + blockIdx.x] = t;
}

__global__ __launch_bounds__(kMwThreads) void mw_start_kernel(int nblk, double tol_sq, int maxiter, MwState W) {
    __shared__ double sh[17];
    const double rz = mw_sum(W.part + nblk, nblk, sh);
    if (threadIdx.x == 0) {
        W.rz[0] = rz;
        *W.err = rz;
        *W.it = rz > tol_sq ? maxiter : 0;
        *W.done = rz > tol_sq ? 0 : 1;
    }
}

// iteration i, part 1: SpMV of p (complete after the previous launch) and partial p.Ap
__global__ __launch_bounds__(kMwThreads) void mw_spmv_kernel(Bsr M, MwState W) {
    __shared__ double sh[17];
    if (*W.done) return;
    const int n = 12 * M.nb;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double part = 0.0;
    if (i < n) {
        const double a = mw_row(M, W.p, i);
        W.Ap[i] = a;
        part = W.p[i] * a;
    }
    const double t = block_sum_fast(part, sh);
    if (threadIdx.x == 0) W.part[blockIdx.x] = t;
}

// part 2: alpha, x / r / z update, partial r.z
template <bool BJ>
__global__ __launch_bounds__(kMwThreads) void mw_update_kernel(int n, int iter, MwState W,
                                                               const double* __restrict__ dinv) {
    __shared__ double sh[17];
    __shared__ double rs[BJ ? kMwThreads : 1];
    if (*W.done) return;  // uniform over the grid
    const int nblk = gridDim.x;
    const double pAp = mw_sum(W.part, nblk, sh);
    const double rz_old = W.rz[iter & 1];
    // unguarded in the reference: p = 0 once the residual vanishes exactly between checks
    const double alpha = pAp != 0.0 ? rz_old / pAp : 0.0;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double part = 0.0, r = 0.0, z = 0.0;
    if (i < n) {
        W.x[i] += alpha * W.p[i];
        r = W.r[i] - alpha * W.Ap[i];
        W.r[i] = r;
        if (!BJ) z = mw_jacobi(W.dg, r, i);
    }
    if (BJ) {
        rs[threadIdx.x] = r;
        __syncthreads();
        if (i < n) z = mw_block_precond(dinv, rs, i);
    }
    if (i < n) {
        W.z[i] = z;
        part = r * z;
    }
    const double t = block_sum_fast(part, sh);
    if (threadIdx.x == 0) W.part[nblk + blockIdx.x] = t;
}

// part 3: stopping test, beta, p = z + beta p
__global__ __launch_bounds__(kMwThreads) void mw_direction_kernel(int n, int iter, int check_every, double tol_sq,
                                                                  MwState W) {
    __shared__ double sh[17];
    if (*W.done) return;
    const int nblk = gridDim.x;
    const double rz_new = mw_sum(W.part + nblk, nblk, sh);
    const double rz_old = W.rz[iter & 1];
    if (((iter + 1) % check_every) == 0 && rz_new <= tol_sq) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            *W.done = 1;
            *W.it = iter;
            *W.err = rz_new;
        }
        return;
    }
    const double beta = rz_new > tol_sq ? rz_new / rz_old : 0.0;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) W.p[i] = W.z[i] + beta * W.p[i];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        W.rz[(iter + 1) & 1] = rz_new;
        *W.err = rz_new;
    }
}

__global__ void mw_finish_kernel(int n, MwState W, double* __restrict__ xg, int32_t* __restrict__ iters,
                                 double* __restrict__ resid) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) xg[i] = W.x[i];
    if (i == 0) {
        if (iters) *iters = *W.it;
        if (resid) *resid = *W.err;
    }
}

}  // namespace abd

using namespace abd;

extern "C" {

namespace {
// Workgroup size of a solve.  It must not depend on the other systems of a launch,
// so that a scene solved on its own or inside a batch reduces in the same order;
// 16 waves per system keep enough loads in flight for the block SpMV.
int solve_threads(int) { return kSolveThreads; }

int launch_solve(hipStream_t s, int32_t method, const Bsr& M, int nsys, const int32_t* sboff, const uint8_t* active,
                 int max_rows, const double* b, double* x, double tol, int32_t maxiter, int32_t check_every,
                 int32_t* iters, double* resid, int threads, const int32_t* pending = nullptr) {
    const int n_all = 12 * M.nb;
    double* ws;
    int rc = arena_get(kSlotSolver, sizeof(double) * 6 * (size_t)n_all, (void**)&ws);
    if (rc) return rc;
    const bool bj = method == ABD_SOLVER_CG_BLOCK;
    double* dinv = nullptr;
    if (bj && (rc = arena_get(kSlotPrecond, sizeof(double) * 144 * (size_t)M.nb, (void**)&dinv))) return rc;
    if (bj) {
        // every body's diagonal-block inverse, one thread per body (inactive scenes included: cheap)
        if (M.nb > 0) {
            // algorithmic bytes: each diagonal block read, its inverse written
            const int ph_inv = pending ? -1 : prof_start(kProfPrecond, s, 2304.0 * M.nb);
            hipLaunchKernelGGL(invert_blocks_kernel, dim3(M.nb), dim3(kInvThreads), 0, s, M, dinv, pending, sboff,
                               nsys);
            ABD_CHECK_LAUNCH();
            prof_stop(ph_inv, s);
        }
    }
    // the fallback of a direct solve is timed with it (its launch returns at once unless a
    // system was declined)
    int ph = pending ? -1 : prof_start(kProfSolve, s, 1.0);
    const size_t vec_bytes = sizeof(double) * 6 * (size_t)max_rows;
    // dynamic LDS: the whole budget of the device (beside the kernels' static LDS) when it
    // holds the structure region, the vectors in it when they fit beside the structure
    int dev = 0, max_lds = 0;
    ABD_CHECK(hipGetDevice(&dev));
    ABD_CHECK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    const size_t usable = (size_t)max_lds > 1024 ? (size_t)max_lds - 1024 : 0;
    const bool staged = usable >= kLdsStructBytes + 8 * 1024;
    const int budget = staged ? (int)usable : -1;
    const bool lds = staged && kLdsStructBytes + vec_bytes <= usable;
    const size_t dyn = staged ? usable : 0;
    static bool attr_set = false;
    static std::mutex attr_mu;  // the first calls of several threads
    std::unique_lock<std::mutex> attr_lock(attr_mu);
    if (!attr_set) {
        const int cap = (int)usable;
        const void* fns[] = {(const void*)bicgstab_kernel<true>,     (const void*)bicgstab_kernel<false>,
                             (const void*)cg_kernel<true, false>,    (const void*)cg_kernel<true, true>,
                             (const void*)cg_kernel<false, false>,   (const void*)cg_kernel<false, true>};
        for (const void* f : fns) ABD_CHECK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, cap));
        attr_set = true;
    }
    attr_lock.unlock();
    if (lds) {
        if (method == ABD_SOLVER_BICGSTAB)
            hipLaunchKernelGGL(bicgstab_kernel<true>, dim3(nsys), dim3(threads), dyn, s, M, sboff, active, b, x, tol,
                               maxiter, check_every, max_rows, budget, ws, iters, resid);
        else if (bj)
            hipLaunchKernelGGL((cg_kernel<true, true>), dim3(nsys), dim3(threads), dyn, s, M, sboff, active, b, x,
                               tol, maxiter, check_every, max_rows, budget, ws, dinv, iters, resid, pending);
        else
            hipLaunchKernelGGL((cg_kernel<true, false>), dim3(nsys), dim3(threads), dyn, s, M, sboff, active, b, x,
                               tol, maxiter, check_every, max_rows, budget, ws, dinv, iters, resid, pending);
    } else {
        if (method == ABD_SOLVER_BICGSTAB)
            hipLaunchKernelGGL(bicgstab_kernel<false>, dim3(nsys), dim3(threads), dyn, s, M, sboff, active, b, x,
                               tol, maxiter, check_every, max_rows, budget, ws, iters, resid);
        else if (bj)
            hipLaunchKernelGGL((cg_kernel<false, true>), dim3(nsys), dim3(threads), dyn, s, M, sboff, active, b, x,
                               tol, maxiter, check_every, max_rows, budget, ws, dinv, iters, resid, pending);
        else
            hipLaunchKernelGGL((cg_kernel<false, false>), dim3(nsys), dim3(threads), dyn, s, M, sboff, active, b, x,
                               tol, maxiter, check_every, max_rows, budget, ws, dinv, iters, resid, pending);
    }
    ABD_CHECK_LAUNCH();
    prof_stop(ph, s);
    if (ph >= 0 && iters) {
        // the byte count needs the iteration counts: computed behind the solve, on its stream
        double* dbytes;
        if ((rc = arena_get(kSlotSolveBytes, sizeof(double), (void**)&dbytes))) return rc;
        hipLaunchKernelGGL(solve_bytes_kernel, dim3(1), dim3(256), 0, s, M.row_ptr, M.nb, sboff, active, nsys, iters,
                           method == ABD_SOLVER_BICGSTAB ? 2 : 1, lds ? 1 : 0, bj ? 1152.0 : 0.0, dbytes);
        ABD_CHECK_LAUNCH();
        prof_stop(ph, s, nullptr, dbytes);
    } else {
        prof_stop(ph, s);
    }
    return ABD_OK;
}
// The direct solve (ldlt_kernel) of every system of at most kLdltMaxBodies bodies, then
// the block-Jacobi CG of the systems it declined (status kLdltFailed in iters).
int launch_ldlt(hipStream_t s, const Bsr& M, int nsys, const int32_t* sboff, const uint8_t* active, int max_bodies,
                const double* b, double* x, double tol, int32_t maxiter, int32_t check_every, int32_t* iters,
                double* resid) {
    int rc;
    int32_t* status = iters;
    if (!status && (rc = arena_get(kSlotLdltStatus, sizeof(int32_t) * (size_t)nsys, (void**)&status))) return rc;
    const int cap = max_bodies < kLdltMaxBodies ? max_bodies : kLdltMaxBodies;
    const int64_t region = ldlt_region(cap);
    double* ws;
    if ((rc = arena_get(kSlotLdlt, sizeof(double) * (size_t)region * (size_t)nsys, (void**)&ws))) return rc;
    // dynamic LDS: the device's budget beside the kernel's static LDS, in whole blocks
    static int dyn_bytes = -1;
    static std::mutex dyn_mu;  // the first calls of several threads
    std::unique_lock<std::mutex> dyn_lock(dyn_mu);
    if (dyn_bytes < 0) {
        int dev = 0, max_lds = 0;
        ABD_CHECK(hipGetDevice(&dev));
        ABD_CHECK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        hipFuncAttributes fa;
        ABD_CHECK(hipFuncGetAttributes(&fa, (const void*)ldlt_kernel));
        const int avail = max_lds - (int)fa.sharedSizeBytes - 256;
        dyn_bytes = avail > 0 ? (avail / 1152) * 1152 : 0;
        if (dyn_bytes > 0)
            ABD_CHECK(hipFuncSetAttribute((const void*)ldlt_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                          dyn_bytes));
    }
    dyn_lock.unlock();
    const int ph = prof_start_ext(kProfSolve, 1.0);
    prof_launch(ph, ldlt_kernel, dim3(nsys), dim3(kSolveThreads), (uint32_t)dyn_bytes, s, M, sboff, active, b, x,
                region, ws, status, resid, dyn_bytes / 1152);
    ABD_CHECK_LAUNCH();
    if (ph >= 0) {
        // algorithmic bytes of the systems solved directly: each block, b and x once
        double* dbytes;
        if ((rc = arena_get(kSlotSolveBytes, sizeof(double), (void**)&dbytes))) return rc;
        hipLaunchKernelGGL(solve_bytes_kernel, dim3(1), dim3(256), 0, s, M.row_ptr, M.nb, sboff, active, nsys,
                           (const int32_t*)status, 1, 1, 192.0, dbytes);
        ABD_CHECK_LAUNCH();
        prof_stop(ph, s, nullptr, dbytes);
    }
    if ((rc = launch_solve(s, ABD_SOLVER_CG_BLOCK, M, nsys, sboff, active, 12 * max_bodies, b, x, tol, maxiter,
                           check_every, status, resid, solve_threads(max_bodies), status)))
        return rc;
    return ABD_OK;
}
}  // namespace

// Jacobi (bj: 12x12 block-Jacobi) CG over several workgroups (systems of more than
// kLargeBodies bodies).  The block-Jacobi form runs kMwBjThreads rows per workgroup so that
// every body's 12 rows share one.
int solve_cg_multi(hipStream_t s, const Bsr& M, const double* b, double* x, double tol, int32_t maxiter,
                   int32_t check_every, int32_t* iters, double* resid, bool bj) {
    const int n = 12 * M.nb;
    const int threads = bj ? kMwBjThreads : kMwThreads;
    const int nblk = (n + threads - 1) / threads;
    if (maxiter <= 0) maxiter = n;
    double* ws;
    const size_t nd = 6 * (size_t)n + 2 * (size_t)nblk + 3;
    int rc = arena_get(kSlotSolver, sizeof(double) * nd + 2 * sizeof(int32_t) + 16, (void**)&ws);
    if (rc) return rc;
    MwState W;
    W.x = ws;
    W.r = ws + n;
    W.p = ws + 2 * (size_t)n;
    W.Ap = ws + 3 * (size_t)n;
    W.z = ws + 4 * (size_t)n;
    W.dg = ws + 5 * (size_t)n;
    W.part = ws + 6 * (size_t)n;
    W.rz = W.part + 2 * nblk;
    W.err = W.rz + 2;
    W.done = reinterpret_cast<int32_t*>(W.err + 1);
    W.it = W.done + 1;
    int32_t* hdone;
    if ((rc = arena_host(sizeof(int32_t), (void**)&hdone))) return rc;
    double* dinv = nullptr;
    if (bj) {
        if ((rc = arena_get(kSlotPrecond, sizeof(double) * 144 * (size_t)M.nb, (void**)&dinv))) return rc;
        const int ph_inv = prof_start(kProfPrecond, s, 2304.0 * M.nb);
        hipLaunchKernelGGL(invert_blocks_kernel, dim3(M.nb), dim3(kInvThreads), 0, s, M, dinv,
                           (const int32_t*)nullptr, (const int32_t*)nullptr, 1);
        ABD_CHECK_LAUNCH();
        prof_stop(ph_inv, s);
    }
    const double tol_sq = tol * tol * (double)n;
    int ph = prof_start(kProfSolve, s, 0.0);
    if (bj)
        hipLaunchKernelGGL(mw_init_kernel<true>, dim3(nblk), dim3(threads), 0, s, M, b, (const double*)x, W,
                           (const double*)dinv);
    else
        hipLaunchKernelGGL(mw_init_kernel<false>, dim3(nblk), dim3(threads), 0, s, M, b, (const double*)x, W,
                           (const double*)dinv);
    ABD_CHECK_LAUNCH();
    hipLaunchKernelGGL(mw_start_kernel, dim3(1), dim3(kMwThreads), 0, s, nblk, tol_sq, maxiter, W);
    ABD_CHECK_LAUNCH();
    for (int i0 = 0; i0 < maxiter; i0 += check_every) {
        const int i1 = i0 + check_every < maxiter ? i0 + check_every : maxiter;
        for (int i = i0; i < i1; ++i) {
            hipLaunchKernelGGL(mw_spmv_kernel, dim3(nblk), dim3(threads), 0, s, M, W);
            if (bj)
                hipLaunchKernelGGL(mw_update_kernel<true>, dim3(nblk), dim3(threads), 0, s, n, i, W,
                                   (const double*)dinv);
            else
                hipLaunchKernelGGL(mw_update_kernel<false>, dim3(nblk), dim3(threads), 0, s, n, i, W,
                                   (const double*)dinv);
            hipLaunchKernelGGL(mw_direction_kernel, dim3(nblk), dim3(threads), 0, s, n, i, (int)check_every,
                               tol_sq, W);
        }
        ABD_CHECK_LAUNCH();
        ABD_CHECK(hipMemcpyAsync(hdone, W.done, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        ABD_CHECK(hipStreamSynchronize(s));
        if (*hdone) break;
    }
    hipLaunchKernelGGL(mw_finish_kernel, dim3(nblk), dim3(kMwThreads), 0, s, n, W, x, iters, resid);
    ABD_CHECK_LAUNCH();
    prof_stop(ph, s);
    return ABD_OK;
}

int abd_bsr_solve(abd_stream_t stream, int32_t method, int32_t nb, const int32_t* row_ptr, const int32_t* cols,
                  const int32_t* slots, const double* blocks, const double* b, double* x, double tol, int32_t maxiter,
                  int32_t check_every, int32_t* iters, double* resid) {
    if (nb <= 0 || !row_ptr || !cols || !slots || !blocks || !b || !x || check_every <= 0) return ABD_ERR_ARG;
    if (method != ABD_SOLVER_BICGSTAB && method != ABD_SOLVER_CG_JACOBI && method != ABD_SOLVER_CG_BLOCK &&
        method != ABD_SOLVER_LDLT)
        return ABD_ERR_ARG;
    Bsr M{nb, row_ptr, cols, slots, blocks};
    if (method == ABD_SOLVER_LDLT) {
        if (nb <= kLdltMaxBodies)
            return launch_ldlt(to_stream(stream), M, 1, nullptr, nullptr, nb, b, x, tol, maxiter, check_every, iters,
                               resid);
        method = ABD_SOLVER_CG_BLOCK;  // above the direct solve's size: its fallback
    }
    if ((method == ABD_SOLVER_CG_JACOBI || method == ABD_SOLVER_CG_BLOCK) && nb > kLargeBodies)
        return solve_cg_multi(to_stream(stream), M, b, x, tol, maxiter, check_every, iters, resid,
                              method == ABD_SOLVER_CG_BLOCK);
    return launch_solve(to_stream(stream), method, M, 1, nullptr, nullptr, 12 * nb, b, x, tol, maxiter, check_every,
                        iters, resid, solve_threads(nb));
}

int abd_bsr_solve_scenes(abd_stream_t stream, int32_t method, const AbdScene* sc, int32_t max_scene_bodies,
                         const int32_t* row_ptr, const int32_t* cols, const int32_t* slots, const double* blocks,
                         const double* b, double* x, double tol, int32_t maxiter, int32_t check_every,
                         const uint8_t* active, int32_t* iters, double* resid) {
    if (!sc || sc->n_scenes <= 0 || !sc->scene_boff || max_scene_bodies <= 0 || !row_ptr || !cols || !slots ||
        !blocks || !b || !x || check_every <= 0)
        return ABD_ERR_ARG;
    if (method != ABD_SOLVER_BICGSTAB && method != ABD_SOLVER_CG_JACOBI && method != ABD_SOLVER_CG_BLOCK &&
        method != ABD_SOLVER_LDLT)
        return ABD_ERR_ARG;
    Bsr M{sc->n_bodies, row_ptr, cols, slots, blocks};
    if (method == ABD_SOLVER_LDLT)
        return launch_ldlt(to_stream(stream), M, sc->n_scenes, sc->scene_boff, active, max_scene_bodies, b, x, tol,
                           maxiter, check_every, iters, resid);
    return launch_solve(to_stream(stream), method, M, sc->n_scenes, sc->scene_boff, active, 12 * max_scene_bodies, b,
                        x, tol, maxiter, check_every, iters, resid, solve_threads(max_scene_bodies));
}

}  // extern "C"
