// abd_geom.hip — per-vertex / per-primitive / per-body geometry: vertex
// projection, state updates and the FP32 bounding boxes of culling2.py.
// All of it is streaming, HBM-bound work: one thread per element, coalesced
// along the flattened arrays.
#include <hip/hip_runtime.h>

#include "abd_block.h"
#include "abd_device.h"
#include "abd_internal.h"
#include "abd_profile.h"

namespace abd {

// abd.py:305-324: x = A^T x0 + p, i.e. x_k = ((p_k + x0_0 A[0][k]) + x0_1 A[1][k]) + x0_2 A[2][k]
__global__ __launch_bounds__(kBlock) void project_kernel(const int32_t* __restrict__ vert_body,
                                                          const double* __restrict__ x0,
                                                          const double* __restrict__ A,
                                                          const double* __restrict__ p, double* __restrict__ x,
                                                          int64_t nv) {
    for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < nv; v += (int64_t)gridDim.x * kBlock) {
        int b = vert_body[v];
        const double* Ab = A + 9 * (int64_t)b;
        double r0 = x0[3 * v], r1 = x0[3 * v + 1], r2 = x0[3 * v + 2];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double r = p[3 * (int64_t)b + k] + r0 * Ab[k];
            r = r + r1 * Ab[3 + k];
            r = r + r2 * Ab[6 + k];
            x[3 * v + k] = r;
        }
    }
}

// The end of a step, x = x(q - alpha dq): per vertex the body's state with the arithmetic of
// update_q_kernel (one alpha_host) or of update_q_scenes_kernel (alpha_dev per scene, or
// alpha_host for the scenes in `active` and 0 for the others; a scene with alpha 0 keeps its
// state bit for bit), placed as project_kernel does.  With lo / hi, the range test of
// within_ranges on this end and on x_other: *flag = gen when some coordinate leaves them.
__global__ __launch_bounds__(kBlock) void project_step_kernel(const int32_t* __restrict__ vert_body,
                                                               const int32_t* __restrict__ body_scene,
                                                               const double* __restrict__ x0,
                                                               const double* __restrict__ A,
                                                               const double* __restrict__ p,
                                                               const double* __restrict__ dq,
                                                               const double* __restrict__ alpha_dev,
                                                               double alpha_host,
                                                               const uint8_t* __restrict__ active,
                                                               double* __restrict__ x,
                                                               const double* __restrict__ x_other,
                                                               const double* __restrict__ lo,
                                                               const double* __restrict__ hi,
                                                               int64_t* __restrict__ flag, int64_t gen, int64_t nv) {
    bool out = false;
    for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < nv; v += (int64_t)gridDim.x * kBlock) {
        const int b = vert_body[v];
        double pb[3], Ab[9];
        if (alpha_dev || active) {
            const int sc = body_scene ? body_scene[b] : 0;
            const double al = alpha_dev ? alpha_dev[sc] : (active[sc] ? alpha_host : 0.0);
#pragma unroll
            for (int r = 0; r < 3; ++r) pb[r] = al == 0.0 ? p[3 * b + r] : p[3 * b + r] - dq[3 * (4 * b) + r] * al;
#pragma unroll
            for (int r = 0; r < 9; ++r)
                Ab[r] = al == 0.0 ? A[9 * b + r] : A[9 * b + r] - dq[3 * (4 * b + 1 + r / 3) + r % 3] * al;
        } else {
#pragma unroll
            for (int r = 0; r < 3; ++r) pb[r] = p[3 * b + r] - dq[3 * (4 * b) + r] * alpha_host;
#pragma unroll
            for (int r = 0; r < 9; ++r) Ab[r] = A[9 * b + r] - dq[3 * (4 * b + 1 + r / 3) + r % 3] * alpha_host;
        }
        const double r0 = x0[3 * v], r1 = x0[3 * v + 1], r2 = x0[3 * v + 2];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double r = pb[k] + r0 * Ab[k];
            r = r + r1 * Ab[3 + k];
            r = r + r2 * Ab[6 + k];
            x[3 * v + k] = r;
            if (lo) {
                const double l = lo[3 * v + k], h = hi[3 * v + k], o = x_other[3 * v + k];
                out |= !(r >= l && r <= h && o >= l && o <= h);  // NaN counts as outside
            }
        }
    }
    if (lo && __any(out) && lane_id() == 0) *flag = gen;  // every writer stores gen
}

// The body state of a step's end for project_step2_kernel: project_step_kernel's arithmetic.
__device__ __forceinline__ void step_end_body(int b, const int32_t* __restrict__ body_scene,
                                              const double* __restrict__ A, const double* __restrict__ p,
                                              const double* __restrict__ dq, const double* __restrict__ alpha_dev,
                                              double alpha_host, const uint8_t* __restrict__ active, double pb[3],
                                              double Ab[9]) {
    if (alpha_dev || active) {
        const int sc = body_scene ? body_scene[b] : 0;
        const double al = alpha_dev ? alpha_dev[sc] : (active[sc] ? alpha_host : 0.0);
#pragma unroll
        for (int r = 0; r < 3; ++r) pb[r] = al == 0.0 ? p[3 * b + r] : p[3 * b + r] - dq[3 * (4 * b) + r] * al;
#pragma unroll
        for (int r = 0; r < 9; ++r)
            Ab[r] = al == 0.0 ? A[9 * b + r] : A[9 * b + r] - dq[3 * (4 * b + 1 + r / 3) + r % 3] * al;
    } else {
#pragma unroll
        for (int r = 0; r < 3; ++r) pb[r] = p[3 * b + r] - dq[3 * (4 * b) + r] * alpha_host;
#pragma unroll
        for (int r = 0; r < 9; ++r) Ab[r] = A[9 * b + r] - dq[3 * (4 * b + 1 + r / 3) + r % 3] * alpha_host;
    }
}

// project_step_kernel over vertex pairs (2i, 2i + 1): the pair's six coordinates of x0, the ranges,
// x_other and x as three 16-byte accesses each (the arrays 16-byte aligned), the body state computed
// again for the second vertex only when it is of another body; an odd last vertex is the last
// thread's.  Same
// arithmetic, same results.
typedef double dvec2 __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(kBlock) void project_step2_kernel(const int32_t* __restrict__ vert_body,
                                                                const int32_t* __restrict__ body_scene,
                                                                const double* __restrict__ x0,
                                                                const double* __restrict__ A,
                                                                const double* __restrict__ p,
                                                                const double* __restrict__ dq,
                                                                const double* __restrict__ alpha_dev,
                                                                double alpha_host,
                                                                const uint8_t* __restrict__ active,
                                                                double* __restrict__ x,
                                                                const double* __restrict__ x_other,
                                                                const double* __restrict__ lo,
                                                                const double* __restrict__ hi,
                                                                int64_t* __restrict__ flag, int64_t gen, int64_t nv) {
    bool out = false;
    const int64_t npair = (nv + 1) / 2;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < npair; i += (int64_t)gridDim.x * kBlock) {
        const int64_t v0 = 2 * i;
        const bool two = v0 + 1 < nv;
        const int b0 = vert_body[v0], b1 = two ? vert_body[v0 + 1] : b0;
        double c[6], l[6], h[6], o[6], rr[6];
        if (two) {
            const dvec2* X0 = reinterpret_cast<const dvec2*>(x0 + 3 * v0);
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const dvec2 a = X0[q];
                c[2 * q] = a.x;
                c[2 * q + 1] = a.y;
            }
            if (lo) {
                const dvec2 *L = reinterpret_cast<const dvec2*>(lo + 3 * v0), *H = reinterpret_cast<const dvec2*>(hi + 3 * v0),
                            *O = reinterpret_cast<const dvec2*>(x_other + 3 * v0);
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const dvec2 a = L[q], b = H[q], e = O[q];
                    l[2 * q] = a.x;
                    l[2 * q + 1] = a.y;
                    h[2 * q] = b.x;
                    h[2 * q + 1] = b.y;
                    o[2 * q] = e.x;
                    o[2 * q + 1] = e.y;
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                c[k] = x0[3 * v0 + k];
                c[3 + k] = 0.0;
                if (lo) {
                    l[k] = lo[3 * v0 + k];
                    h[k] = hi[3 * v0 + k];
                    o[k] = x_other[3 * v0 + k];
                }
            }
        }
        // one body state in registers: vertex 1's is computed over vertex 0's when they differ
        double pb[3], Ab[9];
        step_end_body(b0, body_scene, A, p, dq, alpha_dev, alpha_host, active, pb, Ab);
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            if (u == 1 && b1 != b0) step_end_body(b1, body_scene, A, p, dq, alpha_dev, alpha_host, active, pb, Ab);
            const double r0 = c[3 * u], r1 = c[3 * u + 1], r2 = c[3 * u + 2];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                double r = pb[k] + r0 * Ab[k];
                r = r + r1 * Ab[3 + k];
                r = r + r2 * Ab[6 + k];
                rr[3 * u + k] = r;
                if (lo && (u == 0 || two)) {
                    const int j = 3 * u + k;
                    out |= !(r >= l[j] && r <= h[j] && o[j] >= l[j] && o[j] <= h[j]);  // NaN counts as outside
                }
            }
        }
        if (two) {
            dvec2* Xo = reinterpret_cast<dvec2*>(x + 3 * v0);
#pragma unroll
            for (int q = 0; q < 3; ++q) Xo[q] = dvec2{rr[2 * q], rr[2 * q + 1]};
        } else {
#pragma unroll
            for (int k = 0; k < 3; ++k) x[3 * v0 + k] = rr[k];
        }
    }
    if (lo && __any(out) && lane_id() == 0) *flag = gen;  // every writer stores gen
}

// abd.py:337-353
__global__ void update_q_kernel(int nb, const double* __restrict__ A, const double* __restrict__ p,
                                const double* __restrict__ dq, double alpha, double* A_out, double* p_out,
                                double* A_out2, double* p_out2) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 12 * nb) return;
    int b = i / 12, r = i % 12;
    if (r < 3) {
        const double v = p[3 * b + r] - dq[3 * (4 * b) + r] * alpha;
        p_out[3 * b + r] = v;
        if (p_out2) p_out2[3 * b + r] = v;
    } else {
        int row = (r - 3) / 3, k = (r - 3) % 3;
        const double v = A[9 * b + 3 * row + k] - dq[3 * (4 * b + 1 + row) + k] * alpha;
        A_out[9 * b + 3 * row + k] = v;
        if (A_out2) A_out2[9 * b + 3 * row + k] = v;
    }
}

// abd.py:355-362
__global__ void update_q0qdot_kernel(int nb, double dt, const double* __restrict__ A, const double* __restrict__ p,
                                     double* A0, double* p0, double* Adot, double* pdot) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 12 * nb) return;
    int b = i / 12, r = i % 12;
    if (r < 3) {
        int k = 3 * b + r;
        pdot[k] = (p[k] - p0[k]) / dt;
        p0[k] = p[k];
    } else {
        int k = 9 * b + (r - 3);
        Adot[k] = (A[k] - A0[k]) / dt;
        A0[k] = A[k];
    }
}

// culling2.py:72-79 bvh_points: cast, then FP32 +- f32(dil)
__global__ __launch_bounds__(kBlock) void point_boxes_kernel(const double* __restrict__ x, int64_t n, float dil,
                                                              float* __restrict__ lo, float* __restrict__ hi) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < 3 * n; i += (int64_t)gridDim.x * kBlock) {
        float c = (float)x[i];
        lo[i] = c - dil;
        hi[i] = c + dil;
    }
}
// culling2.py:33-69 bvh_triangles / bvh_edges: FP64 min/max +- dil, then cast.
template <int NV>
__global__ __launch_bounds__(kBlock) void prim_boxes_kernel(const int32_t* __restrict__ prim_body,
                                                             const int32_t* __restrict__ voff,
                                                             const int32_t* __restrict__ topo,
                                                             const double* __restrict__ x, int64_t n, double dil,
                                                             float* __restrict__ lo, float* __restrict__ hi) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        int vb = voff[prim_body[i]];
        int64_t vi[NV];
#pragma unroll
        for (int j = 0; j < NV; ++j) vi[j] = vb + topo[NV * i + j];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double mn = x[3 * vi[0] + k], mx = mn;
#pragma unroll
            for (int j = 1; j < NV; ++j) {
                double c = x[3 * vi[j] + k];
                mn = c < mn ? c : mn;
                mx = c > mx ? c : mx;
            }
            lo[3 * i + k] = (float)(mn - dil);
            hi[3 * i + k] = (float)(mx + dil);
        }
    }
}
// culling2.py:273-280 bvh_point_trajectory
__global__ __launch_bounds__(kBlock) void point_traj_kernel(const double* __restrict__ x, int64_t n,
                                                             float* __restrict__ lo, float* __restrict__ hi) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < 3 * n; i += (int64_t)gridDim.x * kBlock) {
        float c = (float)x[i];
        hi[i] = fmaxf(hi[i], c);
        lo[i] = fminf(lo[i], c);
    }
}
// culling2.py:230-270 bvh_triangle_trajectory / bvh_edge_trajectory
template <int NV>
__global__ __launch_bounds__(kBlock) void prim_traj_kernel(const int32_t* __restrict__ prim_body,
                                                            const int32_t* __restrict__ voff,
                                                            const int32_t* __restrict__ topo,
                                                            const double* __restrict__ x, int64_t n,
                                                            float* __restrict__ lo, float* __restrict__ hi) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        int vb = voff[prim_body[i]];
        int64_t vi[NV];
#pragma unroll
        for (int j = 0; j < NV; ++j) vi[j] = vb + topo[NV * i + j];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double mn = x[3 * vi[0] + k], mx = mn;
#pragma unroll
            for (int j = 1; j < NV; ++j) {
                double c = x[3 * vi[j] + k];
                mn = c < mn ? c : mn;
                mx = c > mx ? c : mx;
            }
            hi[3 * i + k] = fmaxf(hi[3 * i + k], (float)mx);
            lo[3 * i + k] = fminf(lo[3 * i + k], (float)mn);
        }
    }
}

// FP64 extrema of each coordinate over a body's vertices [v0, v1), one wave per body:
// lanes stride the vertices, then a shuffle tree (a minimum or maximum does not depend
// on the order).  Every lane receives the result.
__device__ __forceinline__ void wave_extrema(const double* __restrict__ x, int v0, int v1, double (&mn)[3],
                                             double (&mx)[3]) {
    for (int k = 0; k < 3; ++k) {
        mn[k] = 1.0 / 0.0;
        mx[k] = -1.0 / 0.0;
    }
    for (int v = v0 + lane_id(); v < v1; v += kWave)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double c = x[3 * (int64_t)v + k];
            mn[k] = c < mn[k] ? c : mn[k];
            mx[k] = c > mx[k] ? c : mx[k];
        }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double a = __shfl_xor(mn[k], off, kWave), b = __shfl_xor(mx[k], off, kWave);
            mn[k] = a < mn[k] ? a : mn[k];
            mx[k] = b > mx[k] ? b : mx[k];
        }
}

// culling2.py:171-210: the FP64 extrema of each body's vertices, one wave per body
// (kBlock / 64 bodies per workgroup).  traj = 0: lo/hi = f32(min/max -+ dil);
// traj = 1: lo/hi = min/max(lo/hi, f32(min/max)).
__global__ __launch_bounds__(kBlock) void body_boxes_kernel(int nb, const int32_t* __restrict__ voff,
                                                             const double* __restrict__ x, double dil, int traj,
                                                             float* __restrict__ lo, float* __restrict__ hi) {
    const int b = blockIdx.x * (kBlock / kWave) + wave_id();
    if (b >= nb) return;  // the whole wave (no barrier)
    double mn[3], mx[3];
    wave_extrema(x, voff[b], voff[b + 1], mn, mx);
    if (lane_id() == 0)
        for (int k = 0; k < 3; ++k) {
            if (traj) {
                hi[3 * b + k] = fmaxf(hi[3 * b + k], (float)mx[k]);
                lo[3 * b + k] = fminf(lo[3 * b + k], (float)mn[k]);
            } else {
                lo[3 * b + k] = (float)(mn[k] - dil);
                hi[3 * b + k] = (float)(mx[k] + dil);
            }
        }
}

// Swept boxes with the collision-set dilation at both ends: the union of a primitive's
// dilated box at x0 and at x1 (same R3 arithmetic as the boxes above).  Every pair whose
// dilated boxes overlap at some x0 + t (x1 - x0), t in [0, 1], overlaps here, so the
// lists culled from them hold every candidate of any collision set along the step.
__global__ __launch_bounds__(kBlock) void swept_point_boxes_kernel(const double* __restrict__ x0,
                                                                    const double* __restrict__ x1, int64_t n,
                                                                    float dil, float* __restrict__ lo,
                                                                    float* __restrict__ hi) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < 3 * n; i += (int64_t)gridDim.x * kBlock) {
        const float c0 = (float)x0[i], c1 = (float)x1[i];
        lo[i] = fminf(c0 - dil, c1 - dil);
        hi[i] = fmaxf(c0 + dil, c1 + dil);
    }
}
template <int NV>
__global__ __launch_bounds__(kBlock) void swept_prim_boxes_kernel(const int32_t* __restrict__ prim_body,
                                                                   const int32_t* __restrict__ voff,
                                                                   const int32_t* __restrict__ topo,
                                                                   const double* __restrict__ x0,
                                                                   const double* __restrict__ x1, int64_t n,
                                                                   double dil, float* __restrict__ lo,
                                                                   float* __restrict__ hi) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        int vb = voff[prim_body[i]];
        int64_t vi[NV];
#pragma unroll
        for (int j = 0; j < NV; ++j) vi[j] = vb + topo[NV * i + j];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double mn0 = x0[3 * vi[0] + k], mx0 = mn0, mn1 = x1[3 * vi[0] + k], mx1 = mn1;
#pragma unroll
            for (int j = 1; j < NV; ++j) {
                const double c0 = x0[3 * vi[j] + k], c1 = x1[3 * vi[j] + k];
                mn0 = c0 < mn0 ? c0 : mn0;
                mx0 = c0 > mx0 ? c0 : mx0;
                mn1 = c1 < mn1 ? c1 : mn1;
                mx1 = c1 > mx1 ? c1 : mx1;
            }
            lo[3 * i + k] = fminf((float)(mn0 - dil), (float)(mn1 - dil));
            hi[3 * i + k] = fmaxf((float)(mx0 + dil), (float)(mx1 + dil));
        }
    }
}
// plo / phi (nullable): the body boxes at x0 already computed with the same dilation (the
// collision set's), read instead of recomputed — the same values.
__global__ __launch_bounds__(kBlock) void swept_body_boxes_kernel(int nb, const int32_t* __restrict__ voff,
                                                                   const double* __restrict__ x0,
                                                                   const double* __restrict__ x1, double dil,
                                                                   const float* __restrict__ plo,
                                                                   const float* __restrict__ phi,
                                                                   float* __restrict__ lo, float* __restrict__ hi) {
    const int b = blockIdx.x * (kBlock / kWave) + wave_id();
    if (b >= nb) return;  // the whole wave (no barrier)
    double mn0[3], mx0[3], mn1[3], mx1[3];
    if (!plo) wave_extrema(x0, voff[b], voff[b + 1], mn0, mx0);
    wave_extrema(x1, voff[b], voff[b + 1], mn1, mx1);
    if (lane_id() == 0)
        for (int k = 0; k < 3; ++k) {
            const float l0 = plo ? plo[3 * b + k] : (float)(mn0[k] - dil);
            const float h0 = phi ? phi[3 * b + k] : (float)(mx0[k] + dil);
            lo[3 * b + k] = fminf(l0, (float)(mn1[k] - dil));
            hi[3 * b + k] = fmaxf(h0, (float)(mx1[k] + dil));
        }
}

// abd.py:337-353 with the step length of the body's scene
__global__ void update_q_scenes_kernel(int nb, const int32_t* __restrict__ body_scene, const double* __restrict__ A,
                                       const double* __restrict__ p, const double* __restrict__ dq,
                                       const double* __restrict__ alpha, double* A_out, double* p_out,
                                       double* A_out2, double* p_out2) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 12 * nb) return;
    const int b = i / 12, r = i % 12;
    const double al = alpha[body_scene[b]];
    // a scene that does not move keeps its state bit for bit, whatever dq holds
    double* out = r < 3 ? p_out + 3 * b + r : A_out + 9 * b + r - 3;
    double* out2 = r < 3 ? (p_out2 ? p_out2 + 3 * b + r : nullptr) : (A_out2 ? A_out2 + 9 * b + r - 3 : nullptr);
    double v;
    if (r < 3)
        v = al == 0.0 ? p[3 * b + r] : p[3 * b + r] - dq[3 * (4 * b) + r] * al;
    else {
        const int row = (r - 3) / 3, k = (r - 3) % 3;
        v = al == 0.0 ? A[9 * b + 3 * row + k] : A[9 * b + 3 * row + k] - dq[3 * (4 * b + 1 + row) + k] * al;
    }
    *out = v;
    if (out2) *out2 = v;
}

int project_step(hipStream_t s, const AbdScene* sc, const double* A, const double* p, const double* dq,
                 const double* alpha_dev, double alpha_host, const uint8_t* active, double* x_out,
                 const double* x_other, const double* lo, const double* hi, int64_t* flag, int64_t gen) {
    if (sc->n_verts == 0) return ABD_OK;
    // vertex pairs with 16-byte accesses when every array allows them
    const auto a16 = [](const void* q) { return q == nullptr || ((uintptr_t)q & 15) == 0; };
    if (a16(sc->x0) && a16(x_out) && a16(x_other) && a16(lo) && a16(hi)) {
        hipLaunchKernelGGL(project_step2_kernel, dim3(grid_for(((int64_t)sc->n_verts + 1) / 2, kBlock, 1 << 20)),
                           dim3(kBlock), 0, s, sc->vert_body, sc->n_scenes > 1 ? sc->body_scene : nullptr, sc->x0,
                           A, p, dq, alpha_dev, alpha_host, active, x_out, x_other, lo, hi, flag, gen,
                           (int64_t)sc->n_verts);
        ABD_CHECK_LAUNCH();
        return ABD_OK;
    }
    hipLaunchKernelGGL(project_step_kernel, dim3(grid_for(sc->n_verts, kBlock, 1 << 20)), dim3(kBlock), 0, s,
                       sc->vert_body, sc->n_scenes > 1 ? sc->body_scene : nullptr, sc->x0, A, p, dq, alpha_dev,
                       alpha_host, active, x_out, x_other, lo, hi, flag, gen, (int64_t)sc->n_verts);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

// q_out = q - alpha dq, also written to (A_out2, p_out2) when they are given
int update_q_dual(hipStream_t s, int nb, const double* A, const double* p, const double* dq, double alpha,
                  double* A_out, double* p_out, double* A_out2, double* p_out2) {
    if (nb == 0) return ABD_OK;
    hipLaunchKernelGGL(update_q_kernel, dim3(grid_for(12 * (int64_t)nb)), dim3(kBlock), 0, s, nb, A, p, dq, alpha,
                       A_out, p_out, A_out2, p_out2);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

int update_q_scenes_dual(hipStream_t s, const AbdScene* sc, const double* A, const double* p, const double* dq,
                         const double* alpha, double* A_out, double* p_out, double* A_out2, double* p_out2) {
    const int nb = sc->n_bodies;
    if (nb == 0) return ABD_OK;
    hipLaunchKernelGGL(update_q_scenes_kernel, dim3(grid_for(12 * (int64_t)nb)), dim3(kBlock), 0, s, nb,
                       sc->body_scene, A, p, dq, alpha, A_out, p_out, A_out2, p_out2);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

__global__ __launch_bounds__(kBlock) void vertex_ranges_kernel(int64_t n3, const double* __restrict__ x0,
                                                                const double* __restrict__ x1, double skin,
                                                                double* __restrict__ lo, double* __restrict__ hi) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n3; i += (int64_t)gridDim.x * kBlock) {
        const double a = x0[i], b = x1[i];
        lo[i] = (a < b ? a : b) - skin;
        hi[i] = (a > b ? a : b) + skin;
    }
}

__global__ __launch_bounds__(kBlock) void within_ranges_kernel(int64_t n3, const double* __restrict__ x0,
                                                                const double* __restrict__ x1,
                                                                const double* __restrict__ lo,
                                                                const double* __restrict__ hi,
                                                                int64_t* __restrict__ flag) {
    bool out = false;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n3; i += (int64_t)gridDim.x * kBlock) {
        const double a = x0[i], b = x1[i], l = lo[i], h = hi[i];
        out |= !(a >= l && a <= h && b >= l && b <= h);  // NaN counts as outside
    }
    if (__any(out) && lane_id() == 0) *flag = 1;  // every writer stores 1
}

int vertex_ranges(hipStream_t s, int64_t n3, const double* x0, const double* x1, double skin, double* lo,
                  double* hi) {
    if (n3 <= 0) return ABD_OK;
    hipLaunchKernelGGL(vertex_ranges_kernel, dim3(grid_for(n3)), dim3(kBlock), 0, s, n3, x0, x1, skin, lo, hi);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

int within_ranges(hipStream_t s, int64_t n3, const double* x0, const double* x1, const double* lo, const double* hi,
                  int64_t* flag) {
    if (n3 <= 0) return ABD_OK;
    hipLaunchKernelGGL(within_ranges_kernel, dim3(grid_for(n3, kBlock, 4096)), dim3(kBlock), 0, s, n3, x0, x1, lo,
                       hi, flag);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

int swept_boxes(hipStream_t s, const AbdScene* sc, int kind, const double* x0, const double* x1, double dil,
                float* lo, float* hi, const float* plo, const float* phi) {
    if (!plo != !phi || (plo && kind != -1)) return ABD_ERR_ARG;
    if (kind == ABD_KIND_POINT) {
        if (!sc->n_verts) return ABD_OK;
        hipLaunchKernelGGL(swept_point_boxes_kernel, dim3(grid_for(3 * (int64_t)sc->n_verts)), dim3(kBlock), 0, s, x0,
                           x1, (int64_t)sc->n_verts, (float)dil, lo, hi);
    } else if (kind == ABD_KIND_EDGE) {
        if (!sc->n_edges) return ABD_OK;
        hipLaunchKernelGGL(swept_prim_boxes_kernel<2>, dim3(grid_for(sc->n_edges)), dim3(kBlock), 0, s,
                           sc->edge_body, sc->voff, sc->edges, x0, x1, (int64_t)sc->n_edges, dil, lo, hi);
    } else if (kind == ABD_KIND_TRI) {
        if (!sc->n_tris) return ABD_OK;
        hipLaunchKernelGGL(swept_prim_boxes_kernel<3>, dim3(grid_for(sc->n_tris)), dim3(kBlock), 0, s, sc->tri_body,
                           sc->voff, sc->tris, x0, x1, (int64_t)sc->n_tris, dil, lo, hi);
    } else if (kind == -1) {
        if (!sc->n_bodies) return ABD_OK;
        hipLaunchKernelGGL(swept_body_boxes_kernel, dim3((unsigned)ceil_div(sc->n_bodies, kBlock / kWave)),
                           dim3(kBlock), 0, s, sc->n_bodies, sc->voff, x0, x1, dil, plo, phi, lo, hi);
    } else {
        return ABD_ERR_ARG;
    }
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

}  // namespace abd

using namespace abd;

extern "C" {

int abd_project_vertices(abd_stream_t stream, const AbdScene* sc, const double* A, const double* p, double* x) {
    if (!sc || !A || !p || !x) return ABD_ERR_ARG;
    if (sc->n_verts == 0) return ABD_OK;
    int ph = prof_start(kProfProject, to_stream(stream), 52.0 * sc->n_verts + 96.0 * sc->n_bodies);
    hipLaunchKernelGGL(project_kernel, dim3(grid_for(sc->n_verts)), dim3(kBlock), 0, to_stream(stream),
                       sc->vert_body, sc->x0, A, p, x, (int64_t)sc->n_verts);
    ABD_CHECK_LAUNCH();
    prof_stop(ph, to_stream(stream));
    return ABD_OK;
}

int abd_update_q(abd_stream_t stream, int32_t nb, const double* A, const double* p, const double* dq, double alpha,
                 double* A_out, double* p_out) {
    if (nb < 0 || !A || !p || !dq || !A_out || !p_out) return ABD_ERR_ARG;
    if (nb == 0) return ABD_OK;
    return update_q_dual(to_stream(stream), nb, A, p, dq, alpha, A_out, p_out, nullptr, nullptr);
}

int abd_update_q0qdot(abd_stream_t stream, int32_t nb, double dt, const double* A, const double* p, double* A0,
                      double* p0, double* Adot, double* pdot) {
    if (nb < 0 || !A || !p || !A0 || !p0 || !Adot || !pdot) return ABD_ERR_ARG;
    if (nb == 0) return ABD_OK;
    hipLaunchKernelGGL(update_q0qdot_kernel, dim3(grid_for(12 * (int64_t)nb)), dim3(kBlock), 0, to_stream(stream),
                       nb, dt, A, p, A0, p0, Adot, pdot);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

int abd_prim_boxes(abd_stream_t stream, const AbdScene* sc, int32_t kind, const double* x, double dil, float* lo,
                   float* hi) {
    if (!sc || !x || !lo || !hi) return ABD_ERR_ARG;
    hipStream_t s = to_stream(stream);
    // each vertex read once (24 B) + per primitive its indices and body id and the box written
    // (24 B): the compulsory traffic; a vertex shared by several primitives is re-read from cache
    const double bytes = kind == ABD_KIND_POINT ? 48.0 * sc->n_verts
                         : kind == ABD_KIND_EDGE ? 24.0 * sc->n_verts + (12.0 + 24) * sc->n_edges
                                                 : 24.0 * sc->n_verts + (16.0 + 24) * sc->n_tris;
    const int64_t nprim = kind == ABD_KIND_POINT ? sc->n_verts : kind == ABD_KIND_EDGE ? sc->n_edges : sc->n_tris;
    if (kind < ABD_KIND_POINT || kind > ABD_KIND_TRI) return ABD_ERR_ARG;
    if (nprim == 0) return ABD_OK;
    int ph = prof_start(kProfBoxes, s, bytes);
    if (kind == ABD_KIND_POINT) {
        hipLaunchKernelGGL(point_boxes_kernel, dim3(grid_for(3 * (int64_t)sc->n_verts)), dim3(kBlock), 0, s, x,
                           (int64_t)sc->n_verts, (float)dil, lo, hi);
    } else if (kind == ABD_KIND_EDGE) {
        hipLaunchKernelGGL(prim_boxes_kernel<2>, dim3(grid_for(sc->n_edges)), dim3(kBlock), 0, s, sc->edge_body,
                           sc->voff, sc->edges, x, (int64_t)sc->n_edges, dil, lo, hi);
    } else {
        hipLaunchKernelGGL(prim_boxes_kernel<3>, dim3(grid_for(sc->n_tris)), dim3(kBlock), 0, s, sc->tri_body,
                           sc->voff, sc->tris, x, (int64_t)sc->n_tris, dil, lo, hi);
    }
    ABD_CHECK_LAUNCH();
    prof_stop(ph, s);
    return ABD_OK;
}

int abd_prim_boxes_traj(abd_stream_t stream, const AbdScene* sc, int32_t kind, const double* x_end, float* lo,
                        float* hi) {
    if (!sc || !x_end || !lo || !hi) return ABD_ERR_ARG;
    hipStream_t s = to_stream(stream);
    if (kind == ABD_KIND_POINT) {
        if (!sc->n_verts) return ABD_OK;
        hipLaunchKernelGGL(point_traj_kernel, dim3(grid_for(3 * (int64_t)sc->n_verts)), dim3(kBlock), 0, s, x_end,
                           (int64_t)sc->n_verts, lo, hi);
    } else if (kind == ABD_KIND_EDGE) {
        if (!sc->n_edges) return ABD_OK;
        hipLaunchKernelGGL(prim_traj_kernel<2>, dim3(grid_for(sc->n_edges)), dim3(kBlock), 0, s, sc->edge_body,
                           sc->voff, sc->edges, x_end, (int64_t)sc->n_edges, lo, hi);
    } else if (kind == ABD_KIND_TRI) {
        if (!sc->n_tris) return ABD_OK;
        hipLaunchKernelGGL(prim_traj_kernel<3>, dim3(grid_for(sc->n_tris)), dim3(kBlock), 0, s, sc->tri_body,
                           sc->voff, sc->tris, x_end, (int64_t)sc->n_tris, lo, hi);
    } else {
        return ABD_ERR_ARG;
    }
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

int abd_body_boxes(abd_stream_t stream, const AbdScene* sc, const double* x, double dil, float* lo, float* hi) {
    if (!sc || !x || !lo || !hi) return ABD_ERR_ARG;
    if (!sc->n_bodies) return ABD_OK;
    hipLaunchKernelGGL(body_boxes_kernel, dim3((unsigned)ceil_div(sc->n_bodies, kBlock / kWave)), dim3(kBlock), 0,
                       to_stream(stream), sc->n_bodies, sc->voff, x, dil, 0, lo, hi);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

int abd_body_boxes_traj(abd_stream_t stream, const AbdScene* sc, const double* x_end, float* lo, float* hi) {
    if (!sc || !x_end || !lo || !hi) return ABD_ERR_ARG;
    if (!sc->n_bodies) return ABD_OK;
    hipLaunchKernelGGL(body_boxes_kernel, dim3((unsigned)ceil_div(sc->n_bodies, kBlock / kWave)), dim3(kBlock), 0,
                       to_stream(stream), sc->n_bodies, sc->voff, x_end, 0.0, 1, lo, hi);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

int abd_update_q_scenes(abd_stream_t stream, const AbdScene* sc, const double* A, const double* p, const double* dq,
                        const double* alpha, double* A_out, double* p_out) {
    if (!sc || sc->n_scenes <= 0 || !sc->body_scene || !A || !p || !dq || !alpha || !A_out || !p_out) return ABD_ERR_ARG;
    return update_q_scenes_dual(to_stream(stream), sc, A, p, dq, alpha, A_out, p_out, nullptr, nullptr);
}

}  // extern "C"
