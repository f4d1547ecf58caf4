/*
 * abd_oracle.c — CPU FP64 restatement of the abd-warp per-Newton-iteration
 * contact path.  TEST INFRASTRUCTURE ONLY: it is the checker that the HIP
 * path is compared against (tests/, __graft_entry__.smoke, bench.py's
 * cpu_baseline leg).  Nothing in the product imports, links or calls it.
 *
 * Every function follows the reference source it cites (paths relative to the
 * reference checkout, file:line).  The reference is Python on NVIDIA Warp and
 * cannot run here (no warp, no abdtk); this file re-derives its arithmetic in
 * plain C with the same operation order, built with -ffp-contract=off so that
 * no multiply-add is fused on either side.  The per-pair 12x12 assembly is
 * deliberately done the reference way (dense Kronecker products and
 * Q Lambda Q^-1, ipcsk/pt_ipc.py:194-246), not the rank-one way the GPU uses,
 * so the two paths share no arithmetic beyond the per-pair geometry.
 *
 * Restatement rules kept from the reference:
 *   R1 Warp matrix-from-vectors takes COLUMN vectors (mat33(v1,v2,v3)[r][c] = v_c[r]).
 *   R2 wp.select(c, a, b) returns b when c holds, else a.
 *   R3 FP32 boxes: round-to-nearest casts in the reference's operation order.
 *   R4 2x2 inverse = adjugate * (1/det), zero matrix when |det| <= 0 (Warp mat22
 *      inverse; the epsilon is Warp-internal and assumed 0).
 *   R5 FP64 everywhere except boxes; library log/sqrt.
 *   R6 distances per call site (PT energy |e2p|^2, PT gh l^2, EE energy l^2,
 *      EE gh |e2p|^2).
 * Defect fixes (the reference's orchestration bugs) are marked "Dn".
 *
 * Parity pins: each function is checked in tests/test_oracle_*.py against a
 * closed form, finite differences, numpy eigh, brute force or a known-answer
 * construction.  Functions with no such check are listed as "parity unpinned"
 * below:
 *   - orc_bicgstab: iterate sequence and stopping cadence of Warp's bicgstab
 *     (external, unvendored); pinned only by the residual bound and by
 *     agreement with numpy.linalg.solve.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "abd_oracle.h"
#include <float.h>

typedef struct { double x, y, z; } v3;

static inline v3 mk(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
/* Warp scalar*vec multiplies each component by the scalar */
static inline v3 scl(v3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline v3 dvs(v3 a, double s) { return mk(a.x / s, a.y / s, a.z / s); }
/* wp.dot: ((a.x*b.x + a.y*b.y) + a.z*b.z) */
static inline double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross(v3 a, v3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double len(v3 a) { return sqrt(dot(a, a)); }
static inline v3 ld3(const double* p) { return mk(p[0], p[1], p[2]); }
static inline void st3(double* p, v3 a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }
static inline double comp(v3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }


/* --------------------------------------------------------- vertex mapping */
/* abd.py:305-324 update_mesh_vertex: x = A^T x0 + p, i.e.
 * x_k = ((p_k + x0_0*A[0][k]) + x0_1*A[1][k]) + x0_2*A[2][k] (row r of A is a_{r+1}, R1 at abd.py:344).
 * D12: used for every field (assign_to's A x0 is not used). */
void orc_project(int nb, const int32_t* voff, const double* x0, const double* A, const double* p, double* x) {
    for (int b = 0; b < nb; ++b) {
        const double* Ab = A + 9 * b;
        for (int v = voff[b]; v < voff[b + 1]; ++v) {
            for (int k = 0; k < 3; ++k) {
                double r = p[3 * b + k] + x0[3 * v + 0] * Ab[0 + k];
                r = r + x0[3 * v + 1] * Ab[3 + k];
                r = r + x0[3 * v + 2] * Ab[6 + k];
                x[3 * v + k] = r;
            }
        }
    }
}

/* ------------------------------------------------------------------ boxes */
/* kind: 0 point, 1 edge, 2 triangle.  offs: prim offsets per body, voff: vertex
 * offsets, topo: local vertex indices (edges [E,2], triangles [F,3]). */
static inline float f32(double v) { return (float)v; }

/* culling2.py:33-50 (triangles), :53-69 (edges), :72-79 (points) */
void orc_prim_boxes(int kind, int nb, const int32_t* voff, const int32_t* offs, const int32_t* topo,
                    const double* x, double dil, float* lo, float* hi) {
    for (int b = 0; b < nb; ++b) {
        int n0 = kind == 0 ? voff[b] : offs[b];
        int n1 = kind == 0 ? voff[b + 1] : offs[b + 1];
        for (int i = n0; i < n1; ++i) {
            if (kind == 0) {
                /* wp.vec3f(V) - wp.vec3f(f32(dil)): cast first, FP32 arithmetic */
                float d = f32(dil);
                for (int k = 0; k < 3; ++k) {
                    float c = f32(x[3 * i + k]);
                    lo[3 * i + k] = c - d;
                    hi[3 * i + k] = c + d;
                }
            } else {
                int nv = kind == 1 ? 2 : 3;
                for (int k = 0; k < 3; ++k) {
                    double mn = 0, mx = 0;
                    for (int j = 0; j < nv; ++j) {
                        double c = x[3 * (voff[b] + topo[nv * i + j]) + k];
                        if (j == 0 || c < mn) mn = c;
                        if (j == 0 || c > mx) mx = c;
                    }
                    /* FP64 min/max +- dilation, then cast */
                    lo[3 * i + k] = f32(mn - dil);
                    hi[3 * i + k] = f32(mx + dil);
                }
            }
        }
    }
}

/* culling2.py:230-250, :254-270, :273-280 — union of the stored (start) box with
 * the undilated endpoint box, FP32 max/min of f32 casts. */
void orc_prim_boxes_traj(int kind, int nb, const int32_t* voff, const int32_t* offs, const int32_t* topo,
                         const double* x_end, float* lo, float* hi) {
    for (int b = 0; b < nb; ++b) {
        int n0 = kind == 0 ? voff[b] : offs[b];
        int n1 = kind == 0 ? voff[b + 1] : offs[b + 1];
        for (int i = n0; i < n1; ++i) {
            for (int k = 0; k < 3; ++k) {
                double mn, mx;
                if (kind == 0) {
                    mn = mx = x_end[3 * i + k];
                } else {
                    int nv = kind == 1 ? 2 : 3;
                    mn = mx = x_end[3 * (voff[b] + topo[nv * i]) + k];
                    for (int j = 1; j < nv; ++j) {
                        double c = x_end[3 * (voff[b] + topo[nv * i + j]) + k];
                        if (c < mn) mn = c;
                        if (c > mx) mx = c;
                    }
                }
                float fl = f32(mn), fh = f32(mx);
                if (fh > hi[3 * i + k]) hi[3 * i + k] = fh;
                if (fl < lo[3 * i + k]) lo[3 * i + k] = fl;
            }
        }
    }
}

/* culling2.py:171-186 update_body_bvh: numpy FP64 max/min over the body's
 * vertices +- dilation, cast to f32 on assign. */
void orc_body_boxes(int nb, const int32_t* voff, const double* x, double dil, float* lo, float* hi) {
    for (int b = 0; b < nb; ++b) {
        for (int k = 0; k < 3; ++k) {
            double mn = x[3 * voff[b] + k], mx = mn;
            for (int v = voff[b] + 1; v < voff[b + 1]; ++v) {
                double c = x[3 * v + k];
                if (c < mn) mn = c;
                if (c > mx) mx = c;
            }
            lo[3 * b + k] = f32(mn - dil);
            hi[3 * b + k] = f32(mx + dil);
        }
    }
}

/* culling2.py:188-210 body_bvh_to_traj: np.maximum(FP64 max of x_end, stored
 * f32 box) then cast; equal to the FP32 max of the cast values (monotone rounding). */
void orc_body_boxes_traj(int nb, const int32_t* voff, const double* x_end, float* lo, float* hi) {
    for (int b = 0; b < nb; ++b) {
        for (int k = 0; k < 3; ++k) {
            double mn = x_end[3 * voff[b] + k], mx = mn;
            for (int v = voff[b] + 1; v < voff[b + 1]; ++v) {
                double c = x_end[3 * v + k];
                if (c < mn) mn = c;
                if (c > mx) mx = c;
            }
            double U = mx > (double)hi[3 * b + k] ? mx : (double)hi[3 * b + k];
            double L = mn < (double)lo[3 * b + k] ? mn : (double)lo[3 * b + k];
            hi[3 * b + k] = f32(U);
            lo[3 * b + k] = f32(L);
        }
    }
}

/* ---------------------------------------------------------------- culling */
/* Closed-interval AABB overlap (Warp bvh_query_aabb leaf test; assumed closed). */
static inline int overlap(const float* alo, const float* ahi, const float* blo, const float* bhi) {
    for (int k = 0; k < 3; ++k)
        if (alo[k] > bhi[k] || ahi[k] < blo[k]) return 0;
    return 1;
}

/* culling2.py:214-227 intersection_bodies: every (bi, bj) with bj > bi whose
 * body boxes overlap.  Emitted in lexicographic order (the reference's order
 * follows BVH traversal and is not pinned).  body_scene (nullable) restricts
 * pairs to bodies of the same scene.  Returns the count; writes min(count, cap). */
long orc_cull_bodies(int nb, const int32_t* body_scene, const float* lo, const float* hi, int32_t* ij, long cap) {
    long n = 0;
    for (int i = 0; i < nb; ++i)
        for (int j = i + 1; j < nb; ++j) {
            if (body_scene && body_scene[i] != body_scene[j]) continue;
            if (overlap(lo + 3 * i, hi + 3 * i, lo + 3 * j, hi + 3 * j)) {
                if (n < cap) { ij[2 * n] = i; ij[2 * n + 1] = j; }
                ++n;
            }
        }
    return n;
}

/* culling.py:32-52 cull_ and :86-101 intersection_prims (D1: the live stub
 * cull() is replaced by the BVH semantics): for each body pair ii = (I0, I1),
 * PT (pt=1) emits (I, J, ii, p, t) for every point p of I whose box overlaps
 * triangle t of J, for both orders (I, J) = (I0, I1) then (I1, I0)
 * (culling.py:88-91, R2); EE (pt=0) emits (I0, I1, ii, a, b) for edge a of I0
 * overlapping edge b of I1.  Order: (ii, direction, p, t) ascending.
 * off1/off2: per-body prim offsets of box set 1 / 2. */
long orc_cull_prims(int pt, const int32_t* ij, long nij, const int32_t* off1, const float* lo1, const float* hi1,
                    const int32_t* off2, const float* lo2, const float* hi2, int32_t* out, long cap) {
    long n = 0;
    for (long ii = 0; ii < nij; ++ii) {
        for (int dir = 0; dir < (pt ? 2 : 1); ++dir) {
            int I = dir == 0 ? ij[2 * ii] : ij[2 * ii + 1];
            int J = ij[2 * ii] + ij[2 * ii + 1] - I;
            for (int a = off1[I]; a < off1[I + 1]; ++a)
                for (int b = off2[J]; b < off2[J + 1]; ++b)
                    if (overlap(lo1 + 3 * a, hi1 + 3 * a, lo2 + 3 * b, hi2 + 3 * b)) {
                        if (n < cap) {
                            int32_t* o = out + 5 * n;
                            o[0] = I; o[1] = J; o[2] = (int32_t)ii; o[3] = a - off1[I]; o[4] = b - off2[J];
                        }
                        ++n;
                    }
        }
    }
    return n;
}

/* culling.py:54-76 cull_vg + intersection_ground: bodies whose box lower y is
 * below ground + dhat/2 (FP64 compare, :67) query their point boxes against
 * [(-1e3,-1e3,-1e3), (1e3, f32(ground), 1e3)] (:68-72); emits (b, v) in
 * (b, v) order. */
long orc_cull_vg(int nb, const int32_t* voff, const float* blo, const float* plo, const float* phi,
                 double dhat, double ground, int32_t* out, long cap) {
    long n = 0;
    float qlo[3] = {-1e3f, -1e3f, -1e3f};
    float qhi[3] = {1e3f, f32(ground), 1e3f};
    for (int b = 0; b < nb; ++b) {
        if (!((double)blo[3 * b + 1] < 0.5 * dhat + ground)) continue;
        for (int v = voff[b]; v < voff[b + 1]; ++v)
            if (overlap(plo + 3 * v, phi + 3 * v, qlo, qhi)) {
                if (n < cap) { out[2 * n] = b; out[2 * n + 1] = v - voff[b]; }
                ++n;
            }
    }
    return n;
}

/* --------------------------------------------------------- distance maps */
/* R4: Warp mat22 inverse. */
static void inv22_mul(double m00, double m01, double m10, double m11, double b0, double b1, double* r0, double* r1) {
    double det = m00 * m11 - m10 * m01;
    double i00 = 0, i01 = 0, i10 = 0, i11 = 0;
    if (fabs(det) > 0.0) {
        double s = 1.0 / det;
        i00 = m11 * s; i01 = -m01 * s; i10 = -m10 * s; i11 = m00 * s;
    }
    *r0 = i00 * b0 + i01 * b1;
    *r1 = i10 * b0 + i11 * b1;
}

/* psd/vf.py:4-20 beta_gamma_pt and :23-41 C_vf (x0 point, x1..x3 triangle) */
static void bg_pt(v3 x0, v3 x1, v3 x2, v3 x3, double* beta, double* gamma) {
    v3 e0 = sub(x1, x2), e1 = sub(x3, x2), e2 = sub(x0, x2);
    double e0e0 = dot(e0, e0), e0e1 = dot(e0, e1), e1e1 = dot(e1, e1);
    inv22_mul(e0e0, e0e1, e0e1, e1e1, dot(e0, e2), dot(e1, e2), beta, gamma);
}
static void C_vf(v3 x0, v3 x1, v3 x2, v3 x3, v3* e0p, v3* e1p, v3* e2p) {
    v3 e0 = sub(x1, x2), e1 = sub(x3, x2), e2 = sub(x0, x2);
    double alpha = dot(e0, e1) / dot(e0, e0);
    v3 e1perp = sub(e1, scl(e0, alpha));
    double be, ga;
    double e0e0 = dot(e0, e0), e0e1 = dot(e0, e1), e1e1 = dot(e1, e1);
    inv22_mul(e0e0, e0e1, e0e1, e1e1, dot(e0, e2), dot(e1, e2), &be, &ga);
    *e0p = e0; *e1p = e1perp;
    *e2p = sub(sub(e2, scl(e0, be)), scl(e1, ga));
}
/* psd/ee.py:119-133 beta_gamma_ee and :6-31 C_ee (edge a = x0x1, edge b = x2x3) */
static void bg_ee(v3 x0, v3 x1, v3 x2, v3 x3, double* beta, double* gamma) {
    v3 e0 = sub(x1, x0), e1 = sub(x3, x2), e2 = sub(x2, x0);
    double e0e0 = dot(e0, e0), e0e1 = dot(e0, e1), e1e1 = dot(e1, e1);
    inv22_mul(e0e0, e0e1, e0e1, e1e1, dot(e0, e2), dot(e1, e2), beta, gamma);
}
static void C_ee(v3 x0, v3 x1, v3 x2, v3 x3, v3* e0p, v3* e1p, v3* e2p) {
    v3 e0 = sub(x1, x0), e1 = sub(x3, x2), e2 = sub(x2, x0);
    double alpha = dot(e0, e1) / dot(e0, e0);
    v3 e1perp = sub(e1, scl(e0, alpha));
    double be, ga;
    double e0e0 = dot(e0, e0), e0e1 = dot(e0, e1), e1e1 = dot(e1, e1);
    inv22_mul(e0e0, e0e1, e0e1, e1e1, dot(e0, e2), dot(e1, e2), &be, &ga);
    *e0p = e0; *e1p = e1perp;
    *e2p = sub(sub(e2, scl(e0, be)), scl(e1, ga));
}
/* ccd.py:45-48 verify_root_pt, :51-54 verify_root_ee (strict) */
static int valid_pt(v3 x0, v3 x1, v3 x2, v3 x3) {
    double b, g; bg_pt(x0, x1, x2, x3, &b, &g);
    return 0.0 < b && b < 1.0 && 0.0 < g && g < 1.0 && 0.0 < b + g && b + g < 1.0;
}
static int valid_ee(v3 x0, v3 x1, v3 x2, v3 x3) {
    double b, g; bg_ee(x0, x1, x2, x3, &b, &g);
    return 0.0 < b && b < 1.0 && 0.0 < -g && -g < 1.0;
}
/* psd/hl.py:16-17 */
static double signed_distance(v3 e0p, v3 e1p, v3 e2p) {
    v3 n = cross(e0p, e1p);
    return dot(e2p, n) / len(n);
}
/* psd/vf.py:44-72 dcvfdx_s; D row-major 3x4 */
static void dcdx_pt(v3 x0, v3 x1, v3 x2, v3 x3, double D[12]) {
    v3 e0 = sub(x1, x2), e1 = sub(x3, x2);
    double alpha = dot(e0, e1) / dot(e0, e0);
    double be, ga; bg_pt(x0, x1, x2, x3, &be, &ga);
    memset(D, 0, 12 * sizeof(double));
    D[0 * 4 + 1] = 1.0; D[0 * 4 + 2] = -1.0;
    D[1 * 4 + 1] = -alpha; D[1 * 4 + 2] = alpha - 1.0; D[1 * 4 + 3] = 1.0;
    D[2 * 4 + 0] = 1.0; D[2 * 4 + 1] = -be; D[2 * 4 + 2] = be + ga - 1.0; D[2 * 4 + 3] = -ga;
}
/* psd/ee.py:135-163 dceedx_s */
static void dcdx_ee(v3 x0, v3 x1, v3 x2, v3 x3, double D[12]) {
    v3 e0 = sub(x1, x0), e1 = sub(x3, x2);
    double al = dot(e0, e1) / dot(e0, e0);
    double be, ga; bg_ee(x0, x1, x2, x3, &be, &ga);
    D[0] = -1.0; D[1] = 1.0; D[2] = 0.0; D[3] = 0.0;
    D[4] = al; D[5] = -al; D[6] = -1.0; D[7] = 1.0;
    D[8] = be - 1.0; D[9] = -be; D[10] = ga + 1.0; D[11] = -ga;
}
/* psd/hl.py:156-195 eig_Hl_tid + the packaging of ipcsk/pt_ipc.py:259-280 /
 * ipcsk/ee_ipc.py:26-47.  Q: 15 vec3 (5 eigenvectors of 9), lam: 5. */
static void Q_lambda(v3 e0p, v3 e1p, v3 e2p, int clamp, double Q[45], double lam[5]) {
    double l = signed_distance(e0p, e1p, e2p);
    double e0pn = dot(e0p, e0p), e1pn = dot(e1p, e1p), e2pn = dot(e2p, e2p);
    double f12 = sqrt(1.0 + 4.0 * (e1pn / e2pn));
    double f02 = sqrt(1.0 + 4.0 * (e0pn / e2pn));
    double lam0 = -l / (2.0 * e1pn) * (1.0 + f12);
    double lam1 = -l / (2.0 * e1pn) * (1.0 - f12);
    double lam2 = -l / (2.0 * e0pn) * (1.0 + f02);
    double lam3 = -l / (2.0 * e0pn) * (1.0 - f02);
    double om0 = lam0 / (lam0 - l / e2pn);
    double om1 = lam1 / (lam1 - l / e2pn);
    double om2 = lam2 / (lam2 - l / e2pn);
    double om3 = lam3 / (lam3 - l / e2pn);
    v3 z = mk(0, 0, 0);
    v3 q[15];
    q[0] = z; q[1] = e2p; q[2] = scl(e1p, om0);
    q[3] = z; q[4] = e2p; q[5] = scl(e1p, om1);
    q[6] = e2p; q[7] = z; q[8] = scl(e0p, om2);
    q[9] = e2p; q[10] = z; q[11] = scl(e0p, om3);
    double L[4] = {lam0 * 2.0 * l, lam1 * 2.0 * l, lam2 * 2.0 * l, lam3 * 2.0 * l};
    if (clamp)  /* pt_ipc.py:266-269; EE leaves them unclamped (ee_ipc.py:33-36, D9) */
        for (int k = 0; k < 4; ++k) L[k] = L[k] > 0.0 ? L[k] : 0.0;
    /* psd/hl.py:37-39 gl = (0, 0, e2o * l / |e2o|^2); q4 = gl * 2 * l */
    v3 gl2 = dvs(scl(e2p, l), dot(e2p, e2p));
    q[12] = scl(scl(z, 2.0), l); q[13] = scl(scl(z, 2.0), l); q[14] = scl(scl(gl2, 2.0), l);
    for (int k = 0; k < 4; ++k) lam[k] = L[k];
    lam[4] = 2.0;
    for (int i = 0; i < 15; ++i) { Q[3 * i] = q[i].x; Q[3 * i + 1] = q[i].y; Q[3 * i + 2] = q[i].z; }
}

/* ---------------------------------------------------------------- barrier */
/* ipcsk/barrier.py:4-15, :21-31, :37-48 (s = squared distance) */
double orc_barrier(double d, double d2hat, double kappa) {
    double ret = 0.0;
    if (d < d2hat) {
        double r = d / d2hat;
        ret = kappa * -pow(r - 1.0, 2.0) * log(r);
    }
    return ret;
}
double orc_barrier_d1(double d, double d2hat, double kappa) {
    double ret = 0.0;
    if (d < d2hat) ret = kappa * (d2hat - d) * (2.0 * log(d / d2hat) + (d - d2hat) / d) / (d2hat * d2hat);
    return ret;
}
double orc_barrier_d2(double d, double d2hat, double kappa) {
    double ret = 0.0;
    if (d < d2hat)
        ret = -kappa * (2.0 * log(d / d2hat) + (d - d2hat) / d + (d - d2hat) * (2.0 / d + d2hat / (d * d))) /
              (d2hat * d2hat);
    return ret;
}

/* ------------------------------------------------------ fetch (A1 fetchers) */

/* affine_body.py:148-164 fetch_pt (and the x0/xk twins :166-200) */
static void fetch_pt(const orc_scene* S, const double* x, const int32_t* t5, v3 v[4]) {
    int I = t5[0], J = t5[1];
    const int32_t* T = S->tris + 3 * (S->foff[J] + t5[4]);
    v[0] = ld3(x + 3 * (S->voff[I] + t5[3]));
    for (int k = 0; k < 3; ++k) v[1 + k] = ld3(x + 3 * (S->voff[J] + T[k]));
}
/* affine_body.py:202-221 fetch_ee (and twins :223-263) */
static void fetch_ee(const orc_scene* S, const double* x, const int32_t* t5, v3 v[4]) {
    int I = t5[0], J = t5[1];
    const int32_t* EI = S->edges + 2 * (S->eoff[I] + t5[3]);
    const int32_t* EJ = S->edges + 2 * (S->eoff[J] + t5[4]);
    v[0] = ld3(x + 3 * (S->voff[I] + EI[0]));
    v[1] = ld3(x + 3 * (S->voff[I] + EI[1]));
    v[2] = ld3(x + 3 * (S->voff[J] + EJ[0]));
    v[3] = ld3(x + 3 * (S->voff[J] + EJ[1]));
}

/* ----------------------------------------------------------------- energy */
/* full-IPC mode (psd = 3, no reference counterpart): abd_oracle_full.c */
double orc_full_energy(int kind, const double* x, const double* r, double d2hat, double kappa);
int orc_full_local(int kind, const double* x, const double* r, double d2hat, double kappa, int project, double* E,
                   double* g12, double* H12, int* feat_out);
static double full_energy_list(const orc_scene* S, const orc_params* P, const double* xk, const int32_t* t5s, long n,
                               int kind) {
    double E = 0.0;
    for (long i = 0; i < n; ++i) {
        v3 v[4], r[4];
        double x12[12], r12[12];
        if (kind == 0) {
            fetch_pt(S, xk, t5s + 5 * i, v);
            fetch_pt(S, S->x0, t5s + 5 * i, r);
        } else {
            fetch_ee(S, xk, t5s + 5 * i, v);
            fetch_ee(S, S->x0, t5s + 5 * i, r);
        }
        for (int t = 0; t < 4; ++t) { st3(x12 + 3 * t, v[t]); st3(r12 + 3 * t, r[t]); }
        E += orc_full_energy(kind, x12, r12, P->d2hat, P->kappa);
    }
    return E;
}

/* ipcsk/energy.py:37-48 ipc_energy_pt (|e2p|^2, valid only) */
double orc_energy_pt(const orc_scene* S, const orc_params* P, const double* xk, const int32_t* pt, long n) {
    if (P->psd == 3) return full_energy_list(S, P, xk, pt, n, 0);
    double E = 0.0;
    for (long i = 0; i < n; ++i) {
        v3 v[4]; fetch_pt(S, xk, pt + 5 * i, v);
        int cond = valid_pt(v[0], v[1], v[2], v[3]);
        v3 e0p, e1p, e2p; C_vf(v[0], v[1], v[2], v[3], &e0p, &e1p, &e2p);
        double d2 = dot(e2p, e2p);
        if (cond) E += orc_barrier(d2, P->d2hat, P->kappa);
    }
    return E;
}
/* ipcsk/energy.py:10-23 ipc_energy_ee (l^2, valid only) */
double orc_energy_ee(const orc_scene* S, const orc_params* P, const double* xk, const int32_t* ee, long n) {
    if (P->psd == 3) return full_energy_list(S, P, xk, ee, n, 1);
    double E = 0.0;
    for (long i = 0; i < n; ++i) {
        v3 v[4]; fetch_ee(S, xk, ee + 5 * i, v);
        int cond = valid_ee(v[0], v[1], v[2], v[3]);
        v3 e0p, e1p, e2p; C_ee(v[0], v[1], v[2], v[3], &e0p, &e1p, &e2p);
        double d = signed_distance(e0p, e1p, e2p);
        double d2 = d * d;
        if (cond) E += orc_barrier(d2, P->d2hat, P->kappa);
    }
    return E;
}
/* ipcsk/energy.py:25-35 ipc_energy_vg (no validity; affine_body.py:102-103 vg_distance) */
double orc_energy_vg(const orc_scene* S, const orc_params* P, const double* xk, const int32_t* vg, long n) {
    double E = 0.0;
    for (long i = 0; i < n; ++i) {
        double y = xk[3 * (S->voff[vg[2 * i]] + vg[2 * i + 1]) + 1];
        double d = y - P->ground;
        E += orc_barrier(d * d, P->d2hat, P->kappa);
    }
    return E;
}
/* ipc.py:19-46 IPCContactEnergy.energy (D3: correct lists) */
double orc_contact_energy(const orc_scene* S, const orc_params* P, const double* xk, const int32_t* pt, long npt,
                          const int32_t* ee, long nee, const int32_t* vg, long nvg) {
    double E = 0.0;
    if (P->ee) E += orc_energy_ee(S, P, xk, ee, nee);
    if (P->pt) E += orc_energy_pt(S, P, xk, pt, npt);
    if (P->vg) E += orc_energy_vg(S, P, xk, vg, nvg);
    return E;
}

/* ---------------------------------------------------- per-pair gh (dense) */
/* kron(M (r x c), I3) -> (3r x 3c) */
static void kron3(const double* M, int r, int c, double* out) {
    for (int i = 0; i < 3 * r; ++i)
        for (int j = 0; j < 3 * c; ++j)
            out[i * 3 * c + j] = (i % 3 == j % 3) ? M[(i / 3) * c + (j / 3)] : 0.0;
}
/* C(m x n) = A(m x k) @ B(k x n) */
static void matmul(const double* A, const double* B, double* C, int m, int k, int n) {
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int t = 0; t < k; ++t) s += A[i * k + t] * B[t * n + j];
            C[i * n + j] = s;
        }
}
static void transpose(const double* A, int m, int n, double* T) {
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) T[j * m + i] = A[i * n + j];
}

/* Reference per-contact assembly, ipcsk/pt_ipc.py:111-185 (PT) and
 * ipcsk/ee_ipc.py:94-143 (EE), with helpers extract_g/JTg (:194-209 / :162-177),
 * extract_Q, QLQinv (:221-231 / :146-161), dcTHldc (:233-235 / :154-157), JTH12J
 * (:237-246 / :178-190).  npI: number of contact vertices on body I (PT 1, EE 2).
 * JI: npI x 4, JJ: (4-npI) x 4 rows [1, x0^T].  Outputs g24 (24),
 * Hi/Hj/Hij (12x12 row-major). */
static void map_pair(const double g12[12], const double Hipc[144], int npI, const double* JI, const double* JJ,
                     double B1, double g24[24], double Hi[144], double Hj[144], double Hij[144]);

/* Beyond the reference (psd = 2): the exact derivatives of the squared plane /
 * line distance s = T^2/N, T = det[u v w] = w.(u x v), N = |u|^2|v|^2 - (u.v)^2,
 * over the edge vectors (u, v, w) of C_vf / C_ee (the reference's dc/dx treats
 * alpha, beta, gamma as constants and so drops the curvature of the projection).
 * g: ds/d(u,v,w) (9); H: 9x9 row-major. */
static void skew3(v3 a, double sgn, double* H, int r0, int c0) {
    H[9 * r0 + c0 + 1] -= sgn * a.z; H[9 * r0 + c0 + 2] += sgn * a.y;
    H[9 * (r0 + 1) + c0] += sgn * a.z; H[9 * (r0 + 1) + c0 + 2] -= sgn * a.x;
    H[9 * (r0 + 2) + c0] -= sgn * a.y; H[9 * (r0 + 2) + c0 + 1] += sgn * a.x;
}
static void exact_dist2(v3 u, v3 v, v3 w, double g[9], double H[81]) {
    double uu = dot(u, u), vv = dot(v, v), uv = dot(u, v);
    double T = dot(w, cross(u, v));
    double N = uu * vv - uv * uv;
    double dT[9], dN[9] = {0}, d2N[81] = {0}, d2T[81] = {0};
    v3 t0 = cross(v, w), t1 = cross(w, u), t2 = cross(u, v);
    for (int k = 0; k < 3; ++k) {
        dT[k] = comp(t0, k); dT[3 + k] = comp(t1, k); dT[6 + k] = comp(t2, k);
        dN[k] = 2.0 * vv * comp(u, k) - 2.0 * uv * comp(v, k);
        dN[3 + k] = 2.0 * uu * comp(v, k) - 2.0 * uv * comp(u, k);
    }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double I = r == c ? 1.0 : 0.0;
            d2N[9 * r + c] = 2.0 * vv * I - 2.0 * comp(v, r) * comp(v, c);
            d2N[9 * (3 + r) + 3 + c] = 2.0 * uu * I - 2.0 * comp(u, r) * comp(u, c);
            d2N[9 * r + 3 + c] = 4.0 * comp(u, r) * comp(v, c) - 2.0 * comp(v, r) * comp(u, c) - 2.0 * uv * I;
            d2N[9 * (3 + c) + r] = d2N[9 * r + 3 + c];
        }
    /* T = u.(v x w): d2T/dudv = -[w]x, dudw = [v]x, dvdw = -[u]x, and transposes */
    skew3(w, -1.0, d2T, 0, 3); skew3(w, 1.0, d2T, 3, 0);
    skew3(v, 1.0, d2T, 0, 6); skew3(v, -1.0, d2T, 6, 0);
    skew3(u, -1.0, d2T, 3, 6); skew3(u, 1.0, d2T, 6, 3);
    for (int i = 0; i < 9; ++i) g[i] = 2.0 * T / N * dT[i] - T * T / (N * N) * dN[i];
    for (int i = 0; i < 9; ++i)
        for (int j = 0; j < 9; ++j)
            H[9 * i + j] = 2.0 / N * dT[i] * dT[j] + 2.0 * T / N * d2T[9 * i + j]
                           - 2.0 * T / (N * N) * (dT[i] * dN[j] + dN[i] * dT[j]) - T * T / (N * N) * d2N[9 * i + j]
                           + 2.0 * T * T / (N * N * N) * dN[i] * dN[j];
}

static void pair_gh_dense(const double Q[45], const double lam[5], const double D[12], int npI, const double* JI,
                          const double* JJ, double B1, double B2, int psd, double g24[24], double Hi[144],
                          double Hj[144], double Hij[144]) {
    /* extract_Q: Q.reshape(5,9).T -> 9x5 */
    double Qm[45];
    for (int k = 0; k < 5; ++k)
        for (int r = 0; r < 9; ++r) Qm[r * 5 + k] = Q[9 * k + r];
    /* QLQinv */
    double QT[45], QTQ[25];
    transpose(Qm, 9, 5, QT);
    matmul(QT, Qm, QTQ, 5, 9, 5);
    double Dinv[25] = {0}, Lm[25] = {0};
    for (int i = 0; i < 5; ++i) { Dinv[i * 5 + i] = 1.0 / QTQ[i * 5 + i]; Lm[i * 5 + i] = lam[i]; }
    if (psd) {
        /* Beyond the reference: eigenvalues of the whole contact-space barrier Hessian
         * b' Hl + b'' q4 q4^T (q4 is orthogonal to q0..q3), clamped at zero; the b''
         * term is then part of Hl and not added again below. */
        for (int i = 0; i < 5; ++i) {
            double mu = B1 * lam[i] + (i == 4 ? B2 * QTQ[i * 5 + i] : 0.0);
            Lm[i * 5 + i] = mu > 0.0 ? mu : 0.0;
        }
    }
    double Qinv[45], QL[45], Hl[81];
    matmul(Dinv, QT, Qinv, 5, 5, 9);
    matmul(Qm, Lm, QL, 9, 5, 5);
    matmul(QL, Qinv, Hl, 9, 5, 9);
    /* dcTHldc */
    double K[108], KT[108], tmp[108], H12[144];
    kron3(D, 3, 4, K);  /* 9 x 12 */
    transpose(K, 9, 12, KT);
    matmul(KT, Hl, tmp, 12, 9, 9);
    matmul(tmp, K, H12, 12, 9, 12);
    /* extract_g: gl = Q[12..14]; g12 = kron(dcdx^T, I3) @ gl */
    double DT[12], KD[108], g12[12];
    transpose(D, 3, 4, DT);
    kron3(DT, 4, 3, KD);  /* 12 x 9 */
    double gl[9];
    for (int i = 0; i < 9; ++i) gl[i] = Q[36 + i];
    matmul(KD, gl, g12, 12, 9, 1);
    /* Hipc = H * B' + outer(g12, g12) * B'' */
    double Hipc[144];
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 12; ++j)
            Hipc[i * 12 + j] = psd ? H12[i * 12 + j] : H12[i * 12 + j] * B1 + (g12[i] * g12[j]) * B2;
    map_pair(g12, Hipc, npI, JI, JJ, B1, g24, Hi, Hj, Hij);
}

/* JTg and JTH12J of pt_ipc.py:170-172,237-246 (ee_ipc.py:170-190): the 12-entry
 * vertex-space gradient g12 and Hessian Hipc to the two bodies' affine coordinates. */
static void map_pair(const double g12[12], const double Hipc[144], int npI, const double* JI, const double* JJ,
                     double B1, double g24[24], double Hi[144], double Hj[144], double Hij[144]) {
    /* JTg: [kron(JI^T, I3) @ g12[:3npI]; kron(JJ^T, I3) @ g12[3npI:]] * B' */
    int nJ = 4 - npI;
    double JIT[16], JJT[16], KI[144], KJ[144];
    transpose(JI, npI, 4, JIT);
    transpose(JJ, nJ, 4, JJT);
    kron3(JIT, 4, npI, KI);  /* 12 x 3npI */
    kron3(JJT, 4, nJ, KJ);   /* 12 x 3nJ */
    double gi[12], gj[12];
    matmul(KI, g12, gi, 12, 3 * npI, 1);
    matmul(KJ, g12 + 3 * npI, gj, 12, 3 * nJ, 1);
    for (int i = 0; i < 12; ++i) { g24[i] = gi[i] * B1; g24[12 + i] = gj[i] * B1; }
    /* JTH12J */
    double KIm[144], KJm[144];  /* kron(JI, I3): 3npI x 12; kron(JJ, I3): 3nJ x 12 */
    kron3(JI, npI, 4, KIm);
    kron3(JJ, nJ, 4, KJm);
    double Hpp[144], Hpt[144], Htt[144];
    int a = 3 * npI, c = 3 * nJ;
    for (int i = 0; i < a; ++i)
        for (int j = 0; j < a; ++j) Hpp[i * a + j] = Hipc[i * 12 + j];
    for (int i = 0; i < a; ++i)
        for (int j = 0; j < c; ++j) Hpt[i * c + j] = Hipc[i * 12 + a + j];
    for (int i = 0; i < c; ++i)
        for (int j = 0; j < c; ++j) Htt[i * c + j] = Hipc[(a + i) * 12 + a + j];
    double KImT[144], KJmT[144], t1[144];
    transpose(KIm, a, 12, KImT);
    transpose(KJm, c, 12, KJmT);
    matmul(KImT, Hpp, t1, 12, a, a);
    matmul(t1, KIm, Hi, 12, a, 12);
    matmul(KJmT, Htt, t1, 12, c, c);
    matmul(t1, KJm, Hj, 12, c, 12);
    matmul(KImT, Hpt, t1, 12, a, c);
    matmul(t1, KJm, Hij, 12, c, 12);
}

/* ipcsk/put.py:26-66 _put_hess / _put_grad with D5 fixed (blocks sized
 * 16*(Nb+2Nij)).  blocks: [slots*16][3][3], block (ii, jj) of a slot at
 * 16*slot + ii + 4*jj. */
static void put_pair(int nb, long nij, const int32_t* t5, const double g24[24], const double Hi[144],
                     const double Hj[144], const double Hij[144], double* g, double* blocks) {
    int I = t5[0], J = t5[1];
    long idx = t5[2];
    for (int a = 0; a < 4; ++a)
        for (int k = 0; k < 3; ++k) {
            g[3 * (4 * I + a) + k] += g24[3 * a + k];
            g[3 * (4 * J + a) + k] += g24[12 + 3 * a + k];
        }
    for (int ii = 0; ii < 4; ++ii)
        for (int jj = 0; jj < 4; ++jj)
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) {
                    double hi = Hi[(3 * ii + r) * 12 + 3 * jj + c];
                    double hj = Hj[(3 * ii + r) * 12 + 3 * jj + c];
                    double hij = Hij[(3 * ii + r) * 12 + 3 * jj + c];
                    blocks[9 * (16 * (long)I + ii + jj * 4) + 3 * r + c] += hi;
                    blocks[9 * (16 * (long)J + ii + jj * 4) + 3 * r + c] += hj;
                    if (I < J) {
                        blocks[9 * (16 * (nb + idx) + ii + jj * 4) + 3 * r + c] += hij;
                        blocks[9 * (16 * (nb + idx + nij) + jj + ii * 4) + 3 * c + r] += hij;
                    } else {
                        blocks[9 * (16 * (nb + idx) + jj + ii * 4) + 3 * c + r] += hij;
                        blocks[9 * (16 * (nb + idx + nij) + ii + jj * 4) + 3 * r + c] += hij;
                    }
                }
}

/* One contact's reference assembly.  Returns 1 when active.  kind 0 PT, 1 EE. */
int orc_pair_gh(int kind, const orc_params* P, const v3 v[4], const v3 r[4], double g24[24], double Hi[144],
                double Hj[144], double Hij[144], double* s_out) {
    if (P->psd == 3) {
        /* full IPC: closest feature pair, mollified EE, projected 12x12 (abd_oracle_full.c) */
        double x12[12], r12[12], E, g12[12], H12[144];
        for (int t = 0; t < 4; ++t) { st3(x12 + 3 * t, v[t]); st3(r12 + 3 * t, r[t]); }
        if (s_out) *s_out = 0.0;
        if (!orc_full_local(kind, x12, r12, P->d2hat, P->kappa, 1, &E, g12, H12, NULL)) return 0;
        double JI[8], JJ[12];
        int npI = kind == 0 ? 1 : 2;
        for (int i = 0; i < npI; ++i) { JI[4 * i] = 1.0; JI[4 * i + 1] = r[i].x; JI[4 * i + 2] = r[i].y; JI[4 * i + 3] = r[i].z; }
        for (int i = 0; i < 4 - npI; ++i) {
            v3 q = r[npI + i];
            JJ[4 * i] = 1.0; JJ[4 * i + 1] = q.x; JJ[4 * i + 2] = q.y; JJ[4 * i + 3] = q.z;
        }
        map_pair(g12, H12, npI, JI, JJ, 1.0, g24, Hi, Hj, Hij);
        return 1;
    }
    v3 e0p, e1p, e2p;
    int valid;
    double s, Q[45], lam[5], D[12];
    if (kind == 0) {
        /* pt_ipc.py:14-22 _mask_valid: d = signed_distance^2 */
        valid = valid_pt(v[0], v[1], v[2], v[3]);
        C_vf(v[0], v[1], v[2], v[3], &e0p, &e1p, &e2p);
        double d0 = signed_distance(e0p, e1p, e2p);
        s = d0 * d0;
    } else {
        /* ee_ipc.py:14-23 _mask_valid: d = |e2p|^2 */
        valid = valid_ee(v[0], v[1], v[2], v[3]);
        C_ee(v[0], v[1], v[2], v[3], &e0p, &e1p, &e2p);
        s = dot(e2p, e2p);
    }
    if (s_out) *s_out = s;
    if (!(s < P->d2hat && valid)) return 0;
    double B1 = orc_barrier_d1(s, P->d2hat, P->kappa);
    double B2 = orc_barrier_d2(s, P->d2hat, P->kappa);
    /* pt_ipc.py:249-255 extract_JpJt / ee_ipc.py:201-211 extract_JeiJej */
    double JI[8], JJ[12];
    int npI = kind == 0 ? 1 : 2;
    for (int i = 0; i < npI; ++i) { JI[4 * i] = 1.0; JI[4 * i + 1] = r[i].x; JI[4 * i + 2] = r[i].y; JI[4 * i + 3] = r[i].z; }
    for (int i = 0; i < 4 - npI; ++i) {
        v3 q = r[npI + i];
        JJ[4 * i] = 1.0; JJ[4 * i + 1] = q.x; JJ[4 * i + 2] = q.y; JJ[4 * i + 3] = q.z;
    }
    if (P->psd == 2) {
        /* exact: b' d2s/dx2 + b'' ds/dx ds/dx^T, with (u, v, w) = L x (rows over the
         * contact's four vertices) as in C_vf (x1-x2, x3-x2, x0-x2) / C_ee (x1-x0, x3-x2, x2-x0) */
        static const double Lpt[12] = {0, 1, -1, 0, 0, 0, -1, 1, 1, 0, -1, 0};
        static const double Lee[12] = {-1, 1, 0, 0, 0, 0, -1, 1, -1, 0, 1, 0};
        const double* L = kind == 0 ? Lpt : Lee;
        v3 e[3];
        for (int a = 0; a < 3; ++a) {
            e[a] = mk(0, 0, 0);
            for (int t = 0; t < 4; ++t) e[a] = add(e[a], scl(v[t], L[4 * a + t]));
        }
        double ge[9], He[81], K[108], KT[108], tmp[108], Hx[144], g12[12], Hipc[144];
        exact_dist2(e[0], e[1], e[2], ge, He);
        kron3(L, 3, 4, K);  /* 9 x 12 */
        transpose(K, 9, 12, KT);
        matmul(KT, ge, g12, 12, 9, 1);
        matmul(KT, He, tmp, 12, 9, 9);
        matmul(tmp, K, Hx, 12, 9, 12);
        for (int i = 0; i < 12; ++i)
            for (int j = 0; j < 12; ++j) Hipc[i * 12 + j] = Hx[i * 12 + j] * B1 + (g12[i] * g12[j]) * B2;
        map_pair(g12, Hipc, npI, JI, JJ, B1, g24, Hi, Hj, Hij);
        return 1;
    }
    /* psd mode projects the whole contact-space Hessian, so it starts from the
     * unclamped eigenvalues */
    Q_lambda(e0p, e1p, e2p, kind == 0 && !P->psd, Q, lam);
    if (kind == 0) dcdx_pt(v[0], v[1], v[2], v[3], D); else dcdx_ee(v[0], v[1], v[2], v[3], D);
    pair_gh_dense(Q, lam, D, npI, JI, JJ, B1, B2, P->psd, g24, Hi, Hj, Hij);
    return 1;
}

/* ipc.py:54-62 IPCContactEnergy.gh: ipc_term_pt (pt_ipc.py:24-190), ipc_term_ee
 * (ee_ipc.py:50-143), ipc_term_vg (vg_ipc.py:5-34).  Accumulates into g and
 * blocks; highlight[b] |= 1 for PT-active on tuple[1], |= 2 for EE-active. */
void orc_contact_gh(const orc_scene* S, const orc_params* P, const double* x, const int32_t* pt, long npt,
                    const int32_t* ee, long nee, const int32_t* vg, long nvg, long nij, double* g, double* blocks,
                    uint8_t* highlight) {
    double g24[24], Hi[144], Hj[144], Hij[144];
    if (P->pt)
        for (long i = 0; i < npt; ++i) {
            v3 v[4], r[4];
            fetch_pt(S, x, pt + 5 * i, v);
            fetch_pt(S, S->x0, pt + 5 * i, r);
            if (orc_pair_gh(0, P, v, r, g24, Hi, Hj, Hij, NULL)) {
                put_pair(S->nb, nij, pt + 5 * i, g24, Hi, Hj, Hij, g, blocks);
                if (highlight) highlight[pt[5 * i + 1]] |= 1;
            }
        }
    if (P->ee)
        for (long i = 0; i < nee; ++i) {
            v3 v[4], r[4];
            fetch_ee(S, x, ee + 5 * i, v);
            fetch_ee(S, S->x0, ee + 5 * i, r);
            if (orc_pair_gh(1, P, v, r, g24, Hi, Hj, Hij, NULL)) {
                put_pair(S->nb, nij, ee + 5 * i, g24, Hi, Hj, Hij, g, blocks);
                if (highlight) highlight[ee[5 * i + 1]] |= 2;
            }
        }
    if (P->vg)
        for (long i = 0; i < nvg; ++i) {
            int b = vg[2 * i];
            long vid = S->voff[b] + vg[2 * i + 1];
            v3 v = ld3(x + 3 * vid), vt = ld3(S->x0 + 3 * vid);
            v3 nd = mk(0.0, 1.0, 0.0);
            double d = v.y - P->ground;
            double d2b = orc_barrier_d2(d * d, P->d2hat, P->kappa);
            double db = orc_barrier_d1(d * d, P->d2hat, P->kappa);
            for (int ii = 0; ii < 4; ++ii) {
                double th_i = ii == 0 ? 1.0 : comp(vt, ii - 1);  /* R2 */
                v3 gg = scl(scl(scl(nd, d * 2.0), th_i), db);
                g[3 * (4 * b + ii) + 0] += gg.x;
                g[3 * (4 * b + ii) + 1] += gg.y;
                g[3 * (4 * b + ii) + 2] += gg.z;
                for (int jj = 0; jj < 4; ++jj) {
                    double th_j = jj == 0 ? 1.0 : comp(vt, jj - 1);
                    double sc = th_i * th_j * d2b * 4.0 * d * d;
                    if (P->psd == 2) sc += th_i * th_j * 2.0 * db;  /* exact: the b' term of D11 */
                    if (P->psd == 3) {  /* full IPC: 4 d^2 b'' + 2 b' clamped at 0 */
                        double cy = d2b * 4.0 * d * d + 2.0 * db;
                        sc = th_i * th_j * (cy > 0.0 ? cy : 0.0);
                    }
                    double* B = blocks + 9 * (16 * (long)b + ii + jj * 4);
                    for (int rr = 0; rr < 3; ++rr)
                        for (int cc = 0; cc < 3; ++cc) B[3 * rr + cc] += (comp(nd, rr) * comp(nd, cc)) * sc;
                }
            }
        }
}

/* --------------------------------------------------------------- inertia */
/* orthogonal_energy.py:129-135 tildeA / tildep */
static void tilde(const orc_params* P, const double* A0, const double* Adot, const double* p0, const double* pdot,
                  double At[9], double pt[3]) {
    for (int i = 0; i < 9; ++i) At[i] = A0[i] + P->dt * Adot[i];
    double gv[3] = {0.0, P->gravity, 0.0};
    for (int k = 0; k < 3; ++k) pt[k] = p0[k] + P->dt * pdot[k] + P->dt * P->dt * gv[k];
}
static double row_dot(const double* A, int i, int j) {
    return A[3 * i] * A[3 * j] + A[3 * i + 1] * A[3 * j + 1] + A[3 * i + 2] * A[3 * j + 2];
}
/* orthogonal_energy.py:25-51 energy_inertia / norm_M / energy_ortho, at (Ak, pk).
 * mass/I0 per body (D13); NULL -> P->mass / P->I0. */
double orc_inertia_energy(const orc_params* P, int nb, const double* Ak, const double* pk, const double* A0,
                          const double* p0, const double* Adot, const double* pdot, const double* mass,
                          const double* I0) {
    double E = 0.0;
    for (int b = 0; b < nb; ++b) {
        double At[9], pt[3];
        tilde(P, A0 + 9 * b, Adot + 9 * b, p0 + 3 * b, pdot + 3 * b, At, pt);
        const double* A = Ak + 9 * b;
        double m = mass ? mass[b] : P->mass, i0 = I0 ? I0[b] : P->I0;
        double dq0[3], dq[9];
        for (int k = 0; k < 3; ++k) dq0[k] = pk[3 * b + k] - pt[k];
        for (int k = 0; k < 9; ++k) dq[k] = A[k] - At[k];
        double n0 = dq0[0] * dq0[0] + dq0[1] * dq0[1] + dq0[2] * dq0[2];
        double n1 = dq[0] * dq[0] + dq[1] * dq[1] + dq[2] * dq[2];
        double n2 = dq[3] * dq[3] + dq[4] * dq[4] + dq[5] * dq[5];
        double n3 = dq[6] * dq[6] + dq[7] * dq[7] + dq[8] * dq[8];
        double dqTMdq = n0 * m + (n1 + n2 + n3) * i0;
        double eo = 0.0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double term = row_dot(A, i, j) - (i == j ? 1.0 : 0.0);
                eo += term * term;
            }
        eo = eo * P->stiffness;
        E += eo * P->dt * P->dt + 0.5 * dqTMdq;
    }
    return E;
}
/* orthogonal_energy.py:99-112 flattened_gradient_inertia with :55-61
 * grad_ortho and :121-127 Mdq, at (A, p).  Overwrites g (gradient() zeroes
 * first, :17-19). */
void orc_inertia_gradient(const orc_params* P, int nb, const double* A, const double* p, const double* A0,
                          const double* p0, const double* Adot, const double* pdot, const double* mass,
                          const double* I0, double* g) {
    for (int b = 0; b < nb; ++b) {
        const double* Ab = A + 9 * b;
        double At[9], pt[3];
        tilde(P, A0 + 9 * b, Adot + 9 * b, p0 + 3 * b, pdot + 3 * b, At, pt);
        double m = mass ? mass[b] : P->mass, i0 = I0 ? I0[b] : P->I0;
        for (int k = 0; k < 3; ++k) g[3 * (4 * b) + k] = 0.0;
        for (int ii = 1; ii < 4; ++ii) {
            int i = ii - 1;
            double gr[3] = {-Ab[3 * i], -Ab[3 * i + 1], -Ab[3 * i + 2]};
            for (int j = 0; j < 3; ++j) {
                double d = row_dot(Ab, i, j);
                for (int k = 0; k < 3; ++k) gr[k] += d * Ab[3 * j + k];
            }
            for (int k = 0; k < 3; ++k) g[3 * (4 * b + ii) + k] = P->dt * P->dt * (gr[k] * 4.0 * P->stiffness);
        }
        for (int k = 0; k < 3; ++k) g[3 * (4 * b) + k] += (p[3 * b + k] - pt[k]) * m;
        for (int ii = 1; ii < 4; ++ii)
            for (int k = 0; k < 3; ++k) g[3 * (4 * b + ii) + k] += (Ab[3 * (ii - 1) + k] - At[3 * (ii - 1) + k]) * i0;
    }
}
/* orthogonal_energy.py:84-96 bsr_hessian_inertia with :63-76 hessian_ortho (+=). */
void orc_inertia_hessian(const orc_params* P, int nb, const double* A, const double* mass, const double* I0,
                         double* blocks) {
    for (int b = 0; b < nb; ++b) {
        const double* Ab = A + 9 * b;
        double m_ = mass ? mass[b] : P->mass, i0 = I0 ? I0[b] : P->I0;
        for (int ii = 0; ii < 4; ++ii)
            for (int jj = 0; jj < 4; ++jj) {
                double m = ii == jj ? (ii == 0 ? m_ : i0) : 0.0;  /* R2 at :91 */
                double dh[9] = {m, 0, 0, 0, m, 0, 0, 0, m};
                if (ii > 0 && jj > 0) {
                    int i = ii - 1, j = jj - 1;
                    double h[9] = {0};
                    if (i == j) {
                        double qq = row_dot(Ab, i, i) - 1.0;
                        for (int k = 0; k < 3; ++k)
                            for (int r = 0; r < 3; ++r)
                                for (int c = 0; c < 3; ++c) h[3 * r + c] += Ab[3 * k + r] * Ab[3 * k + c];
                        for (int r = 0; r < 3; ++r)
                            for (int c = 0; c < 3; ++c) h[3 * r + c] += Ab[3 * i + r] * Ab[3 * i + c] + (r == c ? qq : 0.0);
                    } else {
                        double d = row_dot(Ab, j, i);
                        for (int r = 0; r < 3; ++r)
                            for (int c = 0; c < 3; ++c) h[3 * r + c] = Ab[3 * j + r] * Ab[3 * i + c] + (r == c ? d : 0.0);
                    }
                    for (int k = 0; k < 9; ++k) dh[k] += h[k] * 4.0 * P->stiffness * P->dt * P->dt;
                }
                double* B = blocks + 9 * (16 * (long)b + ii + jj * 4);
                for (int k = 0; k < 9; ++k) B[k] += dh[k];
            }
    }
}

/* orthogonal_energy.py:138-165 _set_triplets, launched over all Nb + 2Nij slots (D5). */
void orc_set_triplets(int nb, long nij, const int32_t* ij, int32_t* rows, int32_t* cols) {
    for (long i = 0; i < nb + 2 * nij; ++i) {
        long I, J;
        if (i < nb) { I = J = i; }
        else if (i < nb + nij) { I = ij[2 * (i - nb)]; J = ij[2 * (i - nb) + 1]; }
        else { I = ij[2 * (i - nb - nij) + 1]; J = ij[2 * (i - nb - nij)]; }
        for (int ii = 0; ii < 4; ++ii)
            for (int jj = 0; jj < 4; ++jj) {
                rows[16 * i + ii + jj * 4] = (int32_t)(ii + 4 * I);
                cols[16 * i + ii + jj * 4] = (int32_t)(jj + 4 * J);
            }
    }
}

/* y = H x over 3x3 triplets (bsr_set_from_triplets sums duplicates; bsr_mv). n3 = 4Nb block rows. */
static void trip_mv(long nblk, const int32_t* rows, const int32_t* cols, const double* blocks, const double* x,
                    double* y, long n) {
    for (long i = 0; i < n; ++i) y[i] = 0.0;
    for (long t = 0; t < nblk; ++t) {
        const double* B = blocks + 9 * t;
        long r = rows[t], c = cols[t];
        for (int a = 0; a < 3; ++a) {
            double s = 0.0;
            for (int b = 0; b < 3; ++b) s += B[3 * a + b] * x[3 * c + b];
            y[3 * r + a] += s;
        }
    }
}
static double vdot(const double* a, const double* b, long n) {
    double s = 0.0;
    for (long i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* Unpreconditioned BiCGSTAB (van der Vorst 1992) in the iteration form of
 * warp.optim.linear.bicgstab as called at abd.py:244 (x warm-started, relative
 * tolerance tol on |r| against |b|, residual checked every check_every
 * iterations, maxiter 0 -> system size).  PARITY UNPINNED: Warp is external;
 * pinned by the residual bound and numpy.linalg.solve in the tests.
 * Returns iterations; *res = final |r|^2. */
long orc_bicgstab(long nslot, const int32_t* rows, const int32_t* cols, const double* blocks, const double* b,
                  double* x, long n, double tol, long maxiter, int check_every, double* res) {
    long nblk = 16 * nslot;
    if (maxiter <= 0) maxiter = n;
    double* r = calloc(n, 8); double* r0 = calloc(n, 8); double* p = calloc(n, 8);
    double* v = calloc(n, 8); double* t = calloc(n, 8);
    trip_mv(nblk, rows, cols, blocks, x, t, n);
    for (long i = 0; i < n; ++i) r[i] = b[i] - t[i];
    double rr = vdot(r, r, n);
    double atol = tol * sqrt(vdot(b, b, n));
    double atol_sq = atol * atol;
    long it = 0;
    if (rr > atol_sq) {
        memcpy(r0, r, n * 8); memcpy(p, r, n * 8);
        double rho = rr;
        for (;;) {
            for (int c = 0; c < check_every; ++c) {
                trip_mv(nblk, rows, cols, blocks, p, v, n);
                double r0v = vdot(r0, v, n);
                double alpha = r0v != 0.0 ? rho / r0v : 0.0;
                for (long i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * v[i]; }
                trip_mv(nblk, rows, cols, blocks, r, t, n);
                double tr = vdot(t, r, n), tt = vdot(t, t, n);
                double omega = tt != 0.0 ? tr / tt : 0.0;
                for (long i = 0; i < n; ++i) { x[i] += omega * r[i]; r[i] -= omega * t[i]; }
                rr = vdot(r, r, n);
                double rho_old = rho;
                rho = vdot(r0, r, n);
                double beta = (rho_old != 0.0 && omega != 0.0) ? (rho / rho_old) * (alpha / omega) : 0.0;
                for (long i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
            }
            it += check_every;
            if (rr <= atol_sq || it >= maxiter) break;
        }
    }
    if (res) *res = rr;
    free(r); free(r0); free(p); free(v); free(t);
    return it;
}

/* bsr_utils.py:126-235 bsr_cg with the block-diagonal Jacobi preconditioner
 * (:95-108: z = r / |diag of the 3x3 diagonal block|) and tol_sq = tol^2 n
 * (:195) on r.z, checked every check_every iterations.  Returns iterations. */
long orc_bsr_cg(long nslot, const int32_t* rows, const int32_t* cols, const double* blocks, const double* b,
                double* x, long n, double tol, long maxiter, int check_every, double* res) {
    long nblk = 16 * nslot;
    if (maxiter <= 0) maxiter = n;
    double* r = calloc(n, 8); double* p = calloc(n, 8); double* Ap = calloc(n, 8);
    double* z = calloc(n, 8); double* dg = calloc(n, 8);
    /* bsr_get_diag: diagonal 3x3 blocks (row == col), duplicates summed */
    for (long t = 0; t < nblk; ++t)
        if (rows[t] == cols[t])
            for (int a = 0; a < 3; ++a) dg[3 * rows[t] + a] += blocks[9 * t + 4 * a];
    trip_mv(nblk, rows, cols, blocks, x, Ap, n);
    for (long i = 0; i < n; ++i) r[i] = b[i] - Ap[i];
#define PRECOND() do { for (long i3 = 0; i3 < n / 3; ++i3) { \
        double d0 = dg[3*i3], d1 = dg[3*i3+1], d2 = dg[3*i3+2]; \
        if (d0 * d0 + d1 * d1 + d2 * d2 == 0.0) { for (int a = 0; a < 3; ++a) z[3*i3+a] = r[3*i3+a]; } \
        else { z[3*i3] = r[3*i3] / fabs(d0); z[3*i3+1] = r[3*i3+1] / fabs(d1); z[3*i3+2] = r[3*i3+2] / fabs(d2); } } } while (0)
    PRECOND();
    memcpy(p, z, n * 8);
    double rz_old = vdot(r, z, n), rz_new = 0.0;
    double tol_sq = tol * tol * (double)n;
    double err = rz_old;
    long end_iter = 0;
    if (err > tol_sq) {
        end_iter = maxiter;
        for (long i = 0; i < maxiter; ++i) {
            trip_mv(nblk, rows, cols, blocks, p, Ap, n);
            double pAp = vdot(p, Ap, n);
            /* the reference divides unguarded: a residual that vanishes exactly between
             * two checks gives p = 0 and 0/0 here; the step is skipped instead */
            double alpha = pAp != 0.0 ? rz_old / pAp : 0.0;
            for (long k = 0; k < n; ++k) { x[k] += alpha * p[k]; r[k] -= alpha * Ap[k]; }
            PRECOND();
            rz_new = vdot(r, z, n);
            if (((i + 1) % check_every) == 0 && rz_new <= tol_sq) { end_iter = i; err = rz_new; break; }
            double beta = rz_new > tol_sq ? rz_new / rz_old : 0.0;
            for (long k = 0; k < n; ++k) p[k] = z[k] + beta * p[k];
            double tmp = rz_old; rz_old = rz_new; rz_new = tmp;
            err = rz_old;
        }
    }
#undef PRECOND
    if (res) *res = err;
    free(r); free(p); free(Ap); free(z); free(dg);
    return end_iter;
}

/* Beyond the reference: bsr_cg with the 12x12 diagonal block of each body as the
 * preconditioner (z = D_b^-1 r per body) instead of the diagonal of the 3x3 blocks;
 * same iteration, stopping rule and D16 guard.  D_b^-1 by Gauss-Jordan elimination with
 * partial pivoting (a singular block leaves that body unpreconditioned, z = r). */
static int invert12(const double* D, double* Dinv) {
    double a[12][24];
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 24; ++j) a[i][j] = j < 12 ? D[12 * i + j] : (j - 12 == i ? 1.0 : 0.0);
    for (int c = 0; c < 12; ++c) {
        int piv = c;
        for (int r = c + 1; r < 12; ++r)
            if (fabs(a[r][c]) > fabs(a[piv][c])) piv = r;
        if (a[piv][c] == 0.0) return 0;
        if (piv != c)
            for (int j = 0; j < 24; ++j) { double t = a[c][j]; a[c][j] = a[piv][j]; a[piv][j] = t; }
        double inv = 1.0 / a[c][c];
        for (int j = 0; j < 24; ++j) a[c][j] *= inv;
        for (int r = 0; r < 12; ++r) {
            if (r == c) continue;
            double f = a[r][c];
            if (f != 0.0)
                for (int j = 0; j < 24; ++j) a[r][j] -= f * a[c][j];
        }
    }
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 12; ++j) Dinv[12 * i + j] = a[i][12 + j];
    return 1;
}

long orc_bsr_cg_block(long nslot, const int32_t* rows, const int32_t* cols, const double* blocks, const double* b,
                      double* x, long n, double tol, long maxiter, int check_every, double* res) {
    long nblk = 16 * nslot, nb = n / 12;
    if (maxiter <= 0) maxiter = n;
    double* r = calloc(n, 8); double* p = calloc(n, 8); double* Ap = calloc(n, 8); double* z = calloc(n, 8);
    double* Dinv = calloc(144 * nb, 8);
    int* ok = calloc(nb, sizeof(int));
    /* diagonal 12x12 block of body b: triplets with row == col block and both in body b */
    double* D = calloc(144 * nb, 8);
    for (long t = 0; t < nblk; ++t) {
        int br = rows[t] / 4, bc = cols[t] / 4;
        if (br != bc) continue;
        int ii = rows[t] % 4, jj = cols[t] % 4;
        for (int u = 0; u < 3; ++u)
            for (int v = 0; v < 3; ++v) D[144 * br + 12 * (3 * ii + u) + 3 * jj + v] += blocks[9 * t + 3 * u + v];
    }
    for (long q = 0; q < nb; ++q) ok[q] = invert12(D + 144 * q, Dinv + 144 * q);
#define PRECOND_B() do { for (long q = 0; q < nb; ++q) for (int l = 0; l < 12; ++l) { \
        if (!ok[q]) { z[12 * q + l] = r[12 * q + l]; continue; } \
        double acc = 0.0; for (int k = 0; k < 12; ++k) acc += Dinv[144 * q + 12 * l + k] * r[12 * q + k]; \
        z[12 * q + l] = acc; } } while (0)
    trip_mv(nblk, rows, cols, blocks, x, Ap, n);
    for (long i = 0; i < n; ++i) r[i] = b[i] - Ap[i];
    PRECOND_B();
    memcpy(p, z, n * 8);
    double rz_old = vdot(r, z, n), rz_new = 0.0;
    double tol_sq = tol * tol * (double)n;
    double err = rz_old;
    long end_iter = 0;
    if (err > tol_sq) {
        end_iter = maxiter;
        for (long i = 0; i < maxiter; ++i) {
            trip_mv(nblk, rows, cols, blocks, p, Ap, n);
            double pAp = vdot(p, Ap, n);
            double alpha = pAp != 0.0 ? rz_old / pAp : 0.0;
            for (long k = 0; k < n; ++k) { x[k] += alpha * p[k]; r[k] -= alpha * Ap[k]; }
            PRECOND_B();
            rz_new = vdot(r, z, n);
            if (((i + 1) % check_every) == 0 && rz_new <= tol_sq) { end_iter = i; err = rz_new; break; }
            double beta = rz_new > tol_sq ? rz_new / rz_old : 0.0;
            for (long k = 0; k < n; ++k) p[k] = z[k] + beta * p[k];
            double tmp = rz_old; rz_old = rz_new; rz_new = tmp;
            err = rz_old;
        }
    }
#undef PRECOND_B
    if (res) *res = err;
    free(r); free(p); free(Ap); free(z); free(Dinv); free(ok); free(D);
    return end_iter;
}

/* Beyond the reference: the Newton system solved directly, by a block LDL^T factorization
 * of the 12x12 body blocks (no exchanges between bodies), fill-in kept.  The structural
 * pattern is the slots' (every triplet's block), so it matches the BSR matrix's; the fill
 * is that of eliminating the bodies in their own order.  The bodies are then eliminated
 * level by level of that order's elimination tree (parent of k: the first body above k
 * coupled to it; level: 0 for a leaf, else one more than the highest child's), in
 * ascending order within a level -- an order in which every body comes after all bodies
 * that update it.  Step k: Gauss-Jordan elimination without row exchanges on
 * [P | A'_k,j1 .. A'_k,jd | I], P = A'_kk (column by column: s = a[c][t] / p, a[r][t] -=
 * a[r][c] s for r != c, a[c][t] = s), leaving W_j = P^-1 A'_kj and P^-1; then
 * A'_ij -= A'_ki^T W_j for i <= j among j1..jd (each entry a 12-term sum in row order);
 * the forward solve's step z_k = P^-1 y_k, y_i -= W_i^T y_k (12-term sums); row k becomes
 * W.  The backward solve, levels from the top: x_k = z_k - sum_i W_i x_i with each W_i x_i
 * a 12-term sum, added in ascending i.  Systems above ORC_LDLT_MAX_BODIES bodies, or a
 * pivot that is not positive, return -2 with x untouched (the caller falls back to an
 * iterative solve). */
#define ORC_LDLT_MAX_BODIES 64
long orc_bsr_ldlt(long nslot, const int32_t* rows, const int32_t* cols, const double* blocks, const double* b,
                  double* x, long n) {
    const long m = n / 12, nblk = 16 * nslot;
    if (m > ORC_LDLT_MAX_BODIES) return -2;
    if (m == 0) return 1;
    double* U = calloc((size_t)m * m * 144, 8); /* dense block storage, upper part used */
    unsigned char* nz = calloc((size_t)m * m, 1);
    for (long i = 0; i < m; ++i) nz[i * m + i] = 1;
    for (long t = 0; t < nblk; ++t) {
        int I = rows[t] / 4, J = cols[t] / 4, ii = rows[t] % 4, jj = cols[t] % 4;
        if (I > J) { int q = I; I = J; J = q; }
        nz[I * m + J] = 1;
        if (rows[t] / 4 > cols[t] / 4) continue; /* values from the upper slots */
        for (int u = 0; u < 3; ++u)
            for (int v = 0; v < 3; ++v) U[((size_t)I * m + J) * 144 + 12 * (3 * ii + u) + 3 * jj + v] += blocks[9 * t + 3 * u + v];
    }
    /* fill: eliminating k couples every pair of its later neighbours */
    for (long k = 0; k < m; ++k)
        for (long i = k + 1; i < m; ++i)
            if (nz[k * m + i])
                for (long j = i; j < m; ++j)
                    if (nz[k * m + j]) nz[i * m + j] = 1;
    /* elimination tree levels and the elimination order (level, body) */
    int* lev = calloc((size_t)m, sizeof(int));
    int maxlev = 0;
    for (long k = 0; k < m; ++k) {
        for (long j = k + 1; j < m; ++j)
            if (nz[k * m + j]) {
                if (lev[j] < lev[k] + 1) lev[j] = lev[k] + 1;
                break;
            }
        if (lev[k] > maxlev) maxlev = lev[k];
    }
    long* order = malloc(sizeof(long) * (size_t)m);
    long no = 0;
    for (int L = 0; L <= maxlev; ++L)
        for (long k = 0; k < m; ++k)
            if (lev[k] == L) order[no++] = k;
    double* y = malloc(8 * (size_t)n);
    double* z = malloc(8 * (size_t)n);
    memcpy(y, b, 8 * (size_t)n);
    double* W = malloc(8 * 144 * (size_t)m);
    long* mem = malloc(sizeof(long) * (size_t)m);
    double* aug = malloc(8 * 12 * (size_t)(12 * (m + 2)));
    int ok = 1;
    for (long o = 0; o < m && ok; ++o) {
        const long k = order[o];
        long d = 0;
        for (long j = k + 1; j < m; ++j)
            if (nz[k * m + j]) mem[d++] = j;
        const long nc = 12 * (d + 2);
        /* aug[r * nc + t]: P | row-k blocks | I */
        for (int r = 0; r < 12; ++r)
            for (long t = 0; t < nc; ++t) {
                const long g = t / 12, c = t % 12;
                double v;
                if (g == 0) v = U[((size_t)k * m + k) * 144 + 12 * r + c];
                else if (g <= d) v = U[((size_t)k * m + mem[g - 1]) * 144 + 12 * r + c];
                else v = r == c ? 1.0 : 0.0;
                aug[r * nc + t] = v;
            }
        for (int c = 0; c < 12 && ok; ++c) {
            double f[12];
            for (int r = 0; r < 12; ++r) f[r] = aug[r * nc + c];
            const double piv = f[c];
            if (!(piv > 0.0) || !isfinite(piv)) { ok = 0; break; }
            const double inv = 1.0 / piv;
            for (long t = 0; t < nc; ++t) {
                const double sc = aug[c * nc + t] * inv;
                for (int r = 0; r < 12; ++r)
                    if (r != c) aug[r * nc + t] -= f[r] * sc;
                aug[c * nc + t] = sc;
            }
        }
        if (!ok) break;
        for (long g = 0; g < d; ++g)
            for (int r = 0; r < 12; ++r)
                for (int c = 0; c < 12; ++c) W[144 * g + 12 * r + c] = aug[r * nc + 12 * (g + 1) + c];
        double Dinv[144];
        for (int r = 0; r < 12; ++r)
            for (int c = 0; c < 12; ++c) Dinv[12 * r + c] = aug[r * nc + 12 * (d + 1) + c];
        /* Schur complement of the later rows */
        for (long u = 0; u < d; ++u)
            for (long v = u; v < d; ++v) {
                const double* Aki = U + ((size_t)k * m + mem[u]) * 144;
                double* Uij = U + ((size_t)mem[u] * m + mem[v]) * 144;
                for (int a = 0; a < 12; ++a)
                    for (int c = 0; c < 12; ++c) {
                        double acc = 0.0;
                        for (int r = 0; r < 12; ++r) acc += Aki[12 * r + a] * W[144 * v + 12 * r + c];
                        Uij[12 * a + c] -= acc;
                    }
            }
        /* forward solve, step k */
        for (int a = 0; a < 12; ++a) {
            double acc = 0.0;
            for (int r = 0; r < 12; ++r) acc += Dinv[12 * a + r] * y[12 * k + r];
            z[12 * k + a] = acc;
        }
        for (long u = 0; u < d; ++u)
            for (int a = 0; a < 12; ++a) {
                double acc = 0.0;
                for (int r = 0; r < 12; ++r) acc += W[144 * u + 12 * r + a] * y[12 * k + r];
                y[12 * mem[u] + a] -= acc;
            }
        /* row k of the factor: W */
        for (long u = 0; u < d; ++u) memcpy(U + ((size_t)k * m + mem[u]) * 144, W + 144 * u, 8 * 144);
    }
    if (ok) {
        double* xs = y; /* y_k is consumed once z_k exists */
        for (long o = m - 1; o >= 0; --o) {
            const long k = order[o];
            double acc[12] = {0};
            for (long i = k + 1; i < m; ++i) {
                if (!nz[k * m + i]) continue;
                const double* Wi = U + ((size_t)k * m + i) * 144;
                for (int a = 0; a < 12; ++a) {
                    double p = 0.0;
                    for (int c = 0; c < 12; ++c) p += Wi[12 * a + c] * xs[12 * i + c];
                    acc[a] += p;
                }
            }
            for (int a = 0; a < 12; ++a) xs[12 * k + a] = z[12 * k + a] - acc[a];
        }
        memcpy(x, xs, 8 * (size_t)n);
    }
    free(U); free(nz); free(y); free(z); free(W); free(mem); free(aug); free(lev); free(order);
    return ok ? 1 : -2;
}

/* ------------------------------------------------------------------- CCD */
#define X_ERROR 1e-10
/* Iteration bound the reference does not have; never reached on the test
 * corpus (the GPU path has the same bound so that every wave terminates). */
#define FIND_CLOSED_MAX_IT 1000

/* cubic_roots.py:6-19 */
static double evaluate(const double c[4], double x) { return c[0] + x * (c[1] + x * (c[2] + x * c[3])); }
static double multi_sign(double a, double b) { return b < 0.0 ? -a : a; }
static int diff_sign(double y0, double yr) { return y0 * yr < 0.0; }
/* cubic_roots.py:21-27 */
static void deflate(const double c[4], double root, double d[3]) {
    d[2] = c[3];
    for (int i = 2; i > 0; --i) d[i - 1] = c[i] + root * d[i];
}
/* cubic_roots.py:29-62 */
static double find_closed(const double c[4], const double dv[4], double x0, double x1, double y0, double y1) {
    (void)y1;
    double xr = (x0 + x1) / 2.0;
    double yr = evaluate(c, xr);
    double xb0 = x0, xb1 = x1;
    if (!(x1 - x0 <= X_ERROR * 2.0)) {
        for (int it = 0; it < FIND_CLOSED_MAX_IT; ++it) {
            int side = diff_sign(y0, yr);
            if (side) xb1 = xr; else xb0 = xr;
            double dy = evaluate(dv, xr);
            double dx = yr / dy;
            double xn = xr - dx;
            if (xn > xb0 && xn < xb1) {
                double stepsize = fabs(xr - xn);
                xr = xn;
                if (stepsize > X_ERROR) yr = evaluate(c, xr);
                else break;
            } else {
                xr = (xb0 + xb1) / 2.0;
                yr = evaluate(c, xr);
                if (xb0 == xr || xb1 == xr || xb1 - xb0 <= 2.0 * X_ERROR) break;
            }
        }
    }
    return xr;
}
/* wp.min / wp.max on doubles: a < b ? a : b (a NaN yields the second operand). */
static double wmin(double a, double b) { return a < b ? a : b; }
static double wmax(double a, double b) { return a > b ? a : b; }
/* cubic_roots.py:63-93 (coef constant-first) */
static int quadratic_roots(const double coef[3], double x0, double x1, double roots[2]) {
    roots[0] = roots[1] = 0.0;
    double c = coef[0], b = coef[1], a = coef[2];
    double delta = b * b - 4.0 * a * c;
    int ret = 0;
    if (delta > 0.0) {
        double d = sqrt(delta);
        double q = -(b + multi_sign(d, b)) * 0.5;
        double rv0 = q / a, rv1 = c / q;
        double xa = wmin(rv0, rv1), xb = wmax(rv0, rv1);
        if (xa >= x0 && xa <= x1) roots[ret++] = xa;
        if (xb >= x0 && xb <= x1) roots[ret++] = xb;
    } else if (delta < 0.0) {
        ret = 0;
    } else {
        double r0 = -0.5 * b / a;
        roots[0] = r0;
        if (r0 >= x0 && r0 <= x1) ret = 1;
    }
    return ret;
}
/* cubic_roots.py:95-193 cubic_roots(coef, x0, x1) */
int orc_cubic_roots(const double coef[4], double x0, double x1, double roots[3]) {
    roots[0] = roots[1] = roots[2] = 0.0;
    double y0 = evaluate(coef, x0), y1 = evaluate(coef, x1);
    double a = coef[3] * 3.0, b_2 = coef[2], c = coef[1];
    double deriv[4] = {c, 2.0 * b_2, a, 0.0};
    double delta_4 = b_2 * b_2 - a * c;
    int ret = 0;
    double qr[2], dp[3];
    if (delta_4 > 0.0) {
        double d_2 = sqrt(delta_4);
        double q = -(b_2 + multi_sign(d_2, b_2));
        double rv0 = q / a, rv1 = c / q;
        double xa = wmin(rv0, rv1), xb = wmax(rv0, rv1);
        if (diff_sign(y0, y1)) {
            if (xa >= x1 || xb <= x0 || (xa <= x0 && xb >= x1)) {
                roots[0] = find_closed(coef, deriv, x0, x1, y0, y1);
                return 1;
            }
        } else {
            if (xa >= x1 || xb <= x0 || (xa <= x0 && xb >= x1)) return 0;
        }
        if (xa > x0) {
            double ya = evaluate(coef, xa);
            if (diff_sign(y0, ya)) {
                roots[0] = find_closed(coef, deriv, x0, xa, y0, ya);
                if (diff_sign(ya, y1) || (xb < x1 && diff_sign(ya, evaluate(coef, xb)))) {
                    deflate(coef, roots[0], dp);
                    int r = quadratic_roots(dp, xa, x1, qr);
                    roots[1] = qr[0]; roots[2] = qr[1];
                    return r + 1;
                }
                return 1;
            }
            if (xb < x1) {
                double yb = evaluate(coef, xb);
                if (diff_sign(ya, yb)) {
                    roots[0] = find_closed(coef, deriv, xa, xb, ya, yb);
                    if (diff_sign(yb, y1)) {
                        deflate(coef, roots[0], dp);
                        int r = quadratic_roots(dp, xb, x1, qr);
                        roots[1] = qr[0]; roots[2] = qr[1];
                        return r + 1;
                    }
                    return 1;
                }
                if (diff_sign(yb, y1)) {
                    roots[0] = find_closed(coef, deriv, xb, x1, yb, y1);
                    return 1;
                }
            } else if (diff_sign(ya, y1)) {
                roots[0] = find_closed(coef, deriv, xa, x1, ya, y1);
                return 1;
            }
        } else {
            double yb = evaluate(coef, xb);
            if (diff_sign(y0, yb)) {
                roots[0] = find_closed(coef, deriv, x0, xb, y0, yb);
                if (diff_sign(yb, y1)) {
                    deflate(coef, roots[0], dp);
                    int r = quadratic_roots(dp, xb, x1, qr);
                    roots[1] = qr[0]; roots[2] = qr[1];
                    return r + 1;
                }
                return 1;
            }
            if (diff_sign(yb, y1)) {
                roots[0] = find_closed(coef, deriv, xb, x1, yb, y1);
                return 1;
            }
        }
    } else {
        if (diff_sign(y0, y1)) {
            roots[0] = find_closed(coef, deriv, x0, x1, y0, y1);
            return 1;
        }
        return 0;
    }
    (void)ret;
    return 0;
}

/* ccd.py:122-129 cubic_binomial: prod_i (b_i + t a_i), constant-first */
static void cubic_binomial(const double a[3], const double b[3], double out[4]) {
    out[0] = b[0] * b[1] * b[2];
    out[1] = a[0] * b[1] * b[2] + b[0] * b[1] * a[2] + b[0] * a[1] * b[2];
    out[2] = a[0] * a[1] * b[2] + a[0] * b[1] * a[2] + b[0] * a[1] * a[2];
    out[3] = a[0] * a[1] * a[2];
}
/* ccd.py:82-119 det_polynomial (matrices row-major m[3r+c]) */
static void det_polynomial(const double* a, const double* b, double out[4]) {
    double pos[4] = {0, 0, 0, 0}, neg[4] = {0, 0, 0, 0}, t[4];
    static const int idx[6][3][2] = {
        {{0, 0}, {1, 1}, {2, 2}}, {{0, 1}, {1, 2}, {2, 0}}, {{0, 2}, {1, 0}, {2, 1}},
        {{0, 0}, {1, 2}, {2, 1}}, {{0, 1}, {1, 0}, {2, 2}}, {{0, 2}, {1, 1}, {2, 0}}};
    for (int s = 0; s < 6; ++s) {
        double av[3], bv[3];
        for (int k = 0; k < 3; ++k) {
            av[k] = a[3 * idx[s][k][0] + idx[s][k][1]];
            bv[k] = b[3 * idx[s][k][0] + idx[s][k][1]];
        }
        cubic_binomial(av, bv, t);
        for (int k = 0; k < 4; ++k) {
            if (s < 3) pos[k] += t[k]; else neg[k] += t[k];
        }
    }
    for (int k = 0; k < 4; ++k) out[k] = pos[k] - neg[k];
}
/* R1: mat33(u, v, w) has columns u, v, w */
static void mat_cols(v3 u, v3 v, v3 w, double* m) {
    m[0] = u.x; m[1] = v.x; m[2] = w.x;
    m[3] = u.y; m[4] = v.y; m[5] = w.y;
    m[6] = u.z; m[7] = v.z; m[8] = w.z;
}
/* ccd.py:56-79 build_and_solve_4_points_coplanar; exposed for the KAT tests */
int orc_coplanar_poly(const double* p_t0 /*[4][3]*/, const double* p_t1, double coef[4]) {
    v3 a[4], b[4];
    for (int i = 0; i < 4; ++i) { a[i] = ld3(p_t1 + 3 * i); b[i] = ld3(p_t0 + 3 * i); }
    double A1[9], A2[9], A3[9], A4[9], B1[9], B2[9], B3[9], B4[9];
    mat_cols(a[1], a[2], a[3], A1); mat_cols(a[0], a[2], a[3], A2);
    mat_cols(a[0], a[1], a[3], A3); mat_cols(a[0], a[1], a[2], A4);
    mat_cols(b[1], b[2], b[3], B1); mat_cols(b[0], b[2], b[3], B2);
    mat_cols(b[0], b[1], b[3], B3); mat_cols(b[0], b[1], b[2], B4);
    for (int k = 0; k < 9; ++k) { A1[k] -= B1[k]; A2[k] -= B2[k]; A3[k] -= B3[k]; A4[k] -= B4[k]; }
    double d1[4], d2[4], d3[4], d4[4];
    det_polynomial(A1, B1, d1); det_polynomial(A2, B2, d2);
    det_polynomial(A3, B3, d3); det_polynomial(A4, B4, d4);
    for (int k = 0; k < 4; ++k) coef[k] = d1[k] - d2[k] + d3[k] - d4[k];
    return 0;
}
/* wp.lerp(a, b, t) = a*(1-t) + b*t (assumed Warp definition) */
static v3 lerp(v3 a, v3 b, double t) { return add(scl(a, 1.0 - t), scl(b, t)); }

/* ccd.py:133-156 pt_collision_time (kind 0) and :158-186 ee_collision_time (kind 1) */
double orc_collision_time(int kind, const v3 t0[4], const v3 t1[4]) {
    double P0[12], P1[12], coef[4], roots[3];
    for (int i = 0; i < 4; ++i) { st3(P0 + 3 * i, t0[i]); st3(P1 + 3 * i, t1[i]); }
    orc_coplanar_poly(P0, P1, coef);
    int n = orc_cubic_roots(coef, 0.0, 1.0, roots);
    double root = 1.0;
    int true_root = 0;
    for (int i = 0; i < n; ++i) {
        root = roots[i];
        v3 p[4];
        for (int k = 0; k < 4; ++k) p[k] = lerp(t0[k], t1[k], root);
        true_root = kind == 0 ? valid_pt(p[0], p[1], p[2], p[3]) : valid_ee(p[0], p[1], p[2], p[3]);
        if (true_root) break;
    }
    if (!true_root) root = 1.0;
    return root;
}
double orc_collision_time_flat(int kind, const double* p_t0, const double* p_t1) {
    v3 a[4], b[4];
    for (int i = 0; i < 4; ++i) { a[i] = ld3(p_t0 + 3 * i); b[i] = ld3(p_t1 + 3 * i); }
    return orc_collision_time(kind, a, b);
}
/* ccd.py:188-195 vg_collison_time */
static double vg_time(double y0, double y1, double ground) {
    double toi = 1.0;
    double d1 = y1 - ground;
    if (d1 < 0) { double d0 = y0 - ground; toi = d0 / (d0 - d1); }
    return toi;
}
/* ccd.py:9-40 toi_vg/toi_pt/toi_ee (D4: correct lists, in-repo PT/EE TOI):
 * minimum over all candidates of the TOIs below 1, starting at 1.  x = t0, xk = t1. */
double orc_toi(const orc_scene* S, const orc_params* P, const double* x, const double* xk, const int32_t* pt, long npt,
               const int32_t* ee, long nee, const int32_t* vg, long nvg) {
    double toi = 1.0;
    if (P->vg)
        for (long i = 0; i < nvg; ++i) {
            long vid = S->voff[vg[2 * i]] + vg[2 * i + 1];
            double t = vg_time(x[3 * vid + 1], xk[3 * vid + 1], P->ground);
            if (t < 1.0 && t < toi) toi = t;
        }
    if (P->ee)
        for (long i = 0; i < nee; ++i) {
            v3 a[4], b[4];
            fetch_ee(S, x, ee + 5 * i, a);
            fetch_ee(S, xk, ee + 5 * i, b);
            double t = orc_collision_time(1, a, b);
            if (t < 1.0 && t < toi) toi = t;
        }
    if (P->pt)
        for (long i = 0; i < npt; ++i) {
            v3 a[4], b[4];
            fetch_pt(S, x, pt + 5 * i, a);
            fetch_pt(S, xk, pt + 5 * i, b);
            double t = orc_collision_time(0, a, b);
            if (t < 1.0 && t < toi) toi = t;
        }
    return toi;
}

/* Pair-level entry points for tests (one contact, flat arrays). */
int orc_pair_gh_flat(int kind, const orc_params* P, const double* x4, const double* r4, double* g24, double* Hi,
                     double* Hj, double* Hij, double* s_out) {
    v3 v[4], r[4];
    for (int i = 0; i < 4; ++i) { v[i] = ld3(x4 + 3 * i); r[i] = ld3(r4 + 3 * i); }
    return orc_pair_gh(kind, P, v, r, g24, Hi, Hj, Hij, s_out);
}
/* Geometry of one contact: e0p, e1p, e2p (9), beta, gamma, l, valid, Q (45), lam (5), D (12). */
void orc_pair_geometry(int kind, const double* x4, double* c9, double* bgl, int* valid, double* Q, double* lam,
                       double* D) {
    v3 v[4];
    for (int i = 0; i < 4; ++i) v[i] = ld3(x4 + 3 * i);
    v3 e0p, e1p, e2p;
    double be, ga;
    if (kind == 0) {
        C_vf(v[0], v[1], v[2], v[3], &e0p, &e1p, &e2p);
        bg_pt(v[0], v[1], v[2], v[3], &be, &ga);
        *valid = valid_pt(v[0], v[1], v[2], v[3]);
        dcdx_pt(v[0], v[1], v[2], v[3], D);
    } else {
        C_ee(v[0], v[1], v[2], v[3], &e0p, &e1p, &e2p);
        bg_ee(v[0], v[1], v[2], v[3], &be, &ga);
        *valid = valid_ee(v[0], v[1], v[2], v[3]);
        dcdx_ee(v[0], v[1], v[2], v[3], D);
    }
    st3(c9, e0p); st3(c9 + 3, e1p); st3(c9 + 6, e2p);
    bgl[0] = be; bgl[1] = ga; bgl[2] = signed_distance(e0p, e1p, e2p);
    Q_lambda(e0p, e1p, e2p, kind == 0, Q, lam);
}

/* ------------------------------------------------------------- state ops */
/* abd.py:337-353 _update_q / _update_qk: q_out = q - alpha dq (row i of A is a_{i+1}) */
void orc_update_q(int nb, const double* A, const double* p, const double* dq, double alpha, double* Aout, double* pout) {
    for (int b = 0; b < nb; ++b) {
        for (int k = 0; k < 3; ++k) pout[3 * b + k] = p[3 * b + k] - dq[3 * (4 * b) + k] * alpha;
        for (int r = 0; r < 3; ++r)
            for (int k = 0; k < 3; ++k) Aout[9 * b + 3 * r + k] = A[9 * b + 3 * r + k] - dq[3 * (4 * b + 1 + r) + k] * alpha;
    }
}
/* abd.py:355-362 _update_q0qdot */
void orc_update_q0qdot(int nb, double dt, const double* A, const double* p, double* A0, double* p0, double* Adot,
                       double* pdot) {
    for (int b = 0; b < nb; ++b) {
        for (int k = 0; k < 3; ++k) { pdot[3 * b + k] = (p[3 * b + k] - p0[3 * b + k]) / dt; p0[3 * b + k] = p[3 * b + k]; }
        for (int k = 0; k < 9; ++k) { Adot[9 * b + k] = (A[9 * b + k] - A0[9 * b + k]) / dt; A0[9 * b + k] = A[9 * b + k]; }
    }
}
