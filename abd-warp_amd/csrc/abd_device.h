// abd_device.h — device-side math of the contact path (gfx950, FP64).
//
// Operation order follows the reference Warp functions cited on each helper so
// that validity predicates (strict inequalities on barycentric coordinates)
// and FP32 boxes come out bit-identical to a CPU evaluation of the same
// formulas; the library is compiled with -ffp-contract=off for that reason.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace abd {

struct d3 {
    double x, y, z;
};
__device__ __forceinline__ d3 mk(double x, double y, double z) { return d3{x, y, z}; }
__device__ __forceinline__ d3 operator+(d3 a, d3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ d3 operator-(d3 a, d3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ d3 operator*(d3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ d3 operator/(d3 a, double s) { return mk(a.x / s, a.y / s, a.z / s); }
// wp.dot: (a.x*b.x + a.y*b.y) + a.z*b.z
__device__ __forceinline__ double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ d3 cross(d3 a, d3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double comp(d3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
__device__ __forceinline__ d3 ld3(const double* p) { return mk(p[0], p[1], p[2]); }
__device__ __forceinline__ void st3(double* p, d3 a) {
    p[0] = a.x;
    p[1] = a.y;
    p[2] = a.z;
}

// ipcsk/barrier.py:4-15 barrier(s), s = squared distance
__device__ __forceinline__ double barrier(double d, double d2hat, double kappa) {
    double ret = 0.0;
    if (d < d2hat) {
        double r = d / d2hat;
        double t = r - 1.0;
        ret = kappa * -(t * t) * log(r);
    }
    return ret;
}
// ipcsk/barrier.py:18-24 barrier_derivative
__device__ __forceinline__ double barrier_d1(double d, double d2hat, double kappa) {
    double ret = 0.0;
    if (d < d2hat) ret = kappa * (d2hat - d) * (2.0 * log(d / d2hat) + (d - d2hat) / d) / (d2hat * d2hat);
    return ret;
}
// ipcsk/barrier.py:34-39 barrier_derivative2
__device__ __forceinline__ double barrier_d2(double d, double d2hat, double kappa) {
    double ret = 0.0;
    if (d < d2hat)
        ret = -kappa * (2.0 * log(d / d2hat) + (d - d2hat) / d + (d - d2hat) * (2.0 / d + d2hat / (d * d))) /
              (d2hat * d2hat);
    return ret;
}

// Warp mat22 inverse (adjugate * 1/det, zero matrix for det == 0) applied to b.
__device__ __forceinline__ void inv22_mul(double m00, double m01, double m10, double m11, double b0, double b1,
                                          double& r0, double& r1) {
    double det = m00 * m11 - m10 * m01;
    double i00 = 0.0, i01 = 0.0, i10 = 0.0, i11 = 0.0;
    if (fabs(det) > 0.0) {
        double s = 1.0 / det;
        i00 = m11 * s;
        i01 = -m01 * s;
        i10 = -m10 * s;
        i11 = m00 * s;
    }
    r0 = i00 * b0 + i01 * b1;
    r1 = i10 * b0 + i11 * b1;
}

// Contact-space coordinates c = (e0, e1perp, e2perp) with (alpha, beta, gamma):
// psd/vf.py:4-41 (PT: x0 point, x1..x3 triangle) and psd/ee.py:6-31,295-308
// (EE: edges x0x1, x2x3).  The two maps differ only in the edge vectors.
struct CMap {
    d3 e0, e1p, e2p;
    double alpha, beta, gamma;
};
__device__ __forceinline__ CMap cmap_from_edges(d3 e0, d3 e1, d3 e2) {
    CMap c;
    double e0e0 = dot(e0, e0), e0e1 = dot(e0, e1), e1e1 = dot(e1, e1);
    c.alpha = e0e1 / e0e0;
    c.e0 = e0;
    c.e1p = e1 - e0 * c.alpha;
    inv22_mul(e0e0, e0e1, e0e1, e1e1, dot(e0, e2), dot(e1, e2), c.beta, c.gamma);
    c.e2p = (e2 - e0 * c.beta) - e1 * c.gamma;
    return c;
}
__device__ __forceinline__ CMap cmap_pt(d3 x0, d3 x1, d3 x2, d3 x3) {
    return cmap_from_edges(x1 - x2, x3 - x2, x0 - x2);
}
__device__ __forceinline__ CMap cmap_ee(d3 x0, d3 x1, d3 x2, d3 x3) {
    return cmap_from_edges(x1 - x0, x3 - x2, x2 - x0);
}
// ccd.py:45-54 verify_root_pt / verify_root_ee (strict)
__device__ __forceinline__ bool valid_pt(double b, double g) {
    return 0.0 < b && b < 1.0 && 0.0 < g && g < 1.0 && 0.0 < b + g && b + g < 1.0;
}
__device__ __forceinline__ bool valid_ee(double b, double g) { return 0.0 < b && b < 1.0 && 0.0 < -g && -g < 1.0; }
// psd/hl.py:16-17
__device__ __forceinline__ double signed_distance(d3 e0p, d3 e1p, d3 e2p) {
    d3 n = cross(e0p, e1p);
    return dot(e2p, n) / sqrt(dot(n, n));
}

// ------------------------------------------------------------------ CCD
// cubic_roots.py:6-19
__device__ __forceinline__ double evaluate(const double c[4], double x) {
    return c[0] + x * (c[1] + x * (c[2] + x * c[3]));
}
__device__ __forceinline__ double multi_sign(double a, double b) { return b < 0.0 ? -a : a; }
__device__ __forceinline__ bool diff_sign(double y0, double yr) { return y0 * yr < 0.0; }
__device__ __forceinline__ double wmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double wmax(double a, double b) { return a > b ? a : b; }

constexpr double kXError = 1e-10;
// Bound on the Newton/bisection loop; the reference loop is unbounded, and the
// bracket shrinks strictly every step, so real inputs stop long before this.
constexpr int kFindClosedMaxIt = 1000;

// cubic_roots.py:29-62 find_closed
__device__ inline double find_closed(const double c[4], const double dv[4], double x0, double x1, double y0) {
    double xr = (x0 + x1) / 2.0;
    double yr = evaluate(c, xr);
    double xb0 = x0, xb1 = x1;
    if (!(x1 - x0 <= kXError * 2.0)) {
        for (int it = 0; it < kFindClosedMaxIt; ++it) {
            bool side = diff_sign(y0, yr);
            if (side)
                xb1 = xr;
            else
                xb0 = xr;
            double dy = evaluate(dv, xr);
            double dx = yr / dy;
            double xn = xr - dx;
            if (xn > xb0 && xn < xb1) {
                double stepsize = fabs(xr - xn);
                xr = xn;
                if (stepsize > kXError)
                    yr = evaluate(c, xr);
                else
                    break;
            } else {
                xr = (xb0 + xb1) / 2.0;
                yr = evaluate(c, xr);
                if (xb0 == xr || xb1 == xr || xb1 - xb0 <= 2.0 * kXError) break;
            }
        }
    }
    return xr;
}
// cubic_roots.py:63-93 quadratic_roots (constant-first)
__device__ inline int quadratic_roots(const double coef[3], double x0, double x1, double roots[2]) {
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
// cubic_roots.py:21-27 deflate
__device__ __forceinline__ void deflate(const double c[4], double root, double d[3]) {
    d[2] = c[3];
    d[1] = c[2] + root * d[2];
    d[0] = c[1] + root * d[1];
}
// cubic_roots.py:95-193 cubic_roots (one pass of the reference's while-True block)
__device__ inline int cubic_roots(const double coef[4], double x0, double x1, double roots[3]) {
    roots[0] = roots[1] = roots[2] = 0.0;
    double y0 = evaluate(coef, x0), y1 = evaluate(coef, x1);
    double a = coef[3] * 3.0, b_2 = coef[2], c = coef[1];
    double deriv[4] = {c, 2.0 * b_2, a, 0.0};
    double delta_4 = b_2 * b_2 - a * c;
    double qr[2], dp[3];
    if (delta_4 > 0.0) {
        double d_2 = sqrt(delta_4);
        double q = -(b_2 + multi_sign(d_2, b_2));
        double rv0 = q / a, rv1 = c / q;
        double xa = wmin(rv0, rv1), xb = wmax(rv0, rv1);
        bool outside = xa >= x1 || xb <= x0 || (xa <= x0 && xb >= x1);
        if (diff_sign(y0, y1)) {
            if (outside) {
                roots[0] = find_closed(coef, deriv, x0, x1, y0);
                return 1;
            }
        } else if (outside) {
            return 0;
        }
        if (xa > x0) {
            double ya = evaluate(coef, xa);
            if (diff_sign(y0, ya)) {
                roots[0] = find_closed(coef, deriv, x0, xa, y0);
                if (diff_sign(ya, y1) || (xb < x1 && diff_sign(ya, evaluate(coef, xb)))) {
                    deflate(coef, roots[0], dp);
                    int r = quadratic_roots(dp, xa, x1, qr);
                    roots[1] = qr[0];
                    roots[2] = qr[1];
                    return r + 1;
                }
                return 1;
            }
            if (xb < x1) {
                double yb = evaluate(coef, xb);
                if (diff_sign(ya, yb)) {
                    roots[0] = find_closed(coef, deriv, xa, xb, ya);
                    if (diff_sign(yb, y1)) {
                        deflate(coef, roots[0], dp);
                        int r = quadratic_roots(dp, xb, x1, qr);
                        roots[1] = qr[0];
                        roots[2] = qr[1];
                        return r + 1;
                    }
                    return 1;
                }
                if (diff_sign(yb, y1)) {
                    roots[0] = find_closed(coef, deriv, xb, x1, yb);
                    return 1;
                }
            } else if (diff_sign(ya, y1)) {
                roots[0] = find_closed(coef, deriv, xa, x1, ya);
                return 1;
            }
        } else {
            double yb = evaluate(coef, xb);
            if (diff_sign(y0, yb)) {
                roots[0] = find_closed(coef, deriv, x0, xb, y0);
                if (diff_sign(yb, y1)) {
                    deflate(coef, roots[0], dp);
                    int r = quadratic_roots(dp, xb, x1, qr);
                    roots[1] = qr[0];
                    roots[2] = qr[1];
                    return r + 1;
                }
                return 1;
            }
            if (diff_sign(yb, y1)) {
                roots[0] = find_closed(coef, deriv, xb, x1, yb);
                return 1;
            }
        }
        return 0;
    }
    if (diff_sign(y0, y1)) {
        roots[0] = find_closed(coef, deriv, x0, x1, y0);
        return 1;
    }
    return 0;
}

// ccd.py:122-129 cubic_binomial: prod_k (b_k + t a_k), constant first, accumulated into acc
__device__ __forceinline__ void cubic_binomial_acc(double a0, double a1, double a2, double b0, double b1, double b2,
                                                   double acc[4]) {
    acc[0] += b0 * b1 * b2;
    acc[1] += a0 * b1 * b2 + b0 * b1 * a2 + b0 * a1 * b2;
    acc[2] += a0 * a1 * b2 + a0 * b1 * a2 + b0 * a1 * a2;
    acc[3] += a0 * a1 * a2;
}
// ccd.py:82-119 det_polynomial of M(t) = b + t a, with a, b given by their columns
// (R1: mat33(u, v, w) has columns u, v, w, so m[r][c] = col_c[r]).
__device__ inline void det_polynomial(const d3 ac[3], const d3 bc[3], double out[4]) {
#define M_(m, r, c) comp(m[c], r)
    double pos[4] = {0.0, 0.0, 0.0, 0.0}, neg[4] = {0.0, 0.0, 0.0, 0.0};
    cubic_binomial_acc(M_(ac, 0, 0), M_(ac, 1, 1), M_(ac, 2, 2), M_(bc, 0, 0), M_(bc, 1, 1), M_(bc, 2, 2), pos);
    cubic_binomial_acc(M_(ac, 0, 1), M_(ac, 1, 2), M_(ac, 2, 0), M_(bc, 0, 1), M_(bc, 1, 2), M_(bc, 2, 0), pos);
    cubic_binomial_acc(M_(ac, 0, 2), M_(ac, 1, 0), M_(ac, 2, 1), M_(bc, 0, 2), M_(bc, 1, 0), M_(bc, 2, 1), pos);
    cubic_binomial_acc(M_(ac, 0, 0), M_(ac, 1, 2), M_(ac, 2, 1), M_(bc, 0, 0), M_(bc, 1, 2), M_(bc, 2, 1), neg);
    cubic_binomial_acc(M_(ac, 0, 1), M_(ac, 1, 0), M_(ac, 2, 2), M_(bc, 0, 1), M_(bc, 1, 0), M_(bc, 2, 2), neg);
    cubic_binomial_acc(M_(ac, 0, 2), M_(ac, 1, 1), M_(ac, 2, 0), M_(bc, 0, 2), M_(bc, 1, 1), M_(bc, 2, 0), neg);
#undef M_
    for (int k = 0; k < 4; ++k) out[k] = pos[k] - neg[k];
}
// ccd.py:56-79 build_and_solve_4_points_coplanar: coefficients of the coplanarity cubic
__device__ inline void coplanar_poly(const d3 p0[4], const d3 p1[4], double coef[4]) {
    static constexpr int sel[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    double d[4][4];
    for (int m = 0; m < 4; ++m) {
        d3 ac[3], bc[3];
        for (int k = 0; k < 3; ++k) {
            bc[k] = p0[sel[m][k]];
            ac[k] = p1[sel[m][k]] - bc[k];
        }
        det_polynomial(ac, bc, d[m]);
    }
    for (int k = 0; k < 4; ++k) coef[k] = d[0][k] - d[1][k] + d[2][k] - d[3][k];
}
// wp.lerp(a, b, t) = a*(1-t) + b*t
__device__ __forceinline__ d3 lerp(d3 a, d3 b, double t) { return a * (1.0 - t) + b * t; }
// ccd.py:133-186 pt_collision_time (kind 0) / ee_collision_time (kind 1)
__device__ inline double collision_time(int kind, const d3 t0[4], const d3 t1[4]) {
    double coef[4], roots[3];
    coplanar_poly(t0, t1, coef);
    int n = cubic_roots(coef, 0.0, 1.0, roots);
    double root = 1.0;
    bool true_root = false;
    for (int i = 0; i < n; ++i) {
        root = roots[i];
        d3 p[4];
        for (int k = 0; k < 4; ++k) p[k] = lerp(t0[k], t1[k], root);
        CMap c = kind == 0 ? cmap_pt(p[0], p[1], p[2], p[3]) : cmap_ee(p[0], p[1], p[2], p[3]);
        true_root = kind == 0 ? valid_pt(c.beta, c.gamma) : valid_ee(c.beta, c.gamma);
        if (true_root) break;
    }
    if (!true_root) root = 1.0;
    return root;
}

// ------------------------------------------------------------- fetchers
// affine_body.py:148-263 fetch_pt / fetch_ee over the flattened scene.
struct SceneView {
    const int32_t* voff;
    const int32_t* foff;
    const int32_t* eoff;
    const int32_t* tris;
    const int32_t* edges;
    const double* x0;
};
__device__ __forceinline__ void fetch_pt_idx(const SceneView& S, const int32_t* t5, int64_t v[4]) {
    int I = t5[0], J = t5[1];
    const int32_t* T = S.tris + 3 * (int64_t)(S.foff[J] + t5[4]);
    int vj = S.voff[J];
    v[0] = S.voff[I] + t5[3];
    v[1] = vj + T[0];
    v[2] = vj + T[1];
    v[3] = vj + T[2];
}
__device__ __forceinline__ void fetch_ee_idx(const SceneView& S, const int32_t* t5, int64_t v[4]) {
    int I = t5[0], J = t5[1];
    const int32_t* EI = S.edges + 2 * (int64_t)(S.eoff[I] + t5[3]);
    const int32_t* EJ = S.edges + 2 * (int64_t)(S.eoff[J] + t5[4]);
    int vi = S.voff[I], vj = S.voff[J];
    v[0] = vi + EI[0];
    v[1] = vi + EI[1];
    v[2] = vj + EJ[0];
    v[3] = vj + EJ[1];
}

}  // namespace abd

namespace abd {
// Per-candidate barrier energy (ipcsk/energy.py:10-48) and TOI (ccd.py:9-40), shared
// by the whole-list and the per-scene kernels.
// Vertex positions of the candidate functions: read from an array (Verts), or computed
// on the fly by another source with the same interface (load4: vertices vi[0..nI) of
// body I, the rest of body J; y: the y coordinate of vertex vid of body b).
struct Verts {
    const double* x;
    __device__ __forceinline__ void load4(int, int, const int64_t vi[4], int, d3 v[4]) const {
        for (int k = 0; k < 4; ++k) v[k] = ld3(x + 3 * vi[k]);
    }
    __device__ __forceinline__ double y(int, int64_t vid) const { return x[3 * vid + 1]; }
};

template <class Src>
__device__ __forceinline__ double cand_energy_ee(const SceneView& S, const Src& X, const int32_t* t5, double d2hat,
                                                 double kappa) {
    int64_t vi[4];
    fetch_ee_idx(S, t5, vi);
    d3 v[4];
    X.load4(t5[0], t5[1], vi, 2, v);
    CMap c = cmap_ee(v[0], v[1], v[2], v[3]);
    double d = signed_distance(c.e0, c.e1p, c.e2p);
    return valid_ee(c.beta, c.gamma) ? barrier(d * d, d2hat, kappa) : 0.0;
}
template <class Src>
__device__ __forceinline__ double cand_energy_pt(const SceneView& S, const Src& X, const int32_t* t5, double d2hat,
                                                 double kappa) {
    int64_t vi[4];
    fetch_pt_idx(S, t5, vi);
    d3 v[4];
    X.load4(t5[0], t5[1], vi, 1, v);
    CMap c = cmap_pt(v[0], v[1], v[2], v[3]);
    return valid_pt(c.beta, c.gamma) ? barrier(dot(c.e2p, c.e2p), d2hat, kappa) : 0.0;
}
template <class Src>
__device__ __forceinline__ double cand_energy_vg(const int32_t* __restrict__ voff, const Src& X, const int32_t* t2,
                                                 double ground, double d2hat, double kappa) {
    const int64_t vid = voff[t2[0]] + t2[1];
    const double d = X.y(t2[0], vid) - ground;
    return barrier(d * d, d2hat, kappa);
}
__device__ __forceinline__ double cand_toi(int kind, const SceneView& S, const double* __restrict__ x0,
                                           const double* __restrict__ x1, const int32_t* t) {
    int64_t vi[4];
    if (kind == 0)
        fetch_pt_idx(S, t, vi);
    else
        fetch_ee_idx(S, t, vi);
    d3 a[4], b[4];
    for (int q = 0; q < 4; ++q) {
        a[q] = ld3(x0 + 3 * vi[q]);
        b[q] = ld3(x1 + 3 * vi[q]);
    }
    return collision_time(kind, a, b);
}
__device__ __forceinline__ double cand_toi_vg(const int32_t* __restrict__ voff, const double* __restrict__ x0,
                                              const double* __restrict__ x1, const int32_t* t2, double ground) {
    const int64_t vid = voff[t2[0]] + t2[1];
    const double d1 = x1[3 * vid + 1] - ground;
    double t = 1.0;
    if (d1 < 0) {
        const double d0 = x0[3 * vid + 1] - ground;
        t = d0 / (d0 - d1);
    }
    return t;
}
}  // namespace abd

namespace abd {
// Exact derivatives of the squared plane / line distance used by the barrier,
// s = T^2 / N with T = w.(u x v) and N = |u x v|^2 (u, v, w the three edge
// vectors of cmap_from_edges).  g: ds/d(u, v, w) (9), H: d2s/d(u, v, w)^2 (9x9,
// row-major).  With
//   dT = (v x w, w x u, u x v),  d2T: uv = -[w]x, uw = [v]x, vw = -[u]x (transposes below),
//   dN = (2 v x c, 2 c x u, 0),  d2N: uu = 2(|v|^2 I - v v^T), vv = 2(|u|^2 I - u u^T),
//                                  uv = 2(u v^T - (u.v) I) - 2[c]x   (c = u x v),
//   ds  = (2T/N) dT - (T^2/N^2) dN,
//   d2s = (2/N) dT dT^T + (2T/N) d2T - (2T/N^2)(dT dN^T + dN dT^T) - (T^2/N^2) d2N
//         + (2T^2/N^3) dN dN^T.
__device__ inline void skew_add(double* H, int r0, int c0, d3 a, double s) {
    // H[r0.., c0..] += s [a]x   ([a]x b = a x b)
    H[9 * (r0 + 0) + c0 + 1] += -a.z * s;
    H[9 * (r0 + 0) + c0 + 2] += a.y * s;
    H[9 * (r0 + 1) + c0 + 0] += a.z * s;
    H[9 * (r0 + 1) + c0 + 2] += -a.x * s;
    H[9 * (r0 + 2) + c0 + 0] += -a.y * s;
    H[9 * (r0 + 2) + c0 + 1] += a.x * s;
}
__device__ inline void plane_dist2_derivs(d3 u, d3 v, d3 w, double g[9], double H[81]) {
    const d3 c = cross(u, v);
    const double T = dot(w, c), N = dot(c, c);
    const d3 gT[3] = {cross(v, w), cross(w, u), c};
    const d3 gN[3] = {cross(v, c) * 2.0, cross(c, u) * 2.0, mk(0.0, 0.0, 0.0)};
    const double a1 = 2.0 * T / N, a2 = T * T / (N * N);
    for (int e = 0; e < 3; ++e)
        for (int k = 0; k < 3; ++k) g[3 * e + k] = a1 * comp(gT[e], k) - a2 * comp(gN[e], k);
    const double b0 = 2.0 / N, b1 = 2.0 * T / (N * N), b2 = 2.0 * T * T / (N * N * N);
    for (int i = 0; i < 9; ++i)
        for (int j = 0; j < 9; ++j) {
            const double ti = comp(gT[i / 3], i % 3), tj = comp(gT[j / 3], j % 3);
            const double ni = comp(gN[i / 3], i % 3), nj = comp(gN[j / 3], j % 3);
            H[9 * i + j] = b0 * (ti * tj) - b1 * (ti * nj + ni * tj) + b2 * (ni * nj);
        }
    // (2T/N) d2T
    skew_add(H, 0, 3, w, -a1);
    skew_add(H, 3, 0, w, a1);
    skew_add(H, 0, 6, v, a1);
    skew_add(H, 6, 0, v, -a1);
    skew_add(H, 3, 6, u, -a1);
    skew_add(H, 6, 3, u, a1);
    // -(T^2/N^2) d2N
    const double uu = dot(u, u), vv = dot(v, v), uv = dot(u, v);
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) {
            const double I = r == k ? 1.0 : 0.0;
            H[9 * r + k] -= a2 * (2.0 * (vv * I - comp(v, r) * comp(v, k)));
            H[9 * (3 + r) + 3 + k] -= a2 * (2.0 * (uu * I - comp(u, r) * comp(u, k)));
            H[9 * r + 3 + k] -= a2 * (2.0 * (comp(u, r) * comp(v, k) - uv * I));
            H[9 * (3 + r) + k] -= a2 * (2.0 * (comp(v, r) * comp(u, k) - uv * I));
        }
    skew_add(H, 0, 3, c, 2.0 * a2);
    skew_add(H, 3, 0, c, -2.0 * a2);
}
}  // namespace abd
