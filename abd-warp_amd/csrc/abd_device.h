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
// psd/vf.py:4-41 (PT: x0 point, x1..x3 triangle) and psd/ee.py:6-31,119-133
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
// cubic_roots.py:95-193 cubic_roots, split into its bracketing decisions (which interval
// the Newton / bisection search runs on, and whether the deflated quadratic follows) and
// the searches themselves: a candidate whose bracket is kNone has no root in [x0, x1]
// without any search.
struct CubicBracket {
    int kind;  // kNone: no root; kOne: one root in [lo, hi]; kDeflate: it, then the
               // deflated quadratic's roots in [qlo, x1]
    double lo, hi, ylo, qlo;
    static constexpr int kNone = 0, kOne = 1, kDeflate = 2;
};
__device__ inline CubicBracket cubic_bracket(const double coef[4], double x0, double x1) {
    const double y0 = evaluate(coef, x0), y1 = evaluate(coef, x1);
    const double a = coef[3] * 3.0, b_2 = coef[2], c = coef[1];
    const double delta_4 = b_2 * b_2 - a * c;
    if (delta_4 > 0.0) {
        const double d_2 = sqrt(delta_4);
        const double q = -(b_2 + multi_sign(d_2, b_2));
        const double rv0 = q / a, rv1 = c / q;
        const double xa = wmin(rv0, rv1), xb = wmax(rv0, rv1);
        const bool outside = xa >= x1 || xb <= x0 || (xa <= x0 && xb >= x1);
        if (diff_sign(y0, y1)) {
            if (outside) return {CubicBracket::kOne, x0, x1, y0, 0.0};
        } else if (outside) {
            return {CubicBracket::kNone, 0.0, 0.0, 0.0, 0.0};
        }
        if (xa > x0) {
            const double ya = evaluate(coef, xa);
            if (diff_sign(y0, ya)) {
                const bool more = diff_sign(ya, y1) || (xb < x1 && diff_sign(ya, evaluate(coef, xb)));
                return {more ? CubicBracket::kDeflate : CubicBracket::kOne, x0, xa, y0, xa};
            }
            if (xb < x1) {
                const double yb = evaluate(coef, xb);
                if (diff_sign(ya, yb))
                    return {diff_sign(yb, y1) ? CubicBracket::kDeflate : CubicBracket::kOne, xa, xb, ya, xb};
                if (diff_sign(yb, y1)) return {CubicBracket::kOne, xb, x1, yb, 0.0};
            } else if (diff_sign(ya, y1)) {
                return {CubicBracket::kOne, xa, x1, ya, 0.0};
            }
        } else {
            const double yb = evaluate(coef, xb);
            if (diff_sign(y0, yb))
                return {diff_sign(yb, y1) ? CubicBracket::kDeflate : CubicBracket::kOne, x0, xb, y0, xb};
            if (diff_sign(yb, y1)) return {CubicBracket::kOne, xb, x1, yb, 0.0};
        }
        return {CubicBracket::kNone, 0.0, 0.0, 0.0, 0.0};
    }
    if (diff_sign(y0, y1)) return {CubicBracket::kOne, x0, x1, y0, 0.0};
    return {CubicBracket::kNone, 0.0, 0.0, 0.0, 0.0};
}
// the searches of a bracket (one pass of the reference's while-True block)
__device__ inline int cubic_roots(const double coef[4], double x0, double x1, double roots[3]) {
    roots[0] = roots[1] = roots[2] = 0.0;
    const CubicBracket br = cubic_bracket(coef, x0, x1);
    if (br.kind == CubicBracket::kNone) return 0;
    const double deriv[4] = {coef[1], 2.0 * coef[2], coef[3] * 3.0, 0.0};
    roots[0] = find_closed(coef, deriv, br.lo, br.hi, br.ylo);
    if (br.kind == CubicBracket::kOne) return 1;
    double qr[2], dp[3];
    deflate(coef, roots[0], dp);
    const int r = quadratic_roots(dp, br.qlo, x1, qr);
    roots[1] = qr[0];
    roots[2] = qr[1];
    return r + 1;
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
// affine_body.py:106-234 fetch_pt / fetch_ee over the flattened scene.
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
// The four scene vertex ids of a PT (kind 0) / EE (kind 1) candidate: from `vid` (the ids
// fetch_pt_idx / fetch_ee_idx give, stored beside the list; nullable) or fetched through
// the topology.
__device__ __forceinline__ void cand_idx(int kind, const SceneView& S, const int32_t* t5, const int32_t* vid,
                                         int64_t v[4]) {
    if (vid) {
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = vid[q];
    } else if (kind == 0) {
        fetch_pt_idx(S, t5, v);
    } else {
        fetch_ee_idx(S, t5, v);
    }
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
                                                 double kappa, const int32_t* vid = nullptr) {
    int64_t vi[4];
    cand_idx(1, S, t5, vid, vi);
    d3 v[4];
    X.load4(t5[0], t5[1], vi, 2, v);
    CMap c = cmap_ee(v[0], v[1], v[2], v[3]);
    double d = signed_distance(c.e0, c.e1p, c.e2p);
    return valid_ee(c.beta, c.gamma) ? barrier(d * d, d2hat, kappa) : 0.0;
}
template <class Src>
__device__ __forceinline__ double cand_energy_pt(const SceneView& S, const Src& X, const int32_t* t5, double d2hat,
                                                 double kappa, const int32_t* vid = nullptr) {
    int64_t vi[4];
    cand_idx(0, S, t5, vid, vi);
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
                                           const double* __restrict__ x1, const int32_t* t,
                                           const int32_t* vid = nullptr) {
    int64_t vi[4];
    cand_idx(kind, S, t, vid, vi);
    d3 a[4], b[4];
    for (int q = 0; q < 4; ++q) {
        a[q] = ld3(x0 + 3 * vi[q]);
        b[q] = ld3(x1 + 3 * vi[q]);
    }
    return collision_time(kind, a, b);
}
// The four vertices of a PT (kind 0) / EE (kind 1) candidate at both ends of the step.
__device__ __forceinline__ void cand_verts(int kind, const SceneView& S, const double* __restrict__ x0,
                                           const double* __restrict__ x1, const int32_t* t, d3 a[4], d3 b[4],
                                           const int32_t* vid = nullptr) {
    int64_t vi[4];
    cand_idx(kind, S, t, vid, vi);
    for (int q = 0; q < 4; ++q) {
        a[q] = ld3(x0 + 3 * vi[q]);
        b[q] = ld3(x1 + 3 * vi[q]);
    }
}
// Whether collision_time of the candidate can be below 1: its coplanarity cubic has a
// bracketed root in [0, 1] (cubic_bracket); otherwise it has no root and its TOI is 1.
__device__ __forceinline__ bool cand_toi_bracketed(int kind, const SceneView& S, const double* __restrict__ x0,
                                                   const double* __restrict__ x1, const int32_t* t,
                                                   const int32_t* vid = nullptr) {
    d3 a[4], b[4];
    cand_verts(kind, S, x0, x1, t, a, b, vid);
    double coef[4];
    coplanar_poly(a, b, coef);
    return cubic_bracket(coef, 0.0, 1.0).kind != CubicBracket::kNone;
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

namespace abd {
// ------------------------------------------------------------------ full IPC
// psd = 3 (beyond the reference, which has no distance types): the squared distance of
// a PT / EE candidate to its closest feature pair (Li et al. 2020, Incremental Potential
// Contact), the edge-edge mollifier, and their exact derivatives in the 12 coordinates
// of the contact's four points x0..x3 (PT: point, triangle t0 t1 t2; EE: a0 a1, b0 b1).

// Feature pairs: kFeatPlane (PT: point-plane; EE: line-line), kFeatPE (point x[i] against
// edge x[j] x[k]), kFeatPP (x[i] against x[j]).
enum { kFeatPlane = 0, kFeatPE = 1, kFeatPP = 2 };
struct Feat {
    int kind, i, j, k;
};

// Closest feature of point x0 to triangle x1 x2 x3 by the Voronoi regions of the
// triangle (Ericson, Real-Time Collision Detection, 2005, sec. 5.1.5).
__device__ __forceinline__ Feat pt_feature(const d3 x[4]) {
    const d3 a = x[1], b = x[2], c = x[3], p = x[0];
    const d3 ab = b - a, ac = c - a, ap = p - a;
    const double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) return Feat{kFeatPP, 0, 1, 0};
    const d3 bp = p - b;
    const double d3_ = dot(ab, bp), d4 = dot(ac, bp);
    if (d3_ >= 0.0 && d4 <= d3_) return Feat{kFeatPP, 0, 2, 0};
    const double vc = d1 * d4 - d3_ * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3_ <= 0.0) return Feat{kFeatPE, 0, 1, 2};
    const d3 cp = p - c;
    const double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) return Feat{kFeatPP, 0, 3, 0};
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) return Feat{kFeatPE, 0, 1, 3};
    const double va = d3_ * d6 - d5 * d4;
    if (va <= 0.0 && d4 - d3_ >= 0.0 && d5 - d6 >= 0.0) return Feat{kFeatPE, 0, 2, 3};
    return Feat{kFeatPlane, 0, 0, 0};
}

// Closest feature pair of segments x0 x1 and x2 x3: the line parameters of the closest
// points, clamped to the segments in turn (Sunday, "Distance between 3D lines and
// segments", geomalgorithms.com); (near-)parallel segments start from s = 0.
__device__ __forceinline__ Feat ee_feature(const d3 x[4]) {
    const d3 u = x[1] - x[0], v = x[3] - x[2], w = x[0] - x[2];
    const double a = dot(u, u), b = dot(u, v), c = dot(v, v), d = dot(u, w), e = dot(v, w);
    const double D = a * c - b * b;
    double sN, sD = D, tN, tD = D;
    int s_end = -1, t_end = -1;  // -1: interior, 0 / 1: that end
    if (D <= 1e-12 * a * c) {
        sN = 0.0;
        sD = 1.0;
        tN = e;
        tD = c;
        s_end = 0;
    } else {
        sN = b * e - c * d;
        tN = a * e - b * d;
        if (sN <= 0.0) {
            sN = 0.0;
            tN = e;
            tD = c;
            s_end = 0;
        } else if (sN >= sD) {
            sN = sD;
            tN = e + b;
            tD = c;
            s_end = 1;
        }
    }
    if (tN <= 0.0) {
        t_end = 0;
        if (-d <= 0.0)
            s_end = 0;
        else if (-d >= a)
            s_end = 1;
        else
            s_end = -1;
    } else if (tN >= tD) {
        t_end = 1;
        if (-d + b <= 0.0)
            s_end = 0;
        else if (-d + b >= a)
            s_end = 1;
        else
            s_end = -1;
    }
    if (s_end < 0 && t_end < 0) return Feat{kFeatPlane, 0, 0, 0};
    if (s_end >= 0 && t_end >= 0) return Feat{kFeatPP, s_end, 2 + t_end, 0};
    if (s_end >= 0) return Feat{kFeatPE, s_end, 2, 3};
    return Feat{kFeatPE, 2 + t_end, 0, 1};
}

// C = |a x b|^2 and its derivatives in (a, b): g (6) = (2 b x c, 2 c x a) with c = a x b
// (the cross-product form keeps full relative accuracy when a and b are nearly parallel),
// H (36, row-major) from C = |a|^2 |b|^2 - (a.b)^2.
__device__ inline double cross2_derivs(d3 a, d3 b, double g[6], double H[36]) {
    const double aa = dot(a, a), bb = dot(b, b), ab = dot(a, b);
    const d3 c = cross(a, b), ga = cross(b, c), gb = cross(c, a);
    for (int k = 0; k < 3; ++k) {
        g[k] = 2.0 * comp(ga, k);
        g[3 + k] = 2.0 * comp(gb, k);
    }
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) {
            const double I = r == k ? 1.0 : 0.0;
            H[6 * r + k] = 2.0 * bb * I - 2.0 * comp(b, r) * comp(b, k);
            H[6 * (3 + r) + 3 + k] = 2.0 * aa * I - 2.0 * comp(a, r) * comp(a, k);
            H[6 * r + 3 + k] = 4.0 * comp(a, r) * comp(b, k) - 2.0 * comp(b, r) * comp(a, k) - 2.0 * ab * I;
            H[6 * (3 + r) + k] = 4.0 * comp(b, r) * comp(a, k) - 2.0 * comp(a, r) * comp(b, k) - 2.0 * ab * I;
        }
    return dot(c, c);
}

// The squared distance s of the feature pair f and ds, d2s in the 12 point coordinates.
__device__ inline double feature_dist2(int ckind, const Feat& f, const d3 x[4], double g[12], double H[144]) {
    for (int i = 0; i < 12; ++i) g[i] = 0.0;
    for (int i = 0; i < 144; ++i) H[i] = 0.0;
    if (f.kind == kFeatPP) {
        const d3 dlt = x[f.i] - x[f.j];
        for (int k = 0; k < 3; ++k) {
            g[3 * f.i + k] = 2.0 * comp(dlt, k);
            g[3 * f.j + k] = -2.0 * comp(dlt, k);
            H[12 * (3 * f.i + k) + 3 * f.i + k] = 2.0;
            H[12 * (3 * f.j + k) + 3 * f.j + k] = 2.0;
            H[12 * (3 * f.i + k) + 3 * f.j + k] = -2.0;
            H[12 * (3 * f.j + k) + 3 * f.i + k] = -2.0;
        }
        return dot(dlt, dlt);
    }
    if (f.kind == kFeatPE) {
        // s = C / N, C = |a x b|^2 with a = e0 - p, b = e1 - p; N = |e1 - e0|^2
        const d3 p = x[f.i], e0 = x[f.j], e1 = x[f.k];
        const d3 a = e0 - p, b = e1 - p, e = e1 - e0;
        double gC[6], HC[36];
        const double C = cross2_derivs(a, b, gC, HC);
        const double N = dot(e, e);
        double gN[6];
        for (int k = 0; k < 3; ++k) {
            gN[k] = -2.0 * comp(e, k);
            gN[3 + k] = 2.0 * comp(e, k);
        }
        double gs[6], Hs[36];
        for (int i = 0; i < 6; ++i) gs[i] = gC[i] / N - C * gN[i] / (N * N);
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) {
                const double d2N = ((i < 3) == (j < 3) ? 2.0 : -2.0) * (i % 3 == j % 3 ? 1.0 : 0.0);
                Hs[6 * i + j] = HC[6 * i + j] / N - (gC[i] * gN[j] + gN[i] * gC[j]) / (N * N) - C * d2N / (N * N) +
                                2.0 * C * gN[i] * gN[j] / (N * N * N);
            }
        // (a, b) = (e0 - p, e1 - p): point coefficients per (a | b) row
        const int pts[3] = {f.i, f.j, f.k};
        const double La[3] = {-1.0, 1.0, 0.0}, Lb[3] = {-1.0, 0.0, 1.0};
        for (int t = 0; t < 3; ++t)
            for (int k = 0; k < 3; ++k) g[3 * pts[t] + k] += La[t] * gs[k] + Lb[t] * gs[3 + k];
        for (int t = 0; t < 3; ++t)
            for (int q = 0; q < 3; ++q) {
                const double cf[2][2] = {{La[t] * La[q], La[t] * Lb[q]}, {Lb[t] * La[q], Lb[t] * Lb[q]}};
                for (int k = 0; k < 3; ++k)
                    for (int m = 0; m < 3; ++m)
                        H[12 * (3 * pts[t] + k) + 3 * pts[q] + m] +=
                            cf[0][0] * Hs[6 * k + m] + cf[0][1] * Hs[6 * k + 3 + m] + cf[1][0] * Hs[6 * (3 + k) + m] +
                            cf[1][1] * Hs[6 * (3 + k) + 3 + m];
            }
        return C / N;
    }
    // plane / line: s = (w.(u x v))^2 / |u x v|^2 over edge vectors with point coefficients L
    double L[3][4];
    d3 u, v, w;
    if (ckind == 0) {  // u = t1 - t0, v = t2 - t0, w = p - t0
        const double Lpt[3][4] = {{0, -1, 1, 0}, {0, -1, 0, 1}, {1, -1, 0, 0}};
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 4; ++c) L[r][c] = Lpt[r][c];
        u = x[2] - x[1];
        v = x[3] - x[1];
        w = x[0] - x[1];
    } else {  // u = a1 - a0, v = b1 - b0, w = b0 - a0
        const double Lee[3][4] = {{-1, 1, 0, 0}, {0, 0, -1, 1}, {-1, 0, 1, 0}};
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 4; ++c) L[r][c] = Lee[r][c];
        u = x[1] - x[0];
        v = x[3] - x[2];
        w = x[2] - x[0];
    }
    double ge[9], He[81];
    plane_dist2_derivs(u, v, w, ge, He);
    for (int a = 0; a < 4; ++a)
        for (int k = 0; k < 3; ++k) {
            double acc = 0.0;
            for (int e = 0; e < 3; ++e) acc += L[e][a] * ge[3 * e + k];
            g[3 * a + k] = acc;
        }
    for (int a = 0; a < 4; ++a)
        for (int k = 0; k < 3; ++k)
            for (int b = 0; b < 4; ++b)
                for (int m = 0; m < 3; ++m) {
                    double acc = 0.0;
                    for (int e = 0; e < 3; ++e)
                        for (int f2 = 0; f2 < 3; ++f2) acc += L[e][a] * L[f2][b] * He[9 * (3 * e + k) + 3 * f2 + m];
                    H[12 * (3 * a + k) + 3 * b + m] = acc;
                }
    const d3 cuv = cross(u, v);
    const double T = dot(w, cuv);
    return T * T / dot(cuv, cuv);
}

// Squared distance of the feature pair alone (no derivatives).
__device__ __forceinline__ double feature_dist2_value(int ckind, const Feat& f, const d3 x[4]) {
    if (f.kind == kFeatPP) {
        const d3 dlt = x[f.i] - x[f.j];
        return dot(dlt, dlt);
    }
    if (f.kind == kFeatPE) {
        const d3 c = cross(x[f.j] - x[f.i], x[f.k] - x[f.i]), e = x[f.k] - x[f.j];
        return dot(c, c) / dot(e, e);
    }
    const d3 u = ckind == 0 ? x[2] - x[1] : x[1] - x[0];
    const d3 v = ckind == 0 ? x[3] - x[1] : x[3] - x[2];
    const d3 w = ckind == 0 ? x[0] - x[1] : x[2] - x[0];
    const d3 cuv = cross(u, v);
    const double T = dot(w, cuv);
    return T * T / dot(cuv, cuv);
}

// Edge-edge mollifier (Li et al. 2020): eps_x = 1e-3 |a1 - a0|^2 |b1 - b0|^2 at rest,
// m(c) = (2 - c / eps_x) c / eps_x for c = |(a1 - a0) x (b1 - b0)|^2 < eps_x, else 1.
__device__ __forceinline__ double ee_mollifier_eps(const d3 r[4]) {
    const d3 ua = r[1] - r[0], ub = r[3] - r[2];
    return 1e-3 * dot(ua, ua) * dot(ub, ub);
}
__device__ __forceinline__ double ee_cross2(const d3 x[4]) {
    const d3 c = cross(x[1] - x[0], x[3] - x[2]);
    return dot(c, c);
}
__device__ __forceinline__ double mollifier(double c, double eps) {
    if (c >= eps) return 1.0;
    const double t = c / eps;
    return (2.0 - t) * t;
}

// Full-IPC barrier energy of one candidate (kind 0 PT, 1 EE) at points x, rest points r.
__device__ __forceinline__ double ipc_full_energy(int kind, const d3 x[4], const d3 r[4], double d2hat,
                                                  double kappa) {
    const Feat f = kind == 0 ? pt_feature(x) : ee_feature(x);
    const double s = feature_dist2_value(kind, f, x);
    if (!(s < d2hat)) return 0.0;
    const double b = barrier(s, d2hat, kappa);
    return kind == 0 ? b : mollifier(ee_cross2(x), ee_mollifier_eps(r)) * b;
}

// Full-IPC energy of a PT (ckind 0) or EE (1) candidate tuple, points from X, rest points
// from the scene's rest positions.
template <class Src>
__device__ __forceinline__ double cand_energy_full(int ckind, const SceneView& S, const Src& X, const int32_t* t5,
                                                   double d2hat, double kappa) {
    int64_t vi[4];
    if (ckind == 0)
        fetch_pt_idx(S, t5, vi);
    else
        fetch_ee_idx(S, t5, vi);
    d3 v[4], r[4];
    X.load4(t5[0], t5[1], vi, ckind == 0 ? 1 : 2, v);
    for (int k = 0; k < 4; ++k) r[k] = ld3(S.x0 + 3 * vi[k]);
    return ipc_full_energy(ckind, v, r, d2hat, kappa);
}

// In-place projection of a symmetric 12x12 matrix onto the PSD cone (negative eigenvalues
// set to zero) by cyclic Jacobi rotations (Golub & Van Loan, Matrix Computations, 4th ed.,
// sec. 8.5.2-8.5.3), to a relative off-diagonal norm of 1e-15 or 30 sweeps.
__device__ inline void psd_project12(double A[144]) {
    double V[144];
    for (int i = 0; i < 144; ++i) V[i] = (i / 12 == i % 12) ? 1.0 : 0.0;
    double fro = 0.0;
    for (int i = 0; i < 144; ++i) fro += A[i] * A[i];
    if (fro == 0.0) return;
    for (int sweep = 0; sweep < 30; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < 12; ++p)
            for (int q = p + 1; q < 12; ++q) off += A[12 * p + q] * A[12 * p + q];
        if (off <= 1e-30 * fro) break;
        for (int p = 0; p < 11; ++p)
            for (int q = p + 1; q < 12; ++q) {
                const double apq = A[12 * p + q];
                if (apq == 0.0) continue;
                const double tau = (A[12 * q + q] - A[12 * p + p]) / (2.0 * apq);
                const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
                for (int k = 0; k < 12; ++k) {  // A <- A J (columns p, q)
                    const double akp = A[12 * k + p], akq = A[12 * k + q];
                    A[12 * k + p] = c * akp - s * akq;
                    A[12 * k + q] = s * akp + c * akq;
                }
                for (int k = 0; k < 12; ++k) {  // A <- J^T A (rows p, q)
                    const double apk = A[12 * p + k], aqk = A[12 * q + k];
                    A[12 * p + k] = c * apk - s * aqk;
                    A[12 * q + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 12; ++k) {  // V <- V J
                    const double vkp = V[12 * k + p], vkq = V[12 * k + q];
                    V[12 * k + p] = c * vkp - s * vkq;
                    V[12 * k + q] = s * vkp + c * vkq;
                }
            }
    }
    double lam[12];
    for (int i = 0; i < 12; ++i) lam[i] = A[13 * i] > 0.0 ? A[13 * i] : 0.0;
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 12; ++j) {
            double acc = 0.0;
            for (int k = 0; k < 12; ++k) acc += V[12 * i + k] * lam[k] * V[12 * j + k];
            A[12 * i + j] = acc;
        }
}
}  // namespace abd
