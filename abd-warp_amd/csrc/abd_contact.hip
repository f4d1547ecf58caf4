// abd_contact.hip — barrier energy, gradient and PSD-projected Hessian of the
// point-triangle (PT), edge-edge (EE) and vertex-ground (VG) contacts, and the
// inertia / orthogonality terms (ipc.py, ipcsk/, psd/, orthogonal_energy.py).
//
// Hessian structure.  For one contact the reference forms (ipcsk/pt_ipc.py:111-185)
//   Hl  = sum_k L_k q_k q_k^T / |q_k|^2          (9x9, psd/hl.py:156-195; L_4 = 2)
//   H12 = D~^T Hl D~,  g12 = D~^T q_4           (D~ = dcdx (x) I3, 9x12)
//   Hipc = b' H12 + b'' g12 g12^T,   Hi/Hj/Hij = J^T Hipc J   (J = [1, x0^T] (x) I3)
// which is a sum of five rank-one terms in body space: with V_k = J^T D~^T q_k
// (24 entries: 12 per body),
//   H24 = sum_{k<4} (b' L_k/|q_k|^2) V_k V_k^T + (b' L_4/|q_4|^2 + b'') V_4 V_4^T,
//   g24 = b' V_4.
// A contact is therefore stored as 5 coefficients + 5 x 24 vectors, and the
// body-pair blocks are reduced from those in a fixed order (no atomics):
//   stage A  classify every candidate (active = s < d2hat and valid)      [HBM stream]
//   stage B  ordered compaction of the active candidates
//   stage C  per active contact: coefficients and V_k (one thread each)
//   stage D  per body pair: sum of its contacts' 24x24 terms; the off-diagonal
//            12x12 goes straight into the pair's two slots
//   stage E  per body: diagonal slot and gradient += its pairs' diagonal parts
//            + its vertex-ground terms.
#include <hip/hip_runtime.h>

#include <cstring>

#include "abd_block.h"
#include "abd_device.h"
#include "abd_inertia.h"
#include "abd_internal.h"
#include "abd_profile.h"

namespace abd {

constexpr int kCItems = 8;          // candidates per thread in the compaction passes
constexpr int kGhBoundMib = 8192;  // stored contacts sized by a list bound, at most (default)
constexpr int kContactStride = 128;  // doubles per stored contact: c[0..4], b', pad, then V[5][24]
constexpr int kCoefOff = 0;
constexpr int kB1Off = 5;
constexpr int kVOff = 8;

__device__ __forceinline__ SceneView view_of(const AbdScene& sc) {
    SceneView v;
    v.voff = sc.voff;
    v.foff = sc.foff;
    v.eoff = sc.eoff;
    v.tris = sc.tris;
    v.edges = sc.edges;
    v.x0 = sc.x0;
    return v;
}

__device__ __forceinline__ void load4(const double* __restrict__ x, const int64_t v[4], d3 out[4]) {
#pragma unroll
    for (int k = 0; k < 4; ++k) out[k] = ld3(x + 3 * v[k]);
}


// ------------------------------------------------------------ contact terms
// One contact: writes coefficients and V_k into out[kContactStride]; returns
// active.  kind 0 PT (x0 point; x1..x3 triangle), 1 EE (x0x1, x2x3).
__device__ bool contact_terms(int kind, const d3 v[4], const d3 r[4], double d2hat, double kappa, int psd,
                              double* __restrict__ out) {
    CMap c = kind == 0 ? cmap_pt(v[0], v[1], v[2], v[3]) : cmap_ee(v[0], v[1], v[2], v[3]);
    bool valid = kind == 0 ? valid_pt(c.beta, c.gamma) : valid_ee(c.beta, c.gamma);
    double l = signed_distance(c.e0, c.e1p, c.e2p);
    // _mask_valid: PT gh uses l^2 (pt_ipc.py:20-22), EE gh |e2p|^2 (ee_ipc.py:23)
    double s = kind == 0 ? l * l : dot(c.e2p, c.e2p);
    if (!(s < d2hat && valid)) return false;
    double b1 = barrier_d1(s, d2hat, kappa);
    double b2 = barrier_d2(s, d2hat, kappa);
    // psd/hl.py:156-195 eigensystem of d2(l^2)/dc2
    double e0pn = dot(c.e0, c.e0), e1pn = dot(c.e1p, c.e1p), e2pn = dot(c.e2p, c.e2p);
    double f12 = sqrt(1.0 + 4.0 * (e1pn / e2pn));
    double f02 = sqrt(1.0 + 4.0 * (e0pn / e2pn));
    double lam[4];
    lam[0] = -l / (2.0 * e1pn) * (1.0 + f12);
    lam[1] = -l / (2.0 * e1pn) * (1.0 - f12);
    lam[2] = -l / (2.0 * e0pn) * (1.0 + f02);
    lam[3] = -l / (2.0 * e0pn) * (1.0 - f02);
    const double lt = l / e2pn;
    // dcdx (psd/vf.py:44-72, psd/ee.py:135-163), row-major 3x4
    double D[12];
    if (kind == 0) {
        D[0] = 0.0; D[1] = 1.0; D[2] = -1.0; D[3] = 0.0;
        D[4] = 0.0; D[5] = -c.alpha; D[6] = c.alpha - 1.0; D[7] = 1.0;
        D[8] = 1.0; D[9] = -c.beta; D[10] = c.beta + c.gamma - 1.0; D[11] = -c.gamma;
    } else {
        D[0] = -1.0; D[1] = 1.0; D[2] = 0.0; D[3] = 0.0;
        D[4] = c.alpha; D[5] = -c.alpha; D[6] = -1.0; D[7] = 1.0;
        D[8] = c.beta - 1.0; D[9] = -c.beta; D[10] = c.gamma + 1.0; D[11] = -c.gamma;
    }
    const int nI = kind == 0 ? 1 : 2;  // contact vertices on body I
    const d3 gl2 = (c.e2p * l) / e2pn;  // psd/hl.py:37-39
    const d3 q4 = (gl2 * 2.0) * l;
    for (int k = 0; k < 5; ++k) {
        d3 qb[3];
        double L;
        if (k < 4) {
            double om = lam[k] / (lam[k] - lt);
            L = lam[k] * 2.0 * l;
            // pt_ipc.py:266-269 (EE unclamped, ee_ipc.py:33-36); the psd mode projects the
            // whole contact-space Hessian below and starts from the unclamped eigenvalue
            if (kind == 0 && !psd) L = L > 0.0 ? L : 0.0;
            if (k < 2) {
                qb[0] = mk(0.0, 0.0, 0.0);
                qb[1] = c.e2p;
                qb[2] = c.e1p * om;
            } else {
                qb[0] = c.e2p;
                qb[1] = mk(0.0, 0.0, 0.0);
                qb[2] = c.e0 * om;
            }
        } else {
            L = 2.0;
            qb[0] = mk(0.0, 0.0, 0.0);
            qb[1] = mk(0.0, 0.0, 0.0);
            qb[2] = q4;
        }
        double qn = dot(qb[0], qb[0]) + dot(qb[1], qb[1]) + dot(qb[2], qb[2]);
        if (psd) {
            // contact-space eigenvalue of b' Hl + b'' q4 q4^T, clamped at zero
            double mu = k < 4 ? b1 * L : b1 * L + b2 * qn;
            out[kCoefOff + k] = (mu > 0.0 ? mu : 0.0) * (1.0 / qn);
        } else {
            double w = L * (1.0 / qn);
            out[kCoefOff + k] = k < 4 ? b1 * w : b1 * w + b2;
        }
        // u = D~^T q (12, per contact vertex), then V = J^T u per body
        double V[24];
#pragma unroll
        for (int i = 0; i < 24; ++i) V[i] = 0.0;
#pragma unroll
        for (int vtx = 0; vtx < 4; ++vtx) {
            d3 u = (qb[0] * D[0 * 4 + vtx] + qb[1] * D[1 * 4 + vtx]) + qb[2] * D[2 * 4 + vtx];
            const double Jr[4] = {1.0, r[vtx].x, r[vtx].y, r[vtx].z};
            const int base = vtx < nI ? 0 : 12;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                V[base + 3 * a + 0] += Jr[a] * u.x;
                V[base + 3 * a + 1] += Jr[a] * u.y;
                V[base + 3 * a + 2] += Jr[a] * u.z;
            }
        }
#pragma unroll
        for (int i = 0; i < 24; ++i) out[kVOff + 24 * k + i] = V[i];
    }
    out[kB1Off] = b1;
    return true;
}

// The dense 24x24 block J^T Hx J and gradient gscale J^T gx of one contact (out[0..575]
// row-major, bodies [I | J]; out[576..599]): vertex t of body half h(t) maps through
// [1, r_t^T] (x) I3; PT has one vertex on body I, EE two.
constexpr int kDenseStride = 600;
constexpr int kDenseGOff = 576;
__device__ void dense_body_map(int kind, const d3 r[4], const double gx[12], double gscale, const double Hx[144],
                               double* __restrict__ out) {
    const int nI = kind == 0 ? 1 : 2;
    for (int i = 0; i < 24; ++i) {
        const int hi = i / 12, ai = (i % 12) / 3, ki = i % 3;
        double gacc = 0.0;
        for (int t = 0; t < 4; ++t) {
            if ((t < nI ? 0 : 1) != hi) continue;
            const double Ji = ai == 0 ? 1.0 : comp(r[t], ai - 1);
            gacc += Ji * gx[3 * t + ki];
        }
        out[kDenseGOff + i] = gscale * gacc;
        for (int j = 0; j < 24; ++j) {
            const int hj = j / 12, aj = (j % 12) / 3, kj = j % 3;
            double acc = 0.0;
            for (int t = 0; t < 4; ++t) {
                if ((t < nI ? 0 : 1) != hi) continue;
                const double Ji = ai == 0 ? 1.0 : comp(r[t], ai - 1);
                for (int q = 0; q < 4; ++q) {
                    if ((q < nI ? 0 : 1) != hj) continue;
                    const double Jj = aj == 0 ? 1.0 : comp(r[q], aj - 1);
                    acc += Ji * Jj * Hx[12 * (3 * t + ki) + 3 * q + kj];
                }
            }
            out[24 * i + j] = acc;
        }
    }
}

// Exact mode (psd = 2, beyond the reference): the dense 24x24 block
// J^T (b' d2s/dx2 + b'' ds/dx ds/dx^T) J and the gradient b' J^T ds/dx of one
// contact, written as out[0..575] (row-major, bodies [I | J]) and out[576..599].
__device__ bool contact_terms_exact(int kind, const d3 v[4], const d3 r[4], double d2hat, double kappa,
                                    double* __restrict__ out) {
    CMap c = kind == 0 ? cmap_pt(v[0], v[1], v[2], v[3]) : cmap_ee(v[0], v[1], v[2], v[3]);
    bool valid = kind == 0 ? valid_pt(c.beta, c.gamma) : valid_ee(c.beta, c.gamma);
    double l = signed_distance(c.e0, c.e1p, c.e2p);
    double s = kind == 0 ? l * l : dot(c.e2p, c.e2p);
    if (!(s < d2hat && valid)) return false;
    const double b1 = barrier_d1(s, d2hat, kappa);
    const double b2 = barrier_d2(s, d2hat, kappa);
    // (u, v, w) = L x over the contact's vertices (the edge vectors of cmap_pt / cmap_ee)
    const double Lpt[12] = {0, 1, -1, 0, 0, 0, -1, 1, 1, 0, -1, 0};
    const double Lee[12] = {-1, 1, 0, 0, 0, 0, -1, 1, -1, 0, 1, 0};
    const double* L = kind == 0 ? Lpt : Lee;
    const d3 u = kind == 0 ? v[1] - v[2] : v[1] - v[0];
    const d3 vv = v[3] - v[2];
    const d3 w = kind == 0 ? v[0] - v[2] : v[2] - v[0];
    double ge[9], He[81];
    plane_dist2_derivs(u, vv, w, ge, He);
    double gx[12], Hx[144];
    for (int a = 0; a < 4; ++a)
        for (int k = 0; k < 3; ++k) {
            double acc = 0.0;
            for (int e = 0; e < 3; ++e) acc += L[4 * e + a] * ge[3 * e + k];
            gx[3 * a + k] = acc;
        }
    for (int a = 0; a < 4; ++a)
        for (int k = 0; k < 3; ++k)
            for (int b = 0; b < 4; ++b)
                for (int m = 0; m < 3; ++m) {
                    double acc = 0.0;
                    for (int e = 0; e < 3; ++e)
                        for (int f = 0; f < 3; ++f) acc += L[4 * e + a] * L[4 * f + b] * He[9 * (3 * e + k) + 3 * f + m];
                    Hx[12 * (3 * a + k) + 3 * b + m] = acc * b1 + (gx[3 * a + k] * gx[3 * b + m]) * b2;
                }
    dense_body_map(kind, r, gx, b1, Hx, out);
    return true;
}

// Full-IPC mode (psd = 3, beyond the reference): the barrier of the squared distance to the
// closest feature pair (no validity gate), times the edge-edge mollifier for EE; the 12x12
// Hessian in the points' coordinates projected onto the PSD cone, then mapped to the two
// bodies like contact_terms_exact (same storage).
__device__ bool contact_terms_full(int kind, const d3 v[4], const d3 r[4], double d2hat, double kappa,
                                   double* __restrict__ out) {
    const Feat f = kind == 0 ? pt_feature(v) : ee_feature(v);
    double gs[12], Hx[144];
    const double s = feature_dist2(kind, f, v, gs, Hx);
    if (!(s < d2hat)) return false;
    const double b0 = barrier(s, d2hat, kappa), b1 = barrier_d1(s, d2hat, kappa), b2 = barrier_d2(s, d2hat, kappa);
    // mollifier m(c), c = |u x v|^2, u = x1 - x0, v = x3 - x2 (EE below eps only)
    double m = 1.0, m1 = 0.0, m2 = 0.0, gc6[6], Hc6[36];
    const double eps = kind == 1 ? ee_mollifier_eps(r) : 0.0;
    const double c = kind == 1 ? cross2_derivs(v[1] - v[0], v[3] - v[2], gc6, Hc6) : 0.0;
    const bool moll = kind == 1 && c < eps;
    if (moll) {
        m = mollifier(c, eps);
        m1 = (2.0 - 2.0 * c / eps) / eps;
        m2 = -2.0 / (eps * eps);
    }
    // point coefficients of u (points 0, 1) and v (points 2, 3)
    const double Lc[2][4] = {{-1.0, 1.0, 0.0, 0.0}, {0.0, 0.0, -1.0, 1.0}};
    double gx[12], gc[12];
    for (int a = 0; a < 4; ++a)
        for (int k = 0; k < 3; ++k) {
            gc[3 * a + k] = moll ? Lc[0][a] * gc6[k] + Lc[1][a] * gc6[3 + k] : 0.0;
            gx[3 * a + k] = m * b1 * gs[3 * a + k] + b0 * m1 * gc[3 * a + k];
        }
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 12; ++j) {
            double h = m * (b1 * Hx[12 * i + j] + b2 * (gs[i] * gs[j]));
            if (moll) {
                const int a = i / 3, k = i % 3, bq = j / 3, mm = j % 3;
                double hc = 0.0;
                for (int e = 0; e < 2; ++e)
                    for (int f2 = 0; f2 < 2; ++f2) hc += Lc[e][a] * Lc[f2][bq] * Hc6[6 * (3 * e + k) + 3 * f2 + mm];
                h += b0 * (m2 * (gc[i] * gc[j]) + m1 * hc) + b1 * m1 * (gs[i] * gc[j] + gc[i] * gs[j]);
            }
            Hx[12 * i + j] = h;
        }
    psd_project12(Hx);
    dense_body_map(kind, r, gx, 1.0, Hx, out);
    return true;
}

// stage A: active flags over [PT | EE], plus ordering checks.
// dn (nullable): the list sizes on the device, [nij, nij, npt, nee, nvg] (the arguments are
// then upper bounds that only size the launch)
__global__ __launch_bounds__(kBlock) void classify_kernel(AbdScene sc, AbdParams P, const double* __restrict__ x,
                                                           const int32_t* __restrict__ pt, int64_t npt,
                                                           const int32_t* __restrict__ ee, int64_t nee,
                                                           const int32_t* __restrict__ vg, int64_t nvg,
                                                           const int64_t* __restrict__ dn,
                                                           uint8_t* __restrict__ flag, int* __restrict__ err,
                                                           int64_t* __restrict__ bcnt) {
    const SceneView S = view_of(sc);
    if (dn) {
        npt = dn[2];
        nee = dn[3];
        nvg = dn[4];
    }
    const int64_t n = npt + nee;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n + nvg; i += (int64_t)gridDim.x * kBlock) {
        bool act = false;
        if (i >= n) {
            int64_t k = i - n;
            if (k > 0 && vg[2 * k] < vg[2 * (k - 1)]) *err = 1;
        } else {
        const bool is_ee = i >= npt;
        const int64_t k = is_ee ? i - npt : i;
        const int32_t* t5 = (is_ee ? ee : pt) + 5 * k;
        if (k > 0 && t5[2] < t5[2 - 5]) *err = 1;
        int64_t vi[4];
        d3 v[4];
        if (is_ee)
            fetch_ee_idx(S, t5, vi);
        else
            fetch_pt_idx(S, t5, vi);
        load4(x, vi, v);
        if (P.psd == 3) {  // full IPC: the closest feature pair, no validity gate
            const Feat f = is_ee ? ee_feature(v) : pt_feature(v);
            act = (is_ee ? P.ee : P.pt) && feature_dist2_value(is_ee ? 1 : 0, f, v) < P.d2hat;
        } else if (is_ee) {
            CMap c = cmap_ee(v[0], v[1], v[2], v[3]);
            act = P.ee && dot(c.e2p, c.e2p) < P.d2hat && valid_ee(c.beta, c.gamma);
        } else {
            CMap c = cmap_pt(v[0], v[1], v[2], v[3]);
            double l = signed_distance(c.e0, c.e1p, c.e2p);
            act = P.pt && l * l < P.d2hat && valid_pt(c.beta, c.gamma);
        }
        flag[i] = act ? 1 : 0;
        }
        if (bcnt) {
            // stage B's count: the wave's active candidates (its 64 consecutive ones lie in one
            // compaction block) added to the block's count; lane 0 holds the wave's first one
            const uint64_t m = __ballot(act);
            if (lane_id() == 0 && m != 0)
                atomicAdd(reinterpret_cast<unsigned long long*>(bcnt + i / ((int64_t)kBlock * kCItems)),
                          (unsigned long long)__popcll(m));
        }
    }
}

// stage B: ordered compaction of the flags (count / write passes).
template <bool WRITE>
__global__ __launch_bounds__(kBlock) void compact_flags_kernel(const uint8_t* __restrict__ flag, int64_t n,
                                                                const int64_t* __restrict__ dn,
                                                                int64_t* __restrict__ bcnt,
                                                                const int64_t* __restrict__ boff,
                                                                int32_t* __restrict__ out) {
    __shared__ int shw[kBlock / kWave];
    if (dn) n = dn[2] + dn[3];
    int64_t base = (int64_t)blockIdx.x * kBlock * kCItems;
    int64_t run = WRITE ? boff[blockIdx.x] : 0;
    for (int it = 0; it < kCItems; ++it) {
        int64_t i = base + (int64_t)it * kBlock + threadIdx.x;
        bool pred = i < n && flag[i];
        int tot;
        int r = block_rank(pred, shw, &tot);
        if (WRITE && pred) out[run + r] = (int32_t)i;
        run += tot;
    }
    if (!WRITE && threadIdx.x == 0) bcnt[blockIdx.x] = run;
}

// sort key of candidate i of [PT | EE]: (type << 32) | body pair
__device__ __forceinline__ int64_t active_key(const int32_t* __restrict__ pt, int64_t npt,
                                              const int32_t* __restrict__ ee, int64_t i) {
    const bool is_ee = i >= npt;
    const int32_t* t5 = is_ee ? ee + 5 * (i - npt) : pt + 5 * i;
    return ((int64_t)(is_ee ? 1 : 0) << 32) | (int64_t)t5[2];
}

// stage C: one thread per active contact.  MODE: 1 the contact-space terms (psd 0 and 1),
// 2 the exact Hessian, 3 full IPC; one kernel per mode, so that the contact-space path does
// not carry the dense modes' register and stack footprint.
template <int MODE>
__global__ __launch_bounds__(kBlock) void contact_kernel(AbdScene sc, AbdParams P, const double* __restrict__ x,
                                                          const int32_t* __restrict__ pt, int64_t npt,
                                                          const int32_t* __restrict__ ee,
                                                          const int32_t* __restrict__ act, int64_t nact,
                                                          const int64_t* __restrict__ dnact,
                                                          const int64_t* __restrict__ dn,
                                                          double* __restrict__ cbuf, int64_t* __restrict__ ckey,
                                                          uint8_t* __restrict__ swap, uint8_t* __restrict__ hl_pt,
                                                          uint8_t* __restrict__ hl_ee, uint8_t* __restrict__ pflag,
                                                          int64_t nij, int64_t* __restrict__ pseg) {
    const SceneView S = view_of(sc);
    if (dnact) nact = *dnact;
    if (dn) npt = dn[2];
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < nact; e += (int64_t)gridDim.x * kBlock) {
        const int64_t i = act[e];
        const bool is_ee = i >= npt;
        const int32_t* t5 = is_ee ? ee + 5 * (i - npt) : pt + 5 * i;
        int64_t vi[4];
        if (is_ee)
            fetch_ee_idx(S, t5, vi);
        else
            fetch_pt_idx(S, t5, vi);
        d3 v[4], r[4];
        load4(x, vi, v);
        load4(sc.x0, vi, r);
        if (MODE == 3)
            contact_terms_full(is_ee ? 1 : 0, v, r, P.d2hat, P.kappa, cbuf + e * kDenseStride);
        else if (MODE == 2)
            contact_terms_exact(is_ee ? 1 : 0, v, r, P.d2hat, P.kappa, cbuf + e * kDenseStride);
        else
            contact_terms(is_ee ? 1 : 0, v, r, P.d2hat, P.kappa, P.psd, cbuf + e * kContactStride);
        const int64_t key = ((int64_t)(is_ee ? 1 : 0) << 32) | (int64_t)t5[2];
        ckey[e] = key;
        pflag[t5[2]] = 1;  // the body pair has an active contact (every writer stores 1)
        // the run of this key in the sorted active list: [begin, end) of (type, pair) from
        // the neighbours' keys (recomputed here from the lists, not read from ckey)
        const int64_t kprev = e > 0 ? active_key(pt, npt, ee, act[e - 1]) : -1;
        const int64_t knext = e + 1 < nact ? active_key(pt, npt, ee, act[e + 1]) : -1;
        int64_t* ps = pseg + 2 * ((is_ee ? nij : 0) + t5[2]);
        if (kprev != key) ps[0] = e;
        if (knext != key) ps[1] = e + 1;
        swap[e] = t5[0] > t5[1] ? 1 : 0;  // body I is ij1 (PT second order)
        if (is_ee)
            hl_ee[t5[1]] = 1;
        else
            hl_pt[t5[1]] = 1;
    }
}

constexpr int kPairThreads = 576;  // one thread per entry of the 24x24 body-pair block
constexpr int kPairBufStride = 2 * 144 + 24;  // diag(ij0), diag(ij1), g(ij0), g(ij1)
constexpr int64_t kPairGrid = 2048;           // workgroups of stage D (grid-stride over pairs)

constexpr int kPairChunk = 64;  // contacts of one stage-D work item
constexpr int kPartStride = 600;  // one chunk's partial sums: the 24x24 block, then g (24)

// stage D1: the work items of each body pair: its PT run and its EE run cut into chunks of
// kPairChunk contacts (runs from stage C; empty runs are [0, 0)).  One workgroup scans the pairs'
// item counts into their first items coff and the total.
__device__ __forceinline__ int64_t pair_item_count(const int64_t* __restrict__ pseg, int64_t nij, int64_t ii) {
    const int64_t npt = pseg[2 * ii + 1] - pseg[2 * ii], nee = pseg[2 * (nij + ii) + 1] - pseg[2 * (nij + ii)];
    return (npt + kPairChunk - 1) / kPairChunk + (nee + kPairChunk - 1) / kPairChunk;
}
__global__ __launch_bounds__(kScanThreads) void pair_chunks_scan_kernel(int64_t nij, const int64_t* __restrict__ pseg,
                                                                        int64_t* __restrict__ coff,
                                                                        int64_t* __restrict__ total) {
    scan_block_exclusive_by([=](int64_t ii) { return pair_item_count(pseg, nij, ii); }, coff, nij, total, nullptr);
}

// stage D's items: pair ii owns the items [coff[ii], coff[ii] + cnt[ii]), its PT run's chunks
// then its EE run's; item i's descriptor is its first contact | its contact count << 56.
// The order the rank-one partial sums run them in (slot -> item, odesc / oitem) takes the full
// chunks first (counted in counters[0], from the front) and the others after them (counters[1],
// from the back): the longest sums start first.  Where an item lands in it changes no result.
// One wave per pair, its lanes striding over the pair's items.
constexpr int kItemShift = 56;
__global__ __launch_bounds__(kBlock) void pair_items_kernel(int64_t nij, const int64_t* __restrict__ coff,
                                                            const int64_t* __restrict__ pseg,
                                                            const int64_t* __restrict__ total,
                                                            int64_t* __restrict__ idesc, int* __restrict__ counters,
                                                            int64_t* __restrict__ odesc,
                                                            int64_t* __restrict__ oitem) {
    const int64_t ntot = *total;
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t nw = (int64_t)gridDim.x * (kBlock / kWave);
    for (int64_t ii = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / kWave; ii < nij; ii += nw) {
        const int64_t o = coff[ii];
        const int64_t p0 = pseg[2 * ii], p1 = pseg[2 * ii + 1];
        const int64_t e0 = pseg[2 * (nij + ii)], e1 = pseg[2 * (nij + ii) + 1];
        const int64_t npc = (p1 - p0 + kPairChunk - 1) / kPairChunk;
        const int64_t n = npc + (e1 - e0 + kPairChunk - 1) / kPairChunk;
        for (int64_t c = lane; c < n; c += kWave) {
            const bool pt = c < npc;
            const int64_t c0 = pt ? p0 + c * kPairChunk : e0 + (c - npc) * kPairChunk;
            const int64_t end = pt ? p1 : e1;
            const int64_t nc = end - c0 < kPairChunk ? end - c0 : kPairChunk;
            const int64_t d = c0 | (nc << kItemShift);
            idesc[o + c] = d;
            const bool full = nc == kPairChunk;
            const int64_t k = atomicAdd(counters + (full ? 0 : 1), 1);
            const int64_t slot = full ? k : ntot - 1 - k;
            odesc[slot] = d;
            oitem[slot] = o + c;
        }
    }
}

// stage D2: one work item per workgroup (grid-stride over the *total items): the sums, in
// contact order, of one chunk's contacts into part[item] (24x24 block in the pair's body
// order [ij0 | ij1], then g), for contacts stored whole by contact_terms_exact / _full.
__global__ __launch_bounds__(kPairThreads) void pair_partial_dense_kernel(const double* __restrict__ cbuf,
                                                                           const uint8_t* __restrict__ swap,
                                                                           const int64_t* __restrict__ idesc,
                                                                     const int64_t* __restrict__ total,
                                                                     double* __restrict__ part) {
    constexpr int kStride = kDenseStride;
    constexpr int kBatch = 4;
    __shared__ double sc[kBatch][kStride];
    __shared__ uint8_t ssw[kBatch];
    const int t = threadIdx.x;
    const int R = t / 24, C = t % 24;
    const int64_t nitem = *total;
    for (int64_t gc = blockIdx.x; gc < nitem; gc += gridDim.x) {
        const int64_t d = idesc[gc];
        const int64_t c0 = d & ((int64_t(1) << kItemShift) - 1);
        const int64_t c1 = c0 + (d >> kItemShift);
        double h = 0.0, gacc = 0.0;
        for (int64_t e0 = c0; e0 < c1; e0 += kBatch) {
            const int nbat = (int)((c1 - e0) < kBatch ? (c1 - e0) : kBatch);
            for (int q = t; q < nbat * kStride; q += kPairThreads)
                sc[q / kStride][q % kStride] = cbuf[(e0 + q / kStride) * kStride + q % kStride];
            if (t < nbat) ssw[t] = swap[e0 + t];
            __syncthreads();
            for (int bq = 0; bq < nbat; ++bq) {
                // body order [ij0 | ij1]: swap halves when body I is ij1
                const int sw = ssw[bq] ? 12 : 0;
                const int r = (R + sw) % 24, cc = (C + sw) % 24;
                h += sc[bq][24 * r + cc];
                if (t < 24) gacc += sc[bq][kDenseGOff + (t + sw) % 24];
            }
            __syncthreads();
        }
        double* pp = part + gc * kPartStride;
        pp[t] = h;
        if (t < 24) pp[576 + t] = gacc;
    }
}

// stage D2 for rank-one contacts (the stored form of contact_terms): the items of a
// workgroup (grid-stride) cut into batches of kRankBatch contacts, each staged into LDS
// (without the pad) from registers loaded while the previous batch was summed; one thread
// per entry of the upper triangle of the 24x24 block sums c_k (V_k[r] V_k[c]) over the
// item's contacts in order, and threads 0..23 the gradient; an entry of the lower triangle
// is the same sum (V_k[r] V_k[c] == V_k[c] V_k[r]), written from its mirror.
constexpr int kRankThreads = 320;                   // 300 upper-triangle entries, 5 waves
constexpr int kRankBatch = 32;                      // contacts staged at a time
constexpr int kStaged = 126;                        // doubles per staged contact: c[5], b', V[5][24]
constexpr int kRankPieces = (kRankBatch * 63 + kRankThreads - 1) / kRankThreads;  // 16-byte loads a thread
static_assert(kContactStride == 128 && kCoefOff == 0 && kB1Off == 5 && kVOff == 8, "staged layout");
static_assert(kPairChunk <= kWave, "one swap bit a lane");
typedef double dbl2 __attribute__((ext_vector_type(2)));  // a 16-byte piece (kept in registers)
// read-only data reached through the constant address space: uniform loads become scalar ones
typedef __attribute__((address_space(4))) const double kconst_double;

// A thread's pieces of the batch of n contacts from src (63 pieces of 16 bytes a contact:
// c[0..5] in pieces 0-2, V from piece 4 of the stored row), piece U onwards.  A thread past
// the batch reloads its last piece (stored to a spare slot), so the loads are unconditional
// and issue together; the recursion keeps every index a constant (the pieces in registers).
template <int U>
__device__ __forceinline__ void rank_batch_load(const dbl2* __restrict__ src, int np, int t,
                                                dbl2 (&piece)[kRankPieces]) {
    if constexpr (U < kRankPieces) {
        const int q = min(t + U * kRankThreads, np - 1), e = q / 63, pc = q - 63 * e;
        piece[U] = src[e * (kContactStride / 2) + pc + (pc >= 3 ? 1 : 0)];
        rank_batch_load<U + 1>(src, np, t, piece);
    }
}
// The pieces into the staged batch (the pieces past it into the spare slot).
template <int U>
__device__ __forceinline__ void rank_batch_store(dbl2* dst, int np, int t, const dbl2 (&piece)[kRankPieces]) {
    if constexpr (U < kRankPieces) {
        const int q = t + U * kRankThreads, e = q / 63, pc = q - 63 * e;
        dst[q < np ? e * (kStaged / 2) + pc : kRankBatch * kStaged / 2] = piece[U];
        rank_batch_store<U + 1>(dst, np, t, piece);
    }
}

__global__ __launch_bounds__(kRankThreads, 5) void pair_partial_rank_kernel(const double* __restrict__ cbuf,
                                                                         const uint8_t* __restrict__ swap,
                                                                         const int64_t* __restrict__ odesc,
                                                                         const int64_t* __restrict__ oitem,
                                                                         const int64_t* __restrict__ total,
                                                                         double* __restrict__ part) {
    // the staged batch, then a 16-byte slot the pieces past the batch are stored to
    __shared__ __align__(16) double st[kRankBatch * kStaged + 2];
    const int t = threadIdx.x, lane = t % kWave;
    // this thread's entry (r, c), r <= c, of the upper triangle (rows of 24 - r entries)
    int r = 0, j = t;
    while (r < 24 && j >= 24 - r) {
        j -= 24 - r;
        ++r;
    }
    const bool ent = r < 24;
    const int c = ent ? r + j : 0;
    if (!ent) r = 0;
    // V indices in the stored order, without and with the halves swapped (body I is ij1)
    const int tg = t % 24;
    const int rA = r, rB = (r + 12) % 24, cA = c, cB = (c + 12) % 24, gA = tg, gB = (tg + 12) % 24;
    const int64_t nitem = *total;
    int64_t gc = blockIdx.x;  // the execution slot (its item: oitem[gc])
    if (gc >= nitem) return;  // uniform
    const int64_t kMask = (int64_t(1) << kItemShift) - 1;
    int64_t c0 = odesc[gc] & kMask;
    int nc = (int)(odesc[gc] >> kItemShift), b0 = 0;
    dbl2 piece[kRankPieces];
    rank_batch_load<0>(reinterpret_cast<const dbl2*>(cbuf + c0 * kContactStride), 63 * min(nc, kRankBatch), t,
                       piece);
    // the item's swap flags, one bit a contact, in every wave
    uint64_t swm = __ballot(lane < nc && swap[c0 + lane] != 0);
    double h = 0.0, gacc = 0.0;
    for (;;) {
        const int nbat = min(kRankBatch, nc - b0);
        __syncthreads();  // the previous batch is summed: its LDS is free
        rank_batch_store<0>(reinterpret_cast<dbl2*>(st), 63 * nbat, t, piece);
        __syncthreads();
        // the next batch: the rest of this item, else the workgroup's next item; its loads
        // are in flight while this batch is summed
        int64_t gn = gc, cn = c0;
        int nn = nc, bn = b0 + kRankBatch;
        if (bn >= nc) {
            gn = gc + gridDim.x;
            bn = 0;
            if (gn < nitem) {
                const int64_t dn_ = odesc[gn];
                cn = dn_ & kMask;
                nn = (int)(dn_ >> kItemShift);
            }
        }
        // (past the workgroup's last item this batch is loaded again, unused: the loads stay
        // unconditional and the pieces in registers)
        const bool more = gn < nitem;
        rank_batch_load<0>(reinterpret_cast<const dbl2*>(cbuf + (more ? cn + bn : c0 + b0) * kContactStride),
                           63 * (more ? min(nn - bn, kRankBatch) : nbat), t, piece);
        const uint8_t swn = more && gn != gc && lane < nn ? swap[cn + lane] : 0;
        // every thread also runs the gradient chain of entry t % 24 (kept by threads 0..23),
        // so the loop has no branch and the loads of unrolled contacts issue together
#pragma unroll 4
        for (int e = 0; e < nbat; ++e) {
            const double* S = st + e * kStaged;
            // the coefficients are the same for every lane: scalar loads from the stored row
            const kconst_double* K = (const kconst_double*)cbuf + (c0 + b0 + e) * kContactStride;
            // body order [ij0 | ij1]: halves swapped when body I is ij1
            const bool sw = (swm >> (b0 + e)) & 1;
            const double* Vr = S + 6 + (sw ? rB : rA);
            const double* Vc = S + 6 + (sw ? cB : cA);
#pragma unroll
            for (int k = 0; k < 5; ++k) h += K[k] * (Vr[24 * k] * Vc[24 * k]);
            gacc += K[5] * S[6 + 96 + (sw ? gB : gA)];
        }
        if (gn != gc) {  // uniform: the item is summed
            double* pp = part + oitem[gc] * kPartStride;
            if (ent) {
                pp[24 * r + c] = h;
                pp[24 * c + r] = h;
            }
            if (t < 24) pp[576 + t] = gacc;
            h = gacc = 0.0;
            if (gn >= nitem) break;
            swm = __ballot(swn != 0);
        }
        gc = gn;
        c0 = cn;
        nc = nn;
        b0 = bn;
    }
}

// stage D3: per body pair with active contacts (grid-stride): its chunks' partial sums
// (items [coff[ii], coff[ii + 1]); coff has nij + 1 entries, the last the total) added in
// order (PT chunks, then EE), diagonal parts to the pair buffer (stage E adds
// them per body), the off-diagonal block to the matrix.  nij: pairs of the buffers (an upper
// bound when dn gives the count of the matrix's pairs, dn[0], which places its lower slots).
__global__ __launch_bounds__(kPairThreads) void pair_reduce_kernel(int nb, int64_t nij,
                                                                    const int64_t* __restrict__ dn,
                                                                    const uint8_t* __restrict__ pflag,
                                                                    const int64_t* __restrict__ coff,
                                                                    const double* __restrict__ part,
                                                                    double* __restrict__ pbuf,
                                                                    double* __restrict__ blocks) {
    const int t = threadIdx.x;
    const int R = t / 24, C = t % 24;
    const int64_t nij_m = dn ? dn[0] : nij;
    for (int64_t ii = blockIdx.x; ii < nij; ii += gridDim.x) {
        if (!pflag[ii]) continue;  // uniform over the workgroup
        const int64_t c0 = coff[ii], nc = coff[ii + 1] - c0;  // coff[nij] holds the items' total
        double h = 0.0, gacc = 0.0;
        for (int64_t c = c0; c < c0 + nc; ++c) {
            h += part[c * kPartStride + t];
            if (t < 24) gacc += part[c * kPartStride + 576 + t];
        }
        double* pb = pbuf + ii * kPairBufStride;
        if (R < 12 && C < 12) pb[R * 12 + C] = h;
        if (R >= 12 && C >= 12) pb[144 + (R - 12) * 12 + (C - 12)] = h;
        if (t < 24) pb[288 + t] = gacc;
        // off-diagonal: upper slot nb + ii holds rows ij0 x cols ij1, lower slot its transpose
        if (R < 12 && C >= 12) {
            const int c2 = C - 12;
            const int bi = R / 3, rr = R % 3, bj = c2 / 3, cc = c2 % 3;
            const int64_t up = nb + ii, lo = nb + nij_m + ii;
            blocks[9 * (16 * up + bi + 4 * bj) + 3 * rr + cc] += h;
            blocks[9 * (16 * lo + bj + 4 * bi) + 3 * cc + rr] += h;
        }
    }
}

// The first vertex-ground candidate of body b and of body b + 1 in a list grouped by body
// ascending, found by the whole workgroup at once: each half probes 128 evenly spaced
// entries for its key and the count of entries below the key narrows the range 128-fold,
// so a list of n entries takes about log_128(n) + 1 dependent loads instead of log_2(n).
// Every thread runs the same trip count (the ranges are computed from block-wide counts).
__device__ __forceinline__ void body_vg_range(const int32_t* __restrict__ vg, int64_t n, int b, int64_t& k0,
                                              int64_t& k1) {
    static_assert(kBlock == 256, "two halves of 128 probes");
    __shared__ int s_cnt[2][kBlock / kWave];
    const int t = threadIdx.x, h = t >> 7, j = t & 127, w = t / kWave;
    const int key = b + h;
    int64_t lo[2] = {0, 0}, len[2] = {n, n};
    for (int it = 0; len[0] > 0 || len[1] > 0; ++it) {
        const int64_t L = len[h];
        bool below = false;
        if (L > 128)
            below = vg[2 * (lo[h] + (int64_t)j * L / 128)] < key;
        else if (j < L)
            below = vg[2 * (lo[h] + j)] < key;
        const uint64_t m = __ballot(below);
        if ((t & (kWave - 1)) == 0) s_cnt[it & 1][w] = __popcll(m);
        __syncthreads();
        for (int q = 0; q < 2; ++q) {
            const int c = s_cnt[it & 1][2 * q] + s_cnt[it & 1][2 * q + 1];
            const int64_t Lq = len[q];
            if (Lq <= 128) {
                lo[q] += c;
                len[q] = 0;
            } else {
                // entries before probe c-1 (inclusive) are below the key, probe c is not
                const int64_t a = c > 0 ? lo[q] + (int64_t)(c - 1) * Lq / 128 + 1 : lo[q];
                const int64_t e = c < 128 ? lo[q] + (int64_t)c * Lq / 128 : lo[q] + Lq;
                lo[q] = a;
                len[q] = e - a;
            }
        }
    }
    k0 = lo[0];
    k1 = lo[1];
}

// stage E: one workgroup per body (144 Hessian entries + 12 gradient entries).
__global__ __launch_bounds__(kBlock) void body_finalize_kernel(AbdScene sc, AbdParams P,
                                                                const double* __restrict__ x, int64_t nij,
                                                                const int64_t* __restrict__ dn,
                                                                const int32_t* __restrict__ row_ptr,
                                                                const int32_t* __restrict__ slots,
                                                                const uint8_t* __restrict__ pflag,
                                                                const double* __restrict__ pbuf,
                                                                const int32_t* __restrict__ vg, int64_t nvg,
                                                                double* __restrict__ g, double* __restrict__ blocks,
                                                                const uint8_t* __restrict__ hl_pt,
                                                                const uint8_t* __restrict__ hl_ee,
                                                                uint8_t* __restrict__ highlight) {
    const int b = blockIdx.x;
    const int t = threadIdx.x;
    const int nb = sc.n_bodies;
    if (dn) {
        nij = dn[0];
        nvg = dn[4];
    }
    // threads 0..143: the entries of the diagonal block; 144..155: the gradient; the rest
    // only help to evaluate the vertex-ground terms
    const bool act = t < 156;
    const bool isg = t >= 144;
    const int R = isg ? t - 144 : t / 12, C = isg ? 0 : t % 12;
    double acc = 0.0;
    if (act && nij > 0) {
        for (int e = row_ptr[b]; e < row_ptr[b + 1]; ++e) {
            const int s = slots[e];
            if (s < nb) continue;
            const int64_t ii = s < nb + nij ? s - nb : s - nb - nij;
            if (!pflag[ii]) continue;  // a pair without active contacts adds nothing
            const int side = s < nb + nij ? 0 : 1;  // upper slot in row b: b is ij0
            const double* pb = pbuf + ii * kPairBufStride;
            acc += isg ? pb[288 + 12 * side + R] : pb[144 * side + R * 12 + C];
        }
    }
    if (P.vg && nvg > 0) {
        // ipcsk/vg_ipc.py:5-34 (no 2 b' grad d grad d^T term, D11).  The barrier derivatives
        // of each contact are evaluated once per tile into LDS, then every entry adds the
        // tile's terms in contact order.
        __shared__ double s_d[kBlock], s_db[kBlock], s_d2b[kBlock], s_vt[kBlock][3];
        int64_t k0, k1;
        body_vg_range(vg, nvg, b, k0, k1);
        const int ia = R / 3, ra = R % 3, ja = C / 3, ca = C % 3;
        for (int64_t base = k0; base < k1; base += kBlock) {
            const int nt = (int)(k1 - base < kBlock ? k1 - base : kBlock);
            if (t < nt) {
                const int64_t vid = sc.voff[b] + vg[2 * (base + t) + 1];
                const double d = x[3 * vid + 1] - P.ground;
                s_d[t] = d;
                s_d2b[t] = barrier_d2(d * d, P.d2hat, P.kappa);
                s_db[t] = barrier_d1(d * d, P.d2hat, P.kappa);
                s_vt[t][0] = sc.x0[3 * vid];
                s_vt[t][1] = sc.x0[3 * vid + 1];
                s_vt[t][2] = sc.x0[3 * vid + 2];
            }
            __syncthreads();
            if (act)
                for (int q = 0; q < nt; ++q) {
                    const double d = s_d[q], d2b = s_d2b[q], db = s_db[q];
                    const double th_i = ia == 0 ? 1.0 : s_vt[q][ia - 1];
                    if (isg) {
                        const double nr = ra == 1 ? 1.0 : 0.0;
                        acc += (((d * 2.0) * nr) * th_i) * db;
                    } else {
                        const double th_j = ja == 0 ? 1.0 : s_vt[q][ja - 1];
                        double scl = th_i * th_j * d2b * 4.0 * d * d;
                        if (P.psd == 2) scl += th_i * th_j * 2.0 * db;  // exact mode: the b' term
                        if (P.psd == 3) {  // full IPC: d2b/dy2 = 4 d^2 b'' + 2 b' clamped at 0
                            const double cy = d2b * 4.0 * d * d + 2.0 * db;
                            scl = th_i * th_j * (cy > 0.0 ? cy : 0.0);
                        }
                        const double nn = (ra == 1 ? 1.0 : 0.0) * (ca == 1 ? 1.0 : 0.0);
                        acc += nn * scl;
                    }
                }
            __syncthreads();
        }
    }
    if (!act) return;
    if (isg) {
        g[3 * (4 * (int64_t)b + R / 3) + R % 3] += acc;
    } else {
        const int ia = R / 3, ra = R % 3, ja = C / 3, ca = C % 3;
        blocks[9 * (16 * (int64_t)b + ia + 4 * ja) + 3 * ra + ca] += acc;
    }
    if (t == 0 && highlight) highlight[b] = (uint8_t)((hl_pt[b] ? 1 : 0) | (hl_ee[b] ? 2 : 0));
}

// the gh stage's flags and counters: the ordering flag, n8 bytes at a8, n64 int64 at a64
__global__ void gh_zero_kernel(int* __restrict__ err, uint8_t* __restrict__ a8, int64_t n8,
                               int64_t* __restrict__ a64, int64_t n64, int64_t* __restrict__ b64, int64_t nb64) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nb64) b64[i] = 0;
    if (i == 0) {
        err[0] = 0;
        err[1] = err[2] = 0;  // stage D's counters of full and partial chunks
    }
    if (i < n8) a8[i] = 0;
    if (i < n64) a64[i] = 0;
}

// Per-contact entry (tests, microbenchmarks): dense 24x24 and gradient per contact.
__global__ __launch_bounds__(kBlock) void pair_gh_kernel(int kind, AbdParams P, const double* __restrict__ x4,
                                                          const double* __restrict__ r4, int64_t n,
                                                          uint8_t* __restrict__ active, double* __restrict__ g24,
                                                          double* __restrict__ H24, double* __restrict__ scratch) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        d3 v[4], r[4];
        for (int k = 0; k < 4; ++k) {
            v[k] = ld3(x4 + 12 * i + 3 * k);
            r[k] = ld3(r4 + 12 * i + 3 * k);
        }
        if (P.psd >= 2) {
            double* buf = scratch + i * kDenseStride;
            bool act = P.psd == 3 ? contact_terms_full(kind, v, r, P.d2hat, P.kappa, buf)
                                  : contact_terms_exact(kind, v, r, P.d2hat, P.kappa, buf);
            active[i] = act ? 1 : 0;
            for (int a = 0; a < 24; ++a) {
                g24[24 * i + a] = act ? buf[kDenseGOff + a] : 0.0;
                for (int c = 0; c < 24; ++c) H24[576 * i + 24 * a + c] = act ? buf[24 * a + c] : 0.0;
            }
            continue;
        }
        double* buf = scratch + i * kContactStride;
        bool act = contact_terms(kind, v, r, P.d2hat, P.kappa, P.psd, buf);
        active[i] = act ? 1 : 0;
        for (int a = 0; a < 24; ++a) {
            g24[24 * i + a] = act ? buf[kB1Off] * buf[kVOff + 24 * 4 + a] : 0.0;
            for (int c = 0; c < 24; ++c) {
                double h = 0.0;
                if (act)
                    for (int k = 0; k < 5; ++k) {
                        const double* V = buf + kVOff + 24 * k;
                        h += buf[kCoefOff + k] * (V[a] * V[c]);
                    }
                H24[576 * i + 24 * a + c] = h;
            }
        }
    }
}

// -------------------------------------------------------------- inertia
// orthogonal_energy.py:99-112 (overwrites g)
__global__ void inertia_gradient_kernel(int nb, AbdParams P, const double* __restrict__ mass,
                                        const double* __restrict__ I0v, const double* __restrict__ A,
                                        const double* __restrict__ p, const double* __restrict__ A0,
                                        const double* __restrict__ p0, const double* __restrict__ Adot,
                                        const double* __restrict__ pdot, double* __restrict__ g) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const double* Ab = A + 9 * b;
    double At[9], pt[3];
    tilde(P, A0 + 9 * b, Adot + 9 * b, p0 + 3 * b, pdot + 3 * b, At, pt);
    const double m = mass ? mass[b] : P.mass, i0 = I0v ? I0v[b] : P.I0;
    for (int k = 0; k < 3; ++k) g[3 * (4 * b) + k] = 0.0 + (p[3 * b + k] - pt[k]) * m;
    for (int ii = 1; ii < 4; ++ii) {
        const int i = ii - 1;
        double gr[3] = {-Ab[3 * i], -Ab[3 * i + 1], -Ab[3 * i + 2]};
        for (int j = 0; j < 3; ++j) {
            double d = row_dot(Ab, i, j);
            for (int k = 0; k < 3; ++k) gr[k] += d * Ab[3 * j + k];
        }
        for (int k = 0; k < 3; ++k)
            g[3 * (4 * b + ii) + k] = P.dt * P.dt * (gr[k] * 4.0 * P.stiffness) + (Ab[3 * i + k] - At[3 * i + k]) * i0;
    }
}

// orthogonal_energy.py:63-96 (+= into the diagonal slots); one thread per (body, sub-block).
__global__ void inertia_hessian_kernel(int nb, AbdParams P, const double* __restrict__ mass,
                                       const double* __restrict__ I0v, const double* __restrict__ A,
                                       double* __restrict__ blocks) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 16 * nb) return;
    const int b = t / 16, e = t % 16, ii = e % 4, jj = e / 4;
    const double* Ab = A + 9 * b;
    const double m_ = mass ? mass[b] : P.mass, i0 = I0v ? I0v[b] : P.I0;
    const double m = ii == jj ? (ii == 0 ? m_ : i0) : 0.0;
    double dh[9] = {m, 0, 0, 0, m, 0, 0, 0, m};
    if (ii > 0 && jj > 0) {
        const int i = ii - 1, j = jj - 1;
        double h[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (i == j) {
            const double qq = row_dot(Ab, i, i) - 1.0;
            for (int k = 0; k < 3; ++k)
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) h[3 * r + c] += Ab[3 * k + r] * Ab[3 * k + c];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) h[3 * r + c] += Ab[3 * i + r] * Ab[3 * i + c] + (r == c ? qq : 0.0);
        } else {
            const double d = row_dot(Ab, j, i);
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) h[3 * r + c] = Ab[3 * j + r] * Ab[3 * i + c] + (r == c ? d : 0.0);
        }
        for (int k = 0; k < 9; ++k) dh[k] += h[k] * 4.0 * P.stiffness * P.dt * P.dt;
    }
    double* B = blocks + 9 * (16 * (int64_t)b + ii + jj * 4);
    for (int k = 0; k < 9; ++k) B[k] += dh[k];
}

// inertia_gradient_kernel and inertia_hessian_kernel in one launch: thread (body, sub-block e)
// adds sub-block e of the Hessian, and threads e < 4 also write gradient row e (the rows of
// inertia_gradient_kernel's loop, the same arithmetic).
__global__ void inertia_gh_kernel(int nb, AbdParams P, const double* __restrict__ mass,
                                  const double* __restrict__ I0v, const double* __restrict__ A,
                                  const double* __restrict__ p, const double* __restrict__ A0,
                                  const double* __restrict__ p0, const double* __restrict__ Adot,
                                  const double* __restrict__ pdot, double* __restrict__ g,
                                  double* __restrict__ blocks) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 16 * nb) return;
    const int b = t / 16, e = t % 16;
    if (e < 4) {
        const double* Ab = A + 9 * b;
        double At[9], pt[3];
        tilde(P, A0 + 9 * b, Adot + 9 * b, p0 + 3 * b, pdot + 3 * b, At, pt);
        const double m = mass ? mass[b] : P.mass, i0 = I0v ? I0v[b] : P.I0;
        if (e == 0) {
            for (int k = 0; k < 3; ++k) g[3 * (4 * b) + k] = 0.0 + (p[3 * b + k] - pt[k]) * m;
        } else {
            const int i = e - 1;
            double gr[3] = {-Ab[3 * i], -Ab[3 * i + 1], -Ab[3 * i + 2]};
            for (int j = 0; j < 3; ++j) {
                double d = row_dot(Ab, i, j);
                for (int k = 0; k < 3; ++k) gr[k] += d * Ab[3 * j + k];
            }
            for (int k = 0; k < 3; ++k)
                g[3 * (4 * b + e) + k] = P.dt * P.dt * (gr[k] * 4.0 * P.stiffness) + (Ab[3 * i + k] - At[3 * i + k]) * i0;
        }
    }
    const int ii = e % 4, jj = e / 4;
    const double* Ab = A + 9 * b;
    const double m_ = mass ? mass[b] : P.mass, i0 = I0v ? I0v[b] : P.I0;
    const double m = ii == jj ? (ii == 0 ? m_ : i0) : 0.0;
    double dh[9] = {m, 0, 0, 0, m, 0, 0, 0, m};
    if (ii > 0 && jj > 0) {
        const int i = ii - 1, j = jj - 1;
        double h[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (i == j) {
            const double qq = row_dot(Ab, i, i) - 1.0;
            for (int k = 0; k < 3; ++k)
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) h[3 * r + c] += Ab[3 * k + r] * Ab[3 * k + c];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) h[3 * r + c] += Ab[3 * i + r] * Ab[3 * i + c] + (r == c ? qq : 0.0);
        } else {
            const double d = row_dot(Ab, j, i);
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) h[3 * r + c] = Ab[3 * j + r] * Ab[3 * i + c] + (r == c ? d : 0.0);
        }
        for (int k = 0; k < 9; ++k) dh[k] += h[k] * 4.0 * P.stiffness * P.dt * P.dt;
    }
    double* B = blocks + 9 * (16 * (int64_t)b + ii + jj * 4);
    for (int k = 0; k < 9; ++k) B[k] += dh[k];
}

// The inertia gradient (overwriting g) and Hessian (added into the diagonal slots) in one launch.
int inertia_gh(hipStream_t s, const AbdScene* sc, const AbdParams* P, const double* A, const double* p,
               const double* A0, const double* p0, const double* Adot, const double* pdot, double* g, double* blocks) {
    const int nb = sc->n_bodies;
    if (!nb) return ABD_OK;
    hipLaunchKernelGGL(inertia_gh_kernel, dim3(grid_for(16 * (int64_t)nb)), dim3(kBlock), 0, s, nb, *P, sc->mass,
                       sc->I0, A, p, A0, p0, Adot, pdot, g, blocks);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

// The gh stage (stages A-E).  With L.dn the list sizes are on the device (L's counts bound
// them and size the launches and buffers), nothing is read back and the lists are taken
// as grouped (the drivers' own culling output); n_active_dev (nullable) then receives the
// active count on the device.  Without it the call reads the active count and the
// ordering flag back (n_active, host, nullable) and returns ABD_ERR_UNSORTED for lists
// that are not grouped.
int contact_gh(hipStream_t s, const AbdScene* sc, const AbdParams* P, const double* x, const GhLists& L,
               const int32_t* row_ptr, const int32_t* slots, double* g, double* blocks, uint8_t* highlight,
               int64_t* n_active, int64_t* n_active_dev) {
    const int nb = sc->n_bodies;
    if (nb == 0) return ABD_OK;
    const int32_t *pt = L.pt, *ee = L.ee, *vg = L.vg;
    const int64_t npt = L.npt, nee = L.nee, nij = L.nij;
    const int64_t nvg = P->vg ? L.nvg : 0;
    const int64_t* dn = L.dn;
    const int64_t n = npt + nee;
    int rc;
    // misc: err flag, highlight bytes [2 nb], active-pair flags [nij]
    uint8_t* misc;
    if ((rc = arena_get(kSlotMisc, 64 + 2 * (size_t)nb + (size_t)nij, (void**)&misc))) return rc;
    int* err = reinterpret_cast<int*>(misc);
    uint8_t* hl_pt = misc + 64;
    uint8_t* hl_ee = hl_pt + nb;
    uint8_t* pflag = hl_ee + nb;
    // the pair runs of stage C (empty unless written), zeroed with the flags
    int64_t* pseg = nullptr;  // [2][nij][2]
    if (nij > 0 && (rc = arena_get(kSlotPairSeg, sizeof(int64_t) * 4 * (size_t)nij, (void**)&pseg))) return rc;
    // stage B's per-block counts, added to by stage A (zeroed here with the flags)
    const int64_t nblk = n > 0 ? ceil_div(n, (int64_t)kBlock * kCItems) : 0;
    int64_t *bcnt = nullptr, *boff = nullptr;
    if (n > 0) {
        if ((rc = arena_get(kSlotScanA, sizeof(int64_t) * (nblk + 1), (void**)&bcnt))) return rc;
        if ((rc = arena_get(kSlotScanB, sizeof(int64_t) * (nblk + 1), (void**)&boff))) return rc;
    }
    {
        const int64_t n8 = 2 * (int64_t)nb + nij, n64 = 4 * nij;
        const int64_t nz = std::max(std::max(n8, n64), nblk);
        hipLaunchKernelGGL(gh_zero_kernel, dim3(grid_for(nz, kBlock, INT32_MAX)), dim3(kBlock), 0, s, err, hl_pt, n8,
                           pseg, n64, bcnt, nblk);
        ABD_CHECK_LAUNCH();
    }
    int64_t nact = 0;
    int32_t* act = nullptr;
    const int64_t* dev_nact = nullptr;
    if (n + nvg > 0) {
        uint8_t* flag;
        if ((rc = arena_get(kSlotFlags, (size_t)n + 1, (void**)&flag))) return rc;
        int ph = prof_start(kProfClassify, s, dn ? 1.0 : 129.0 * npt + 133.0 * nee + 8.0 * nvg);
        hipLaunchKernelGGL(classify_kernel, dim3(grid_for(n + nvg, kBlock, 8192)), dim3(kBlock), 0, s, *sc, *P, x,
                           pt, npt, ee, nee, vg, nvg, dn, flag, err, bcnt);
        ABD_CHECK_LAUNCH();
        prof_stop(ph, s);
        if ((rc = prof_count_bytes(ph, s, dn, 129.0, 133.0, 8.0, 0.0))) return rc;
        if (n > 0) {
            if ((rc = arena_get(kSlotList, sizeof(int32_t) * (size_t)n, (void**)&act))) return rc;
            if ((rc = scan_exclusive_i64(s, bcnt, boff, nblk, boff + nblk, n_active_dev))) return rc;
            hipLaunchKernelGGL(compact_flags_kernel<true>, dim3((unsigned)nblk), dim3(kBlock), 0, s,
                               (const uint8_t*)flag, n, dn, bcnt, (const int64_t*)boff, act);
            ABD_CHECK_LAUNCH();
            dev_nact = boff + nblk;
        }
    }
    if (!dev_nact && n_active_dev && (rc = zero_fill(s, n_active_dev, sizeof(int64_t)))) return rc;
    const bool dense = P->psd >= 2;
    const int stride = dense ? kDenseStride : kContactStride;
    if (dn) {
        nact = dev_nact ? n : 0;  // an upper bound: the stages below read the count on the device
        // The bound is the capacity of the candidate lists, which a superset swept by a very
        // long step can make many times the contacts: past the tuning's contact_bound_mib of
        // stored contacts the active count is read back and sizes the buffers instead (never
        // while a graph is recorded: that iteration then runs eagerly)
        const int mib = tuning().contact_bound_mib;
        if ((double)nact * (8.0 * stride + 9.0) > (mib > 0 ? mib : kGhBoundMib) * 1048576.0) {
            if (capturing()) return ABD_ERR_CAPACITY;
            int64_t* hv;
            if ((rc = arena_host(sizeof(int64_t), (void**)&hv))) return rc;
            ABD_CHECK(hipMemcpyAsync(hv, dev_nact, sizeof(int64_t), hipMemcpyDeviceToHost, s));
            ABD_CHECK(hipStreamSynchronize(s));
            nact = hv[0];
        }
    } else {
        // the active count and the ordering flag in one read-back
        int64_t* hv;
        if ((rc = arena_host(2 * sizeof(int64_t), (void**)&hv))) return rc;
        hv[0] = 0;
        if (dev_nact) ABD_CHECK(hipMemcpyAsync(hv, dev_nact, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        ABD_CHECK(hipMemcpyAsync(hv + 1, err, sizeof(int), hipMemcpyDeviceToHost, s));
        ABD_CHECK(hipStreamSynchronize(s));
        nact = hv[0];
        int herr = 0;
        memcpy(&herr, hv + 1, sizeof(int));
        if (herr) return ABD_ERR_UNSORTED;
        if (n_active) *n_active = nact;
    }
    double* pbuf = nullptr;
    if (nij > 0) {
        if ((rc = arena_get(kSlotPair, sizeof(double) * kPairBufStride * (size_t)nij, (void**)&pbuf))) return rc;
    }
    if (nact > 0) {
        double* cbuf;
        int64_t* ckey;
        uint8_t* swp;
        if ((rc = arena_get(kSlotContact, (sizeof(double) * stride + sizeof(int64_t) + 1) * (size_t)nact,
                            (void**)&cbuf)))
            return rc;
        ckey = reinterpret_cast<int64_t*>(cbuf + stride * nact);
        swp = reinterpret_cast<uint8_t*>(ckey + nact);
        const double cbytes = 4.0 + 20 + 16 + 192 + 8.0 * stride + 9;
        int ph = prof_start(kProfContact, s, dn ? 1.0 : cbytes * nact);
        const int64_t* dnact = dn ? dev_nact : nullptr;
        const dim3 cgrid(grid_for(nact, kBlock, 65535));
        if (P->psd == 3)
            hipLaunchKernelGGL(contact_kernel<3>, cgrid, dim3(kBlock), 0, s, *sc, *P, x, pt, npt, ee,
                               (const int32_t*)act, nact, dnact, dn, cbuf, ckey, swp, hl_pt, hl_ee, pflag, nij, pseg);
        else if (P->psd == 2)
            hipLaunchKernelGGL(contact_kernel<2>, cgrid, dim3(kBlock), 0, s, *sc, *P, x, pt, npt, ee,
                               (const int32_t*)act, nact, dnact, dn, cbuf, ckey, swp, hl_pt, hl_ee, pflag, nij, pseg);
        else
            hipLaunchKernelGGL(contact_kernel<1>, cgrid, dim3(kBlock), 0, s, *sc, *P, x, pt, npt, ee,
                               (const int32_t*)act, nact, dnact, dn, cbuf, ckey, swp, hl_pt, hl_ee, pflag, nij, pseg);
        ABD_CHECK_LAUNCH();
        prof_stop(ph, s);
        if ((rc = prof_count_bytes(ph, s, dnact, cbytes, 0.0, 0.0, 0.0, true))) return rc;
        // stage D: chunks of each pair's runs (count, scan), their partial sums, the per-pair sums
        // items: at most one per kPairChunk contacts plus one per run (a run per pair and type)
        const int64_t pgrid = tuning().pair_grid > 0 ? tuning().pair_grid : kPairGrid;
        const int64_t runs = 2 * nij < nact ? 2 * nij : nact;
        const int64_t cap_items = nact / kPairChunk + runs + 1;
        int64_t* pcoff;
        if ((rc = arena_get(kSlotPairChunk, sizeof(int64_t) * ((size_t)nij + 2 + 3 * (size_t)cap_items),
                            (void**)&pcoff)))
            return rc;
        int64_t* ptotal = pcoff + nij;
        int64_t* idesc = ptotal + 2;
        int64_t* odesc = idesc + cap_items;
        int64_t* oitem = odesc + cap_items;
        double* part;
        if ((rc = arena_get(kSlotPairPart, sizeof(double) * kPartStride * (size_t)cap_items, (void**)&part)))
            return rc;
        hipLaunchKernelGGL(pair_chunks_scan_kernel, dim3(1), dim3(kScanThreads), 0, s, nij, (const int64_t*)pseg,
                           pcoff, ptotal);
        ABD_CHECK_LAUNCH();
        hipLaunchKernelGGL(pair_items_kernel, dim3(grid_for(nij, kBlock / kWave)), dim3(kBlock), 0, s, nij,
                           (const int64_t*)pcoff, (const int64_t*)pseg, (const int64_t*)ptotal,
                           idesc, err + 1, odesc, oitem);
        ABD_CHECK_LAUNCH();
        // timed: the partial sums (each stored contact read once, its body-order flag once)
        ph = prof_start_ext(kProfPairReduce, dn ? 1.0 : (8.0 * stride + 1) * nact);
        if (dense)
            prof_launch(ph, pair_partial_dense_kernel, dim3((unsigned)pgrid), dim3(kPairThreads), 0, s,
                        (const double*)cbuf, (const uint8_t*)swp, (const int64_t*)idesc, (const int64_t*)ptotal,
                        part);
        else
            prof_launch(ph, pair_partial_rank_kernel, dim3((unsigned)pgrid), dim3(kRankThreads), 0, s,
                        (const double*)cbuf, (const uint8_t*)swp, (const int64_t*)odesc, (const int64_t*)oitem,
                        (const int64_t*)ptotal, part);
        ABD_CHECK_LAUNCH();
        if ((rc = prof_count_bytes(ph, s, dnact, 8.0 * stride + 1, 0.0, 0.0, 0.0, true))) return rc;
        hipLaunchKernelGGL(pair_reduce_kernel, dim3(grid_for(nij, 1, pgrid)), dim3(kPairThreads), 0, s, nb, nij,
                           dn, (const uint8_t*)pflag, (const int64_t*)pcoff,
                           (const double*)part, pbuf, blocks);
        ABD_CHECK_LAUNCH();
    }
    const double fbytes = 2.0 * 8 * 156 * nb;
    int phf = prof_start(kProfBodyFinalize, s, dn ? 1.0 : fbytes + 8.0 * kPairBufStride * nij + 40.0 * nvg);
    hipLaunchKernelGGL(body_finalize_kernel, dim3(nb), dim3(kBlock), 0, s, *sc, *P, x, nij, dn, row_ptr, slots,
                       (const uint8_t*)pflag, (const double*)pbuf, vg, nvg, g, blocks, (const uint8_t*)hl_pt,
                       (const uint8_t*)hl_ee, highlight);
    ABD_CHECK_LAUNCH();
    prof_stop(phf, s);
    if ((rc = prof_count_bytes(phf, s, dn, 0.0, 0.0, 40.0, 8.0 * kPairBufStride, false, fbytes))) return rc;
    return ABD_OK;
}

}  // namespace abd

using namespace abd;

extern "C" {

int abd_contact_gh(abd_stream_t stream, const AbdScene* sc, const AbdParams* P, const double* x, const int32_t* pt,
                   int64_t npt, const int32_t* ee, int64_t nee, const int32_t* vg, int64_t nvg, int64_t nij,
                   const int32_t* row_ptr, const int32_t* slots, double* g, double* blocks, uint8_t* highlight,
                   int64_t* n_active) {
    if (!sc || !P || !x || !g || !blocks || npt < 0 || nee < 0 || nvg < 0 || nij < 0) return ABD_ERR_ARG;
    if (P->psd < 0 || P->psd > 3) return ABD_ERR_ARG;
    if ((npt && !pt) || (nee && !ee) || (nvg && !vg) || (nij && (!row_ptr || !slots))) return ABD_ERR_ARG;
    if (nij == 0 && (npt || nee)) return ABD_ERR_ARG;
    if (n_active) *n_active = 0;
    GhLists L{pt, npt, ee, nee, vg, nvg, nij, nullptr};
    return contact_gh(to_stream(stream), sc, P, x, L, row_ptr, slots, g, blocks, highlight, n_active, nullptr);
}

int abd_pair_gh(abd_stream_t stream, int32_t kind, const AbdParams* P, const double* x4, const double* r4, int64_t n,
                uint8_t* active, double* g24, double* H24) {
    if (!P || n < 0 || (n && (!x4 || !r4 || !active || !g24 || !H24)) || (kind != 0 && kind != 1)) return ABD_ERR_ARG;
    if (P->psd < 0 || P->psd > 3) return ABD_ERR_ARG;
    if (n == 0) return ABD_OK;
    double* scratch;
    const int stride = P->psd >= 2 ? kDenseStride : kContactStride;
    int rc = arena_get(kSlotContact, sizeof(double) * stride * (size_t)n, (void**)&scratch);
    if (rc) return rc;
    hipLaunchKernelGGL(pair_gh_kernel, dim3(grid_for(n)), dim3(kBlock), 0, to_stream(stream), kind, *P, x4, r4, n,
                       active, g24, H24, scratch);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

int abd_inertia_gradient(abd_stream_t stream, const AbdScene* sc, const AbdParams* P, const double* A, const double* p,
                         const double* A0, const double* p0, const double* Adot, const double* pdot, double* g) {
    if (!sc || !P || !A || !p || !A0 || !p0 || !Adot || !pdot || !g) return ABD_ERR_ARG;
    const int nb = sc->n_bodies;
    if (!nb) return ABD_OK;
    hipLaunchKernelGGL(inertia_gradient_kernel, dim3(grid_for(nb)), dim3(kBlock), 0, to_stream(stream), nb, *P,
                       sc->mass, sc->I0, A, p, A0, p0, Adot, pdot, g);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

int abd_inertia_hessian(abd_stream_t stream, const AbdScene* sc, const AbdParams* P, const double* A,
                        double* blocks) {
    if (!sc || !P || !A || !blocks) return ABD_ERR_ARG;
    const int nb = sc->n_bodies;
    if (!nb) return ABD_OK;
    hipLaunchKernelGGL(inertia_hessian_kernel, dim3(grid_for(16 * (int64_t)nb)), dim3(kBlock), 0, to_stream(stream),
                       nb, *P, sc->mass, sc->I0, A, blocks);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

}  // extern "C"
