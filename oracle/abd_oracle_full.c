/*
 * abd_oracle_full.c — CPU restatement of the full-IPC contact mode (psd = 3; no
 * reference counterpart: the reference gates contacts by the strict interior
 * validity of ccd.py:45-54 and has no distance types).  TEST INFRASTRUCTURE ONLY,
 * like abd_oracle.c.
 *
 * The formulation is that of Li et al. 2020 (Incremental Potential Contact): the
 * barrier of the squared distance to the closest feature pair of a PT / EE
 * candidate, the edge-edge mollifier eps_x = 1e-3 |a1 - a0|^2 |b1 - b0|^2 (rest),
 * m(c) = (2 - c/eps_x) c/eps_x for c = |(a1 - a0) x (b1 - b0)|^2 < eps_x, and the
 * local 12x12 Hessian projected onto the PSD cone.  This restatement shares no
 * arithmetic with the GPU's on purpose:
 *   - the closest feature is the minimum over every admissible feature pair
 *     (interior projections tested explicitly), not a Voronoi-region walk;
 *   - derivatives come from second-order forward differentiation of the plain
 *     distance formulas (struct dd: value, gradient and Hessian in the 12 point
 *     coordinates), not from hand-derived closed forms;
 *   - the projection uses classical Jacobi (largest off-diagonal pivot), not the
 *     cyclic sweep.
 * Pins (tests/test_oracle_full.py): feature distance against dense sampling of the
 * two features, gradient and Hessian against central differences, projection
 * against numpy eigh.
 */
#include <math.h>
#include <string.h>

#define NV 12

typedef struct {
    double v, g[NV], h[NV * NV];
} dd;

static void dd_const(dd* a, double c) {
    memset(a, 0, sizeof(dd));
    a->v = c;
}
static void dd_var(dd* a, double c, int i) {
    dd_const(a, c);
    a->g[i] = 1.0;
}
static void dd_add(const dd* a, const dd* b, dd* o) {
    o->v = a->v + b->v;
    for (int i = 0; i < NV; ++i) o->g[i] = a->g[i] + b->g[i];
    for (int i = 0; i < NV * NV; ++i) o->h[i] = a->h[i] + b->h[i];
}
static void dd_sub(const dd* a, const dd* b, dd* o) {
    o->v = a->v - b->v;
    for (int i = 0; i < NV; ++i) o->g[i] = a->g[i] - b->g[i];
    for (int i = 0; i < NV * NV; ++i) o->h[i] = a->h[i] - b->h[i];
}
static void dd_mul(const dd* a, const dd* b, dd* o) {
    dd r;
    r.v = a->v * b->v;
    for (int i = 0; i < NV; ++i) r.g[i] = a->v * b->g[i] + b->v * a->g[i];
    for (int i = 0; i < NV; ++i)
        for (int j = 0; j < NV; ++j)
            r.h[NV * i + j] = a->v * b->h[NV * i + j] + b->v * a->h[NV * i + j] + a->g[i] * b->g[j] + b->g[i] * a->g[j];
    *o = r;
}
/* o = a / b from a = o b */
static void dd_div(const dd* a, const dd* b, dd* o) {
    dd q;
    q.v = a->v / b->v;
    for (int i = 0; i < NV; ++i) q.g[i] = (a->g[i] - q.v * b->g[i]) / b->v;
    for (int i = 0; i < NV; ++i)
        for (int j = 0; j < NV; ++j)
            q.h[NV * i + j] = (a->h[NV * i + j] - q.v * b->h[NV * i + j] - q.g[i] * b->g[j] - b->g[i] * q.g[j]) / b->v;
    *o = q;
}
/* o = f(a) given f, f', f'' at a */
static void dd_fn(const dd* a, double f, double f1, double f2, dd* o) {
    dd r;
    r.v = f;
    for (int i = 0; i < NV; ++i) r.g[i] = f1 * a->g[i];
    for (int i = 0; i < NV; ++i)
        for (int j = 0; j < NV; ++j) r.h[NV * i + j] = f1 * a->h[NV * i + j] + f2 * a->g[i] * a->g[j];
    *o = r;
}

typedef struct {
    dd c[3];
} dv3;

static void dv_point(const double* x, int t, dv3* o) {
    for (int k = 0; k < 3; ++k) dd_var(&o->c[k], x[3 * t + k], 3 * t + k);
}
static void dv_sub(const dv3* a, const dv3* b, dv3* o) {
    for (int k = 0; k < 3; ++k) dd_sub(&a->c[k], &b->c[k], &o->c[k]);
}
static void dv_dot(const dv3* a, const dv3* b, dd* o) {
    dd t, acc;
    dd_mul(&a->c[0], &b->c[0], &acc);
    dd_mul(&a->c[1], &b->c[1], &t);
    dd_add(&acc, &t, &acc);
    dd_mul(&a->c[2], &b->c[2], &t);
    dd_add(&acc, &t, o);
}
static void dv_cross(const dv3* a, const dv3* b, dv3* o) {
    dd t1, t2;
    dv3 r;
    dd_mul(&a->c[1], &b->c[2], &t1);
    dd_mul(&a->c[2], &b->c[1], &t2);
    dd_sub(&t1, &t2, &r.c[0]);
    dd_mul(&a->c[2], &b->c[0], &t1);
    dd_mul(&a->c[0], &b->c[2], &t2);
    dd_sub(&t1, &t2, &r.c[1]);
    dd_mul(&a->c[0], &b->c[1], &t1);
    dd_mul(&a->c[1], &b->c[0], &t2);
    dd_sub(&t1, &t2, &r.c[2]);
    *o = r;
}

/* feature pair codes (shared with the GPU's numbering only through the tests):
 * type 0 plane / line, 1 point x[i] - edge x[j] x[k], 2 point x[i] - point x[j] */
typedef struct {
    int type, i, j, k;
} feat;

static double pd(const double* x, int a, int b, int k) { return x[3 * a + k] - x[3 * b + k]; }
static double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static void sub3(const double* x, int a, int b, double* o) {
    for (int k = 0; k < 3; ++k) o[k] = pd(x, a, b, k);
}
static void cross3(const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

/* plain values of the three distance families */
static double pp2(const double* x, int i, int j) {
    double d[3];
    sub3(x, i, j, d);
    return dot3(d, d);
}
static double pe2(const double* x, int p, int e0, int e1, double* t_out) {
    double e[3], w[3];
    sub3(x, e1, e0, e);
    sub3(x, p, e0, w);
    const double t = dot3(w, e) / dot3(e, e);
    if (t_out) *t_out = t;
    double c[3];
    cross3(w, e, c);
    return dot3(c, c) / dot3(e, e);
}

/* Closest feature pair: the smallest squared distance over the admissible pairs.
 * kind 0 PT (point 0, triangle 1 2 3), 1 EE (a 0 1, b 2 3). */
static feat closest_feature(int kind, const double* x, double* s_out) {
    feat best = {2, 0, kind == 0 ? 1 : 2, 0};
    double bs = INFINITY;
#define TRY(T, I, J, K, S)              \
    do {                                \
        double s_ = (S);                \
        if (s_ < bs) {                  \
            bs = s_;                    \
            best.type = (T);            \
            best.i = (I);               \
            best.j = (J);               \
            best.k = (K);               \
        }                               \
    } while (0)
    if (kind == 0) {
        for (int t = 1; t <= 3; ++t) TRY(2, 0, t, 0, pp2(x, 0, t));
        const int E[3][2] = {{1, 2}, {1, 3}, {2, 3}};
        for (int e = 0; e < 3; ++e) {
            double t;
            const double s = pe2(x, 0, E[e][0], E[e][1], &t);
            if (t > 0.0 && t < 1.0) TRY(1, 0, E[e][0], E[e][1], s);
        }
        /* plane: barycentric coordinates of the projection by least squares */
        double u[3], v[3], w[3];
        sub3(x, 2, 1, u);
        sub3(x, 3, 1, v);
        sub3(x, 0, 1, w);
        const double uu = dot3(u, u), uv = dot3(u, v), vv = dot3(v, v), wu = dot3(w, u), wv = dot3(w, v);
        const double det = uu * vv - uv * uv;
        if (det > 0.0) {
            const double b1 = (vv * wu - uv * wv) / det, b2 = (uu * wv - uv * wu) / det;
            if (b1 > 0.0 && b2 > 0.0 && b1 + b2 < 1.0) {
                double n[3];
                cross3(u, v, n);
                const double T = dot3(w, n);
                TRY(0, 0, 0, 0, T * T / dot3(n, n));
            }
        }
    } else {
        for (int a = 0; a < 2; ++a)
            for (int b = 2; b < 4; ++b) TRY(2, a, b, 0, pp2(x, a, b));
        for (int a = 0; a < 2; ++a) {
            double t;
            const double s = pe2(x, a, 2, 3, &t);
            if (t > 0.0 && t < 1.0) TRY(1, a, 2, 3, s);
        }
        for (int b = 2; b < 4; ++b) {
            double t;
            const double s = pe2(x, b, 0, 1, &t);
            if (t > 0.0 && t < 1.0) TRY(1, b, 0, 1, s);
        }
        double u[3], v[3], w[3];
        sub3(x, 1, 0, u);
        sub3(x, 3, 2, v);
        sub3(x, 2, 0, w);
        const double uu = dot3(u, u), uv = dot3(u, v), vv = dot3(v, v), wu = dot3(w, u), wv = dot3(w, v);
        const double det = uu * vv - uv * uv;
        if (det > 1e-12 * uu * vv) {
            /* a0 + s u closest to b0 + t v: [uu -uv; uv -vv][s t] = [wu wv] */
            const double s = (wu * vv - wv * uv) / det, t = (wu * uv - wv * uu) / det;
            if (s > 0.0 && s < 1.0 && t > 0.0 && t < 1.0) {
                double n[3];
                cross3(u, v, n);
                const double T = dot3(w, n);
                TRY(0, 0, 0, 0, T * T / dot3(n, n));
            }
        }
    }
#undef TRY
    if (s_out) *s_out = bs;
    return best;
}

/* the squared distance of feature f with its derivatives */
static void feature_dd(int kind, feat f, const double* x, dd* s) {
    dv3 P[4];
    for (int t = 0; t < 4; ++t) dv_point(x, t, &P[t]);
    if (f.type == 2) {
        dv3 d;
        dv_sub(&P[f.i], &P[f.j], &d);
        dv_dot(&d, &d, s);
        return;
    }
    if (f.type == 1) {
        dv3 e, w, c;
        dd cc, ee;
        dv_sub(&P[f.k], &P[f.j], &e);
        dv_sub(&P[f.i], &P[f.j], &w);
        dv_cross(&w, &e, &c);
        dv_dot(&c, &c, &cc);
        dv_dot(&e, &e, &ee);
        dd_div(&cc, &ee, s);
        return;
    }
    dv3 u, v, w, n;
    dd T, T2, nn;
    if (kind == 0) {
        dv_sub(&P[2], &P[1], &u);
        dv_sub(&P[3], &P[1], &v);
        dv_sub(&P[0], &P[1], &w);
    } else {
        dv_sub(&P[1], &P[0], &u);
        dv_sub(&P[3], &P[2], &v);
        dv_sub(&P[2], &P[0], &w);
    }
    dv_cross(&u, &v, &n);
    dv_dot(&w, &n, &T);
    dd_mul(&T, &T, &T2);
    dv_dot(&n, &n, &nn);
    dd_div(&T2, &nn, s);
}

/* classical Jacobi: largest off-diagonal pivot; A (12x12, symmetric) -> PSD part */
static void project_psd12(double* A) {
    double V[NV * NV];
    for (int i = 0; i < NV * NV; ++i) V[i] = (i / NV == i % NV) ? 1.0 : 0.0;
    double fro = 0.0;
    for (int i = 0; i < NV * NV; ++i) fro += A[i] * A[i];
    for (int it = 0; it < 4000 && fro > 0.0; ++it) {
        int p = 0, q = 1;
        double big = 0.0;
        for (int i = 0; i < NV; ++i)
            for (int j = i + 1; j < NV; ++j)
                if (fabs(A[NV * i + j]) > big) {
                    big = fabs(A[NV * i + j]);
                    p = i;
                    q = j;
                }
        if (big * big <= 1e-32 * fro) break;
        /* rotation angle zeroing A[p][q]: tan(2 theta) = 2 apq / (aqq - app) */
        const double theta = 0.5 * atan2(2.0 * A[NV * p + q], A[NV * q + q] - A[NV * p + p]);
        const double c = cos(theta), s = sin(theta);
        for (int k = 0; k < NV; ++k) {
            const double akp = A[NV * k + p], akq = A[NV * k + q];
            A[NV * k + p] = c * akp - s * akq;
            A[NV * k + q] = s * akp + c * akq;
        }
        for (int k = 0; k < NV; ++k) {
            const double apk = A[NV * p + k], aqk = A[NV * q + k];
            A[NV * p + k] = c * apk - s * aqk;
            A[NV * q + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < NV; ++k) {
            const double vkp = V[NV * k + p], vkq = V[NV * k + q];
            V[NV * k + p] = c * vkp - s * vkq;
            V[NV * k + q] = s * vkp + c * vkq;
        }
    }
    double lam[NV];
    for (int i = 0; i < NV; ++i) lam[i] = A[(NV + 1) * i] > 0.0 ? A[(NV + 1) * i] : 0.0;
    for (int i = 0; i < NV; ++i)
        for (int j = 0; j < NV; ++j) {
            double acc = 0.0;
            for (int k = 0; k < NV; ++k) acc += V[NV * i + k] * lam[k] * V[NV * j + k];
            A[NV * i + j] = acc;
        }
}

double orc_barrier(double d, double d2hat, double kappa);
double orc_barrier_d1(double d, double d2hat, double kappa);
double orc_barrier_d2(double d, double d2hat, double kappa);

/* One candidate in the 12 point coordinates x (rest points r): returns 1 when within
 * dhat; E, g12, H12 (projected when `project`), the closest feature (type, i, j, k). */
int orc_full_local(int kind, const double* x, const double* r, double d2hat, double kappa, int project, double* E,
                   double* g12, double* H12, int* feat_out) {
    double s0;
    const feat f = closest_feature(kind, x, &s0);
    if (feat_out) {
        feat_out[0] = f.type;
        feat_out[1] = f.i;
        feat_out[2] = f.j;
        feat_out[3] = f.k;
    }
    if (E) *E = 0.0;
    if (!(s0 < d2hat)) return 0;
    static dd s, b, c, m, e;  /* large: kept off the stack */
    feature_dd(kind, f, x, &s);
    dd_fn(&s, orc_barrier(s.v, d2hat, kappa), orc_barrier_d1(s.v, d2hat, kappa), orc_barrier_d2(s.v, d2hat, kappa),
          &b);
    e = b;
    if (kind == 1) {
        double ru[3], rv[3];
        sub3(r, 1, 0, ru);
        sub3(r, 3, 2, rv);
        const double eps = 1e-3 * dot3(ru, ru) * dot3(rv, rv);
        dv3 P[4], u, v, n;
        for (int t = 0; t < 4; ++t) dv_point(x, t, &P[t]);
        dv_sub(&P[1], &P[0], &u);
        dv_sub(&P[3], &P[2], &v);
        dv_cross(&u, &v, &n);
        dv_dot(&n, &n, &c);
        if (c.v < eps) {
            const double t = c.v / eps;
            dd_fn(&c, (2.0 - t) * t, (2.0 - 2.0 * t) / eps, -2.0 / (eps * eps), &m);
            dd_mul(&m, &b, &e);
        }
    }
    if (E) *E = e.v;
    if (g12) memcpy(g12, e.g, sizeof(double) * NV);
    if (H12) {
        memcpy(H12, e.h, sizeof(double) * NV * NV);
        if (project) project_psd12(H12);
    }
    return 1;
}

/* Full-IPC energy of one candidate (0 when farther than dhat). */
double orc_full_energy(int kind, const double* x, const double* r, double d2hat, double kappa) {
    double s0;
    const feat f = closest_feature(kind, x, &s0);
    (void)f;
    if (!(s0 < d2hat)) return 0.0;
    double E = orc_barrier(s0, d2hat, kappa);
    if (kind == 1) {
        double ru[3], rv[3], u[3], v[3], n[3];
        sub3(r, 1, 0, ru);
        sub3(r, 3, 2, rv);
        sub3(x, 1, 0, u);
        sub3(x, 3, 2, v);
        cross3(u, v, n);
        const double eps = 1e-3 * dot3(ru, ru) * dot3(rv, rv), c = dot3(n, n);
        if (c < eps) E *= (2.0 - c / eps) * (c / eps);
    }
    return E;
}
