/*
 * abd_oracle_driver.c — CPU FP64 restatement of the reference Newton driver
 * (abd.py:124-362).  TEST INFRASTRUCTURE ONLY (checker and timed CPU baseline);
 * the product never links it.
 *
 * One frame = AffineBodySimulator.step (abd.py:212-276) with the defect fixes
 * D1 (BVH-semantics culling instead of the stub), D3/D4 (correct PT/EE lists in
 * the line-search energy and CCD), D5 (full off-diagonal coupling), D7 (runtime
 * parameters).  Culling here is brute force over every primitive pair of every
 * overlapping body pair (the definition the GPU's candidate sets are compared
 * with).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "abd_oracle.h"


void orc_project(int nb, const int32_t* voff, const double* x0, const double* A, const double* p, double* x);
void orc_prim_boxes(int kind, int nb, const int32_t* voff, const int32_t* offs, const int32_t* topo, const double* x,
                    double dil, float* lo, float* hi);
void orc_prim_boxes_traj(int kind, int nb, const int32_t* voff, const int32_t* offs, const int32_t* topo,
                         const double* x_end, float* lo, float* hi);
void orc_body_boxes(int nb, const int32_t* voff, const double* x, double dil, float* lo, float* hi);
void orc_body_boxes_traj(int nb, const int32_t* voff, const double* x_end, float* lo, float* hi);
long orc_cull_bodies(int nb, const int32_t* body_scene, const float* lo, const float* hi, int32_t* ij, long cap);
long orc_cull_prims(int pt, const int32_t* ij, long nij, const int32_t* off1, const float* lo1, const float* hi1,
                    const int32_t* off2, const float* lo2, const float* hi2, int32_t* out, long cap);
long orc_cull_vg(int nb, const int32_t* voff, const float* blo, const float* plo, const float* phi, double dhat,
                 double ground, int32_t* out, long cap);
double orc_contact_energy(const orc_scene* S, const orc_params* P, const double* xk, const int32_t* pt, long npt,
                          const int32_t* ee, long nee, const int32_t* vg, long nvg);
void orc_contact_gh(const orc_scene* S, const orc_params* P, const double* x, const int32_t* pt, long npt,
                    const int32_t* ee, long nee, const int32_t* vg, long nvg, long nij, double* g, double* blocks,
                    uint8_t* highlight);
double orc_inertia_energy(const orc_params* P, int nb, const double* Ak, const double* pk, const double* A0,
                          const double* p0, const double* Adot, const double* pdot, const double* mass,
                          const double* I0);
void orc_inertia_gradient(const orc_params* P, int nb, const double* A, const double* p, const double* A0,
                          const double* p0, const double* Adot, const double* pdot, const double* mass,
                          const double* I0, double* g);
void orc_inertia_hessian(const orc_params* P, int nb, const double* A, const double* mass, const double* I0,
                         double* blocks);
void orc_set_triplets(int nb, long nij, const int32_t* ij, int32_t* rows, int32_t* cols);
long orc_bicgstab(long nslot, const int32_t* rows, const int32_t* cols, const double* blocks, const double* b,
                  double* x, long n, double tol, long maxiter, int check_every, double* res);
long orc_bsr_cg(long nslot, const int32_t* rows, const int32_t* cols, const double* blocks, const double* b,
                double* x, long n, double tol, long maxiter, int check_every, double* res);
long orc_bsr_cg_block(long nslot, const int32_t* rows, const int32_t* cols, const double* blocks, const double* b,
                      double* x, long n, double tol, long maxiter, int check_every, double* res);
long orc_bsr_ldlt(long nslot, const int32_t* rows, const int32_t* cols, const double* blocks, const double* b,
                  double* x, long n);
double orc_toi(const orc_scene* S, const orc_params* P, const double* x, const double* xk, const int32_t* pt, long npt,
               const int32_t* ee, long nee, const int32_t* vg, long nvg);
void orc_update_q(int nb, const double* A, const double* p, const double* dq, double alpha, double* Aout, double* pout);
void orc_update_q0qdot(int nb, double dt, const double* A, const double* p, double* A0, double* p0, double* Adot,
                       double* pdot);

typedef struct { int32_t* d; long n, cap, w; } ilist;
static void il_fit(ilist* l, long need) {
    if (need > l->cap) {
        l->cap = need + need / 2 + 64;
        l->d = realloc(l->d, (size_t)l->cap * l->w * sizeof(int32_t));
    }
}

typedef struct {
    const orc_scene* S;
    const orc_params* P;
    const int32_t* body_scene;
    const double *mass, *I0;
    int nb, nv, nf, ne;
    double *A, *p, *Ak, *pk, *A0, *p0, *Adot, *pdot, *dq;
    double *x, *xk;
    float *blo, *bhi, *plo, *phi, *elo, *ehi, *tlo, *thi;        /* current (collision_set) */
    float *Tblo, *Tbhi, *Tplo, *Tphi, *Telo, *Tehi, *Ttlo, *Tthi; /* trajectory */
    ilist ij, pt, ee, vg;
    int64_t* stats;
} sim;

static long cull_b(sim* s, const float* lo, const float* hi, ilist* out) {
    long n = orc_cull_bodies(s->nb, s->body_scene, lo, hi, out->d, out->cap);
    if (n > out->cap) { il_fit(out, n); n = orc_cull_bodies(s->nb, s->body_scene, lo, hi, out->d, out->cap); }
    out->n = n;
    return n;
}
static long cull_p(int pt, const ilist* ij, const int32_t* o1, const float* l1, const float* h1, const int32_t* o2,
                   const float* l2, const float* h2, ilist* out) {
    long n = orc_cull_prims(pt, ij->d, ij->n, o1, l1, h1, o2, l2, h2, out->d, out->cap);
    if (n > out->cap) { il_fit(out, n); n = orc_cull_prims(pt, ij->d, ij->n, o1, l1, h1, o2, l2, h2, out->d, out->cap); }
    out->n = n;
    return n;
}
static long cull_v(sim* s, const float* blo, const float* plo, const float* phi, ilist* out) {
    long n = orc_cull_vg(s->nb, s->S->voff, blo, plo, phi, s->P->dhat, s->P->ground, out->d, out->cap);
    if (n > out->cap) {
        il_fit(out, n);
        n = orc_cull_vg(s->nb, s->S->voff, blo, plo, phi, s->P->dhat, s->P->ground, out->d, out->cap);
    }
    out->n = n;
    return n;
}

/* abd.py:187-207 collision_set (on xk) */
static void collision_set(sim* s) {
    const orc_scene* S = s->S;
    double dh = s->P->dhat * 0.5;
    orc_body_boxes(s->nb, S->voff, s->xk, dh, s->blo, s->bhi);
    cull_b(s, s->blo, s->bhi, &s->ij);
    orc_prim_boxes(2, s->nb, S->voff, S->foff, S->tris, s->xk, 0.0, s->tlo, s->thi);
    orc_prim_boxes(1, s->nb, S->voff, S->eoff, S->edges, s->xk, dh, s->elo, s->ehi);
    orc_prim_boxes(0, s->nb, S->voff, S->voff, NULL, s->xk, s->P->dhat, s->plo, s->phi);
    cull_p(0, &s->ij, S->eoff, s->elo, s->ehi, S->eoff, s->elo, s->ehi, &s->ee);
    cull_p(1, &s->ij, S->voff, s->plo, s->phi, S->foff, s->tlo, s->thi, &s->pt);
    cull_v(s, s->blo, s->plo, s->phi, &s->vg);
}

/* abd.py:170-185 trajectory_intersection_set (stored boxes from the last
 * collision_set, endpoint xk) */
static void trajectory_set(sim* s) {
    const orc_scene* S = s->S;
    memcpy(s->Tblo, s->blo, 12 * (size_t)s->nb); memcpy(s->Tbhi, s->bhi, 12 * (size_t)s->nb);
    orc_body_boxes_traj(s->nb, S->voff, s->xk, s->Tblo, s->Tbhi);
    cull_b(s, s->Tblo, s->Tbhi, &s->ij);
    memcpy(s->Ttlo, s->tlo, 12 * (size_t)s->nf); memcpy(s->Tthi, s->thi, 12 * (size_t)s->nf);
    memcpy(s->Telo, s->elo, 12 * (size_t)s->ne); memcpy(s->Tehi, s->ehi, 12 * (size_t)s->ne);
    memcpy(s->Tplo, s->plo, 12 * (size_t)s->nv); memcpy(s->Tphi, s->phi, 12 * (size_t)s->nv);
    orc_prim_boxes_traj(2, s->nb, S->voff, S->foff, S->tris, s->xk, s->Ttlo, s->Tthi);
    orc_prim_boxes_traj(1, s->nb, S->voff, S->eoff, S->edges, s->xk, s->Telo, s->Tehi);
    orc_prim_boxes_traj(0, s->nb, S->voff, S->voff, NULL, s->xk, s->Tplo, s->Tphi);
    cull_p(0, &s->ij, S->eoff, s->Telo, s->Tehi, S->eoff, s->Telo, s->Tehi, &s->ee);
    cull_p(1, &s->ij, S->voff, s->Tplo, s->Tphi, S->foff, s->Ttlo, s->Tthi, &s->pt);
    cull_v(s, s->Tblo, s->Tplo, s->Tphi, &s->vg);
}

/* abd.py:133-141 compute_energy */
static double compute_energy(sim* s, double alpha) {
    orc_update_q(s->nb, s->A, s->p, s->dq, alpha, s->Ak, s->pk);
    orc_project(s->nb, s->S->voff, s->S->x0, s->Ak, s->pk, s->xk);
    collision_set(s);
    double ec = orc_contact_energy(s->S, s->P, s->xk, s->pt.d, s->pt.n, s->ee.d, s->ee.n, s->vg.d, s->vg.n);
    double ei = orc_inertia_energy(s->P, s->nb, s->Ak, s->pk, s->A0, s->p0, s->Adot, s->pdot, s->mass, s->I0);
    if (s->stats) s->stats[5] += 1;
    return ec + ei;
}
static double vdot(const double* a, const double* b, long n) {
    double r = 0.0;
    for (long i = 0; i < n; ++i) r += a[i] * b[i];
    return r;
}

/* abd.py:278-303 ccd */
static double ccd(sim* s) {
    orc_update_q(s->nb, s->A, s->p, s->dq, 1.0, s->Ak, s->pk);
    orc_project(s->nb, s->S->voff, s->S->x0, s->Ak, s->pk, s->xk);
    orc_project(s->nb, s->S->voff, s->S->x0, s->A, s->p, s->x);
    trajectory_set(s);
    double t = orc_toi(s->S, s->P, s->x, s->xk, s->pt.d, s->pt.n, s->ee.d, s->ee.n, s->vg.d, s->vg.n);
    if (s->stats) { s->stats[6] += s->pt.n; s->stats[7] += s->ee.n; }
    t = t < 1.0 ? t : 1.0;
    t = t > 0.0 ? t : 0.0;
    if (t < 1.0) t *= 0.9;
    return t;
}

/* abd.py:212-276 step (one frame).  States are caller-owned and updated in
 * place; dq is the warm start carried across iterations and frames.
 * stats (nullable, int64[8]): [0] Newton iterations, [1] gh PT candidates,
 * [2] gh EE candidates, [3] gh VG candidates, [4] solver iterations,
 * [5] energy evaluations, [6] CCD PT candidates, [7] CCD EE candidates. */
int orc_step(const orc_scene* S, const orc_params* P, const int32_t* body_scene, const double* mass, const double* I0,
             int nv, int nf, int ne, double* A, double* p, double* Ak, double* pk, double* A0, double* p0, double* Adot,
             double* pdot, double* dq, int64_t* stats, double* E_out) {
    sim s;
    memset(&s, 0, sizeof(s));
    s.S = S; s.P = P; s.body_scene = body_scene; s.mass = mass; s.I0 = I0;
    s.nb = S->nb; s.nv = nv; s.nf = nf; s.ne = ne;
    s.A = A; s.p = p; s.Ak = Ak; s.pk = pk; s.A0 = A0; s.p0 = p0; s.Adot = Adot; s.pdot = pdot; s.dq = dq;
    s.stats = stats;
    s.x = malloc(24 * (size_t)nv); s.xk = malloc(24 * (size_t)nv);
    float** fb[] = {&s.blo, &s.bhi, &s.Tblo, &s.Tbhi};
    for (int i = 0; i < 4; ++i) *fb[i] = malloc(12 * (size_t)s.nb);
    s.plo = malloc(12 * (size_t)nv); s.phi = malloc(12 * (size_t)nv);
    s.Tplo = malloc(12 * (size_t)nv); s.Tphi = malloc(12 * (size_t)nv);
    s.elo = malloc(12 * (size_t)ne + 12); s.ehi = malloc(12 * (size_t)ne + 12);
    s.Telo = malloc(12 * (size_t)ne + 12); s.Tehi = malloc(12 * (size_t)ne + 12);
    s.tlo = malloc(12 * (size_t)nf + 12); s.thi = malloc(12 * (size_t)nf + 12);
    s.Ttlo = malloc(12 * (size_t)nf + 12); s.Tthi = malloc(12 * (size_t)nf + 12);
    s.ij.w = 2; s.pt.w = 5; s.ee.w = 5; s.vg.w = 2;
    il_fit(&s.ij, 1024); il_fit(&s.pt, 4096); il_fit(&s.ee, 4096); il_fit(&s.vg, 1024);
    long n12 = 12L * s.nb;
    double* g = malloc(8 * (size_t)n12);

    /* q_gets_q0 (abd.py:124-128) */
    memcpy(A, A0, 72 * (size_t)s.nb); memcpy(p, p0, 24 * (size_t)s.nb);
    memcpy(Ak, A0, 72 * (size_t)s.nb); memcpy(pk, p0, 24 * (size_t)s.nb);
    double E0 = compute_energy(&s, 0.0);
    int it = 0;
    for (;;) {
        orc_project(s.nb, S->voff, S->x0, A, p, s.x);
        orc_project(s.nb, S->voff, S->x0, Ak, pk, s.xk);
        collision_set(&s);
        long nij = s.ij.n;
        long nslot = s.nb + 2 * nij;
        double* blocks = calloc((size_t)nslot * 16 * 9, 8);
        int32_t* rows = malloc(4 * (size_t)nslot * 16);
        int32_t* cols = malloc(4 * (size_t)nslot * 16);
        memset(g, 0, 8 * (size_t)n12);
        orc_inertia_gradient(P, s.nb, A, p, A0, p0, Adot, pdot, mass, I0, g);
        orc_inertia_hessian(P, s.nb, A, mass, I0, blocks);
        orc_contact_gh(S, P, s.x, s.pt.d, s.pt.n, s.ee.d, s.ee.n, s.vg.d, s.vg.n, nij, g, blocks, NULL);
        if (stats) { stats[1] += s.pt.n; stats[2] += s.ee.n; stats[3] += s.vg.n; }
        orc_set_triplets(s.nb, nij, s.ij.d, rows, cols);
        double res;
        long sit;
        if (P->solver == 3) {
            /* direct block LDL^T; the block-Jacobi CG where it does not apply */
            sit = orc_bsr_ldlt(nslot, rows, cols, blocks, g, dq, n12);
            if (sit < 0)
                sit = orc_bsr_cg_block(nslot, rows, cols, blocks, g, dq, n12, P->solver_tol, P->solver_maxiter, 10,
                                       &res);
        } else if (P->solver == 2)
            sit = orc_bsr_cg_block(nslot, rows, cols, blocks, g, dq, n12, P->solver_tol, P->solver_maxiter, 10, &res);
        else if (P->solver == 1)
            sit = orc_bsr_cg(nslot, rows, cols, blocks, g, dq, n12, P->solver_tol, P->solver_maxiter, 10, &res);
        else
            sit = orc_bicgstab(nslot, rows, cols, blocks, g, dq, n12, P->solver_tol, P->solver_maxiter, 10, &res);
        if (stats) stats[4] += sit;
        free(blocks); free(rows); free(cols);
        double alpha_cap = ccd(&s);
        /* abd.py:146-160 line_search (D8: Armijo sign kept as written) */
        double qTg = vdot(dq, g, n12);
        double alpha = alpha_cap, E1;
        for (;;) {
            E1 = compute_energy(&s, alpha);
            int wolfe = E1 < E0 + P->c1 * alpha * qTg;
            if (wolfe || alpha < P->alpha_min) break;
            alpha *= 0.5;
        }
        E0 = E1;
        orc_update_q(s.nb, A, p, dq, alpha, A, p);
        ++it;
        double gTg = vdot(g, g, n12);
        if (gTg < P->tol || it >= P->max_iter) break;
    }
    orc_update_q0qdot(s.nb, P->dt, A, p, A0, p0, Adot, pdot);
    if (stats) stats[0] += it;
    if (E_out) *E_out = E0;
    free(g); free(s.x); free(s.xk);
    for (int i = 0; i < 4; ++i) free(*fb[i]);
    free(s.plo); free(s.phi); free(s.Tplo); free(s.Tphi); free(s.elo); free(s.ehi); free(s.Telo); free(s.Tehi);
    free(s.tlo); free(s.thi); free(s.Ttlo); free(s.Tthi);
    free(s.ij.d); free(s.pt.d); free(s.ee.d); free(s.vg.d);
    return it;
}
