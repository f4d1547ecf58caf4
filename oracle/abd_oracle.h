/*
 * abd_oracle.h — shared types of the CPU restatement (test infrastructure only;
 * see abd_oracle.c for what it is and what it is not).
 */
#ifndef ABD_ORACLE_H
#define ABD_ORACLE_H
#include <stdint.h>

/* Runtime parameters; the reference reads module constants instead
 * (const_params.py:24-39) — defect D7: they become arguments.  Field order is
 * shared with the C-ABI's AbdParams and the Python ctypes mirror. */
typedef struct {
    double dt, dhat, d2hat, kappa, stiffness, gravity, ground, mass, I0, c1, alpha_min, tol;
    double solver_tol;
    int32_t max_iter, solver_maxiter, pt, ee, vg, solver;  /* solver: 0 BiCGSTAB, 1 Jacobi CG, 2 block-Jacobi CG, 3 block LDL^T */
    int32_t psd;  /* 0 reference contact Hessian; 1 contact-space PSD projection; 2 exact Hessian;
                     3 full IPC (closest feature pairs, EE mollifier, projected 12x12) —
                     1, 2 and 3 are beyond the reference */
} orc_params;

/* Flattened scene: per-body offsets into vertex / triangle / edge arrays,
 * topology in body-local vertex indices, rest positions x0 [V,3]. */
typedef struct {
    int32_t nb;
    const int32_t* voff;
    const int32_t* foff;
    const int32_t* eoff;
    const int32_t* tris;   /* [F,3] local */
    const int32_t* edges;  /* [E,2] local */
    const double* x0;      /* [V,3] */
} orc_scene;
#endif
