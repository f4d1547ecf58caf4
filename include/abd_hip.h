/*
 * abd_hip.h — C ABI of libabd_hip.so, the MI355X (gfx950) implementation of
 * the abd-warp per-Newton-iteration contact path.
 *
 * Conventions
 *  - Every pointer argument is a DEVICE pointer (contiguous arrays in the
 *    layouts below) unless its comment says "host".
 *  - Every call is stream-ordered on `stream` (a hipStream_t; NULL = the
 *    default stream).  Calls that must return a size to the host (candidate
 *    counts, active counts) wait for `stream` up to the point where that size
 *    is known — as the reference's `compress()` does (simulator/fenwick.py:152-170);
 *    the list itself is written by work still queued on `stream`.
 *  - Return value: ABD_OK (0); ABD_ERR_OVERFLOW (1) when an output list is too
 *    small (the required count is written to *count and nothing is written past
 *    `cap`); negative on error (ABD_ERR_* below, or -hipError_t).
 *  - Scratch memory (partial sums, compaction counters, filtered primitive
 *    lists) comes from an arena that the library grows on demand.  The
 *    stateless entry points share one arena per device, kept until
 *    abd_release(): those calls must not overlap in time on one device.  Each
 *    simulator (abd_sim_*) owns an arena of its own, freed by abd_sim_destroy,
 *    so simulators may be stepped concurrently from several host threads.
 *  - No C++ exception crosses this boundary.
 *
 * Data layouts (shared with the Python mirror and the CPU oracle)
 *  - vertices: double [V][3]; per body b the range voff[b]..voff[b+1].
 *  - triangles int32 [F][3], edges int32 [E][2], in BODY-LOCAL vertex ids
 *    (the reference's WarpMesh.triangles / .edges, affine_body.py:85-95).
 *  - A: double [Nb][3][3], row r = a_{r+1}, x = A^T x0 + p (abd.py:313-316);
 *    p: double [Nb][3]; gradient / dq: double [4 Nb][3] (dp, da1, da2, da3).
 *  - boxes: float [n][3] lower and upper.
 *  - candidate tuples: int32 [n][5] = (I, J, ii, prim_I, prim_J) with body-local
 *    primitive ids and ii the index of the body pair in the ij list; VG tuples
 *    int32 [n][2] = (body, local vertex) (the reference's vec5i / vec2i).
 *  - body pairs ij: int32 [Nij][2], i < j, sorted lexicographically; ji: the
 *    same pairs sorted by (j, i).
 *  - Hessian slots: double [(Nb + 2 Nij) * 16][3][3]; slot b < Nb is body b's
 *    diagonal 12x12 block, slot Nb + ii the (ij[ii][0], ij[ii][1]) block, slot
 *    Nb + Nij + ii its transpose; 3x3 sub-block (ii, jj) of a slot at
 *    16*slot + ii + 4*jj (ipcsk/put.py:26-49, orthogonal_energy.py:138-165).
 */
#ifndef ABD_HIP_H
#define ABD_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* abd_stream_t; /* hipStream_t */

#define ABD_OK 0
#define ABD_ERR_OVERFLOW 1
#define ABD_ERR_ARG (-1000)      /* bad argument (null pointer, negative size, ...) */
#define ABD_ERR_UNSORTED (-1001) /* candidate list not grouped by body pair / body */
#define ABD_ERR_NODEVICE (-1002) /* no HIP device */
#define ABD_ERR_CAPACITY (-1003) /* input larger than the fixed scratch of an allocation-free entry point */

#define ABD_KIND_POINT 0
#define ABD_KIND_EDGE 1
#define ABD_KIND_TRI 2
#define ABD_CULL_EE 0
#define ABD_CULL_PT 1
#define ABD_SOLVER_BICGSTAB 0
#define ABD_SOLVER_CG_JACOBI 1
#define ABD_SOLVER_CG_BLOCK 2 /* CG, 12x12 block-Jacobi preconditioner (no reference counterpart) */
/* Direct block LDL^T of the 12x12 body blocks (no reference counterpart) for systems of at
 * most 64 bodies; a larger system, or one whose factorization meets a pivot that is not
 * positive, is solved by the 12x12 block-Jacobi CG (ABD_SOLVER_CG_BLOCK) instead.  iters: 1 for a direct solve (else the
 * CG's count); resid: |b - H x|^2 for a direct solve. */
#define ABD_SOLVER_LDLT 3

/* Launch shapes and buffer sizing that change no result (no reference counterpart):
 * candidate lists are exact and emitted in sorted order whatever the tiling.  cull_chunk:
 * primitive-pair tests per chunk of abd_cull_prims (power of two, 256..65536; default
 * 4096); cull_grid: most workgroups of its chunk passes (64..2^20; default 16384);
 * contact_bound_mib: the simulator's per-iteration contact buffer is sized by its candidate
 * lists' capacity while that needs at most this many MiB (0: 8192), else by the active
 * count read back (one host wait; such an iteration is launched, not replayed from a graph);
 * pair_grid: workgroups of the contact Hessian's per-pair sums (64..65536; 0: 2048: each work
 * item is summed by one workgroup in a fixed order whatever the grid); toi_grid: most
 * workgroups of the CCD's root searches (64..65536; 0: 2048; the minima are exact).
 * Process-wide; a change invalidates recorded graphs.  Zeroed trailing fields are the defaults. */
typedef struct AbdTuning {
    int32_t cull_chunk, cull_grid;
    int32_t contact_bound_mib;
    int32_t pair_grid, toi_grid;
} AbdTuning;
int abd_get_tuning(AbdTuning* tuning);
int abd_set_tuning(const AbdTuning* tuning);

/* Runtime parameters (the reference's module constants, const_params.py:24-39). */
/* psd: 0 = the reference's contact Hessian (PT eigenvalues of d2(l^2)/dc2 clamped, EE
 * not, then b' H + b'' g g^T: indefinite since b' < 0); 1 = the whole barrier Hessian
 * projected in contact space (eigenvalues b' L_k and b' L_4 + b'' |q_4|^2 clamped at 0),
 * which keeps the Newton matrix positive semi-definite; 2 = the exact Hessian of the
 * barrier energy (b' d2s/dx2 + b'' ds/dx ds/dx^T with the curvature of the plane / line
 * projection, and the b' term of the vertex-ground block), which the adjoint needs;
 * 3 = full IPC (Li et al. 2020): each PT / EE candidate's barrier of the squared
 * distance to its closest feature pair (point-plane / line-line, point-edge,
 * point-point; no validity gate), EE times the mollifier of nearly parallel edges,
 * the 12x12 Hessian in the points' coordinates projected onto the PSD cone, the
 * vertex-ground block's 4 d^2 b'' + 2 b' clamped at 0.
 * 1, 2 and 3 have no reference counterpart.  The entry points that read psd (contact
 * gradient/Hessian, pair terms, contact energy, simulator creation) return ABD_ERR_ARG
 * for any other value. */
typedef struct AbdParams {
    double dt, dhat, d2hat, kappa, stiffness, gravity, ground, mass, I0, c1, alpha_min, tol;
    double solver_tol;
    int32_t max_iter, solver_maxiter, pt, ee, vg, solver, psd;
} AbdParams;

/* Flattened scene (replaces the per-body WarpMesh structs, affine_body.py:85-95).
 * All pointers device; body_scene / mass / I0 may be NULL. */
typedef struct AbdScene {
    int32_t n_bodies, n_verts, n_tris, n_edges;
    int32_t max_verts, max_tris, max_edges; /* per-body maxima */
    const int32_t* voff;                     /* [Nb+1] */
    const int32_t* foff;                     /* [Nb+1] */
    const int32_t* eoff;                     /* [Nb+1] */
    const int32_t* vert_body;                /* [V] */
    const int32_t* tri_body;                 /* [F] */
    const int32_t* edge_body;                /* [E] */
    const int32_t* tris;                     /* [F][3] local */
    const int32_t* edges;                    /* [E][2] local */
    const double* x0;                        /* [V][3] rest shape */
    const int32_t* body_scene;               /* [Nb] scene id; bodies of different scenes never pair */
    const double* mass;                      /* [Nb] (NULL: params.mass) */
    const double* I0;                        /* [Nb] (NULL: params.I0) */
    /* Independent scenes (batching; no reference counterpart): the bodies of scene s are
     * scene_boff[s]..scene_boff[s+1] (contiguous); scene_poff[s] = number of body pairs
     * i < j inside scenes before s; n_scene_pairs = scene_poff[n_scenes] (host copy).
     * n_scenes <= 1 or NULL offsets: one scene (all pairs enumerated). */
    int32_t n_scenes;
    const int32_t* scene_boff;               /* [S+1] */
    const int64_t* scene_poff;               /* [S+1] */
    int64_t n_scene_pairs;
} AbdScene;

/* ---- library ---- */
int abd_version(void);
const char* abd_error_string(int code);
/* Frees the scratch arena of the current device. */
int abd_release(void);

/* ---- state updates ---- */
/* abd.py:305-324 update_mesh_vertex: x = A^T x0 + p for every vertex. */
int abd_project_vertices(abd_stream_t stream, const AbdScene* scene, const double* A, const double* p, double* x);
/* abd.py:337-353 _update_q / _update_qk: (A_out, p_out) = (A, p) - alpha * dq (in place allowed). */
int abd_update_q(abd_stream_t stream, int32_t nb, const double* A, const double* p, const double* dq, double alpha,
                 double* A_out, double* p_out);
/* abd.py:355-362 _update_q0qdot: qdot = (q - q0) / dt; q0 = q. */
int abd_update_q0qdot(abd_stream_t stream, int32_t nb, double dt, const double* A, const double* p, double* A0,
                      double* p0, double* Adot, double* pdot);
/* abd.py:143-144 dot (warp.utils.array_inner): *out (device) = a . b, fixed-order reduction. */
int abd_dot(abd_stream_t stream, const double* a, const double* b, int64_t n, double* out);

/* ---- bounding boxes (culling2.py) ---- */
/* culling2.py:33-79 bvh_triangles / bvh_edges / bvh_points: FP32 boxes of every primitive
 * of `kind` (ABD_KIND_*) dilated by `dilation` (points: cast then +-f32(dil); edges and
 * triangles: FP64 min/max +- dil, then cast). */
int abd_prim_boxes(abd_stream_t stream, const AbdScene* scene, int32_t kind, const double* x, double dilation,
                   float* lo, float* hi);
/* culling2.py:230-280 bvh_*_trajectory: lo/hi <- union(lo/hi, f32 box of x_end, undilated). */
int abd_prim_boxes_traj(abd_stream_t stream, const AbdScene* scene, int32_t kind, const double* x_end, float* lo,
                        float* hi);
/* culling2.py:171-186 update_body_bvh: per-body FP64 min/max of x +- dilation, cast to f32. */
int abd_body_boxes(abd_stream_t stream, const AbdScene* scene, const double* x, double dilation, float* lo, float* hi);
/* culling2.py:188-210 body_bvh_to_traj: lo/hi <- union(lo/hi, per-body box of x_end, undilated). */
int abd_body_boxes_traj(abd_stream_t stream, const AbdScene* scene, const double* x_end, float* lo, float* hi);

/* ---- culling ---- */
/* culling2.py:214-227 intersection_bodies + abd.py:163-168 ij_list: every body pair
 * (i < j, same scene) whose boxes overlap (closed intervals).  Writes ij sorted by (i, j)
 * and ji (the same pairs sorted by (j, i)); *count (host) = number of pairs.
 * Synchronizes.  ABD_ERR_OVERFLOW when *count > cap. */
int abd_cull_bodies(abd_stream_t stream, const AbdScene* scene, const float* lo, const float* hi, int32_t* ij,
                    int32_t* ji, int64_t cap, int64_t* count);
/* culling.py:32-52 cull_ + culling.py:86-101 intersection_prims (the BVH semantics the
 * live stub culling.py:8-30 stands in for): for each body pair ii,
 *   ABD_CULL_PT: (I, J, ii, p, t) for every point p of I (set 1 = point boxes) whose box
 *                overlaps triangle t of J (set 2 = triangle boxes), both orders
 *                (I, J) = (ij0, ij1) then (ij1, ij0);
 *   ABD_CULL_EE: (ij0, ij1, ii, a, b) for edge a of ij0 overlapping edge b of ij1
 *                (set 1 = set 2 = edge boxes).
 * Output is sorted by (ii, order, prim_I, prim_J).  *count (host).  Synchronizes.
 * ABD_ERR_OVERFLOW when *count > cap. */
int abd_cull_prims(abd_stream_t stream, const AbdScene* scene, int32_t mode, const int32_t* ij, int64_t nij,
                   const float* lo1, const float* hi1, const float* lo2, const float* hi2, int32_t* out, int64_t cap,
                   int64_t* count);
/* culling.py:54-76 cull_vg + intersection_ground: (b, v) for bodies with f64(blo.y) <
 * ground + dhat/2 and vertices whose point box overlaps [(-1e3,-1e3,-1e3),(1e3,f32(ground),1e3)].
 * Sorted by (b, v).  *count (host).  Synchronizes. */
int abd_cull_vg(abd_stream_t stream, const AbdScene* scene, const float* blo, const float* plo, const float* phi,
                double dhat, double ground, int32_t* out, int64_t cap, int64_t* count);
/* The collision set from a superset of it (no reference counterpart: the reference culls
 * from scratch at every state, abd.py:187-207): the body pairs (ij, and ji when not NULL),
 * PT / EE / VG tuples of a candidate set made from larger boxes (e.g. boxes swept over a
 * step), kept when they pass the culling predicates on the current boxes (body boxes
 * blo/bhi, point / edge / triangle boxes); tuples take their pairs' new indices.  When the
 * current boxes lie inside those the superset was made from, the outputs equal
 * abd_cull_bodies / abd_cull_prims / abd_cull_vg on the current boxes.  Outputs sized like
 * the inputs, not aliasing them; counts (host) [5] = nij, nij, npt, nee, nvg.
 * Synchronizes. */
int abd_cull_keep(abd_stream_t stream, const AbdScene* scene, const int32_t* ij, const int32_t* ji, int64_t nij,
                  const int32_t* pt, int64_t npt, const int32_t* ee, int64_t nee, const int32_t* vg, int64_t nvg,
                  const float* blo, const float* bhi, const float* plo, const float* phi, const float* elo,
                  const float* ehi, const float* tlo, const float* thi, double dhat, double ground, int32_t* ij_out,
                  int32_t* ji_out, int32_t* pt_out, int32_t* ee_out, int32_t* vg_out, int64_t* counts);

/* ---- sparse structure ---- */
/* orthogonal_energy.py:138-165 _set_triplets (over all Nb + 2 Nij slots): block row/col of
 * every 3x3 sub-block, rows[16 s + ii + 4 jj] = ii + 4 I, cols[...] = jj + 4 J. */
int abd_set_triplets(abd_stream_t stream, int32_t nb, const int32_t* ij, int64_t nij, int32_t* rows, int32_t* cols);
/* abd.py:240-243 bsr_set_from_triplets, as a CSR over bodies with 12x12 blocks:
 * row b holds (col, slot) for its lower pairs (from ji), its diagonal, its upper pairs
 * (from ij), columns ascending.  row_ptr [Nb+1], cols/slots [Nb + 2 Nij]. */
int abd_bsr_structure(abd_stream_t stream, int32_t nb, const int32_t* ij, const int32_t* ji, int64_t nij,
                      int32_t* row_ptr, int32_t* cols, int32_t* slots);

/* ---- energies, gradient, Hessian ---- */
/* ipc.py:19-46 IPCContactEnergy.energy at xk: *E (device) = sum of barrier energies of
 * valid PT (|e2p|^2), valid EE (l^2) and all VG candidates (ipcsk/energy.py). */
int abd_contact_energy(abd_stream_t stream, const AbdScene* scene, const AbdParams* params, const double* xk,
                       const int32_t* pt, int64_t npt, const int32_t* ee, int64_t nee, const int32_t* vg, int64_t nvg,
                       double* E);
/* ipc.py:54-62 IPCContactEnergy.gh at x (ipcsk/pt_ipc.py:24-190, ipcsk/ee_ipc.py:50-143,
 * ipcsk/vg_ipc.py:5-34, ipcsk/put.py): ADDS the barrier gradient into g [4Nb][3] and the
 * PSD-projected Hessian into the slots `blocks`.  pt/ee must be grouped by ii ascending
 * and vg by body ascending (as abd_cull_* emit them; else ABD_ERR_UNSORTED).  row_ptr/
 * slots: the body CSR of abd_bsr_structure for the nij body pairs the tuples index.
 * highlight (nullable) [Nb] gets bit 0 for bodies J of active PT contacts and bit 1
 * for EE.  *n_active (host, nullable) = active PT + EE contacts.  Synchronizes. */
int abd_contact_gh(abd_stream_t stream, const AbdScene* scene, const AbdParams* params, const double* x,
                   const int32_t* pt, int64_t npt, const int32_t* ee, int64_t nee, const int32_t* vg, int64_t nvg,
                   int64_t nij, const int32_t* row_ptr, const int32_t* slots, double* g, double* blocks,
                   uint8_t* highlight, int64_t* n_active);
/* Per-contact form of the gh stage for n independent contacts (kind 0 PT, 1 EE): x4/r4
 * [n][4][3] current / rest positions; writes active [n] (uint8), g24 [n][24] and the 24x24
 * body-pair Hessian H24 [n][24][24] (rows/cols 0-11 body I, 12-23 body J). */
int abd_pair_gh(abd_stream_t stream, int32_t kind, const AbdParams* params, const double* x4, const double* r4,
                int64_t n, uint8_t* active, double* g24, double* H24);
/* orthogonal_energy.py:25-32 energy_inertia at (Ak, pk): *E (device). */
int abd_inertia_energy(abd_stream_t stream, const AbdScene* scene, const AbdParams* params, const double* Ak,
                       const double* pk, const double* A0, const double* p0, const double* Adot, const double* pdot,
                       double* E);
/* orthogonal_energy.py:17-19,99-112: g <- inertia gradient at (A, p) (overwrites g). */
int abd_inertia_gradient(abd_stream_t stream, const AbdScene* scene, const AbdParams* params, const double* A,
                         const double* p, const double* A0, const double* p0, const double* Adot, const double* pdot,
                         double* g);
/* orthogonal_energy.py:21-22,84-96: diagonal slots += inertia Hessian at A. */
int abd_inertia_hessian(abd_stream_t stream, const AbdScene* scene, const AbdParams* params, const double* A,
                        double* blocks);

/* ---- linear solve ---- */
/* abd.py:244 bicgstab(hess, g, dq, tol) (method ABD_SOLVER_BICGSTAB, warp.optim.linear) or
 * bsr_utils.py:126-235 bsr_cg with the diagonal preconditioner (ABD_SOLVER_CG_JACOBI):
 * solves H x = b with x as the warm start, H the BSR (row_ptr/cols/slots over 12x12
 * slots).  tol: BiCGSTAB relative tol on |r| (|r| <= tol |b|); CG absolute tol^2 n on r.z.
 * maxiter 0 -> 12 Nb.  iters/resid (device, nullable) receive the iteration count and
 * the final |r|^2 (BiCGSTAB) or r.z (CG).  One workgroup per system. */
int abd_bsr_solve(abd_stream_t stream, int32_t method, int32_t nb, const int32_t* row_ptr, const int32_t* cols,
                  const int32_t* slots, const double* blocks, const double* b, double* x, double tol, int32_t maxiter,
                  int32_t check_every, int32_t* iters, double* resid);

/* ---- CCD ---- */
/* ccd.py:9-40 toi_vg / toi_pt / toi_ee + abd.py:284-299 (in-repo cubic TOI; the reference
 * calls the absent abdtk instead): *toi (device) = min(1, TOIs < 1 of every candidate),
 * trajectories x_t0 -> x_t1. */
int abd_toi(abd_stream_t stream, const AbdScene* scene, const AbdParams* params, const double* x_t0,
            const double* x_t1, const int32_t* pt, int64_t npt, const int32_t* ee, int64_t nee, const int32_t* vg,
            int64_t nvg, double* toi);
/* abd_toi without allocation (no reference counterpart; for callers that cannot call
 * hipMalloc: code under HIP graph capture, serving loops, harnesses that forbid it).
 * Same arguments, kernels and result as abd_toi, bit for bit; its scratch (the chunk plan,
 * the chunks' partial minima and the queue of candidates whose cubic has a bracketed root)
 * is a fixed array in device memory of about 4.1 MiB, held by every device the library
 * runs on until the process ends: npt + nee + nvg must not exceed
 * ABD_TOI_NOALLOC_MAX_CANDIDATES, else ABD_ERR_CAPACITY and nothing is launched.  The
 * first call on a device looks the array up (no allocation).  All calls on one device
 * share the array: calls that overlap in time on one device race. */
#define ABD_TOI_NOALLOC_MAX_CANDIDATES 1048576
int abd_toi_noalloc(abd_stream_t stream, const AbdScene* scene, const AbdParams* params, const double* x_t0,
                    const double* x_t1, const int32_t* pt, int64_t npt, const int32_t* ee, int64_t nee,
                    const int32_t* vg, int64_t nvg, double* toi);
/* cubic_roots.py:95-193 for n polynomials coef [n][4] (constant first) on [x0, x1]:
 * nroots [n], roots [n][3]. */
int abd_cubic_roots(abd_stream_t stream, const double* coef, int64_t n, double x0, double x1, int32_t* nroots,
                    double* roots);
/* ccd.py:133-186 pt_collision_time (kind 0) / ee_collision_time (kind 1) for n vertex
 * quadruples p_t0, p_t1 [n][4][3]: t [n]. */
int abd_collision_time(abd_stream_t stream, int32_t kind, const double* p_t0, const double* p_t1, int64_t n,
                       double* t);

/* ---- independent scenes (batching; no reference counterpart) ----
 * Per-scene forms of the reductions above for a scene with n_scenes > 0 and
 * scene_boff set; outputs are [n_scenes] device arrays, one workgroup per scene. */
int abd_scene_contact_energy(abd_stream_t stream, const AbdScene* scene, const AbdParams* params, const double* xk,
                             const int32_t* pt, int64_t npt, const int32_t* ee, int64_t nee, const int32_t* vg,
                             int64_t nvg, double* E);
int abd_scene_inertia_energy(abd_stream_t stream, const AbdScene* scene, const AbdParams* params, const double* Ak,
                             const double* pk, const double* A0, const double* p0, const double* Adot,
                             const double* pdot, double* E);
int abd_scene_toi(abd_stream_t stream, const AbdScene* scene, const AbdParams* params, const double* x_t0,
                  const double* x_t1, const int32_t* pt, int64_t npt, const int32_t* ee, int64_t nee,
                  const int32_t* vg, int64_t nvg, double* toi);
/* out[s] = a . b over the 12 n_s entries of scene s's bodies */
int abd_scene_dot(abd_stream_t stream, const AbdScene* scene, const double* a, const double* b, double* out);
/* abd_update_q with alpha[body_scene[b]] (device alpha [n_scenes]) */
int abd_update_q_scenes(abd_stream_t stream, const AbdScene* scene, const double* A, const double* p,
                        const double* dq, const double* alpha, double* A_out, double* p_out);
/* abd_bsr_solve for every scene's block-diagonal system, one workgroup per scene; a
 * scene with active[s] == 0 (device, nullable) is left untouched; iters/resid [n_scenes].
 * maxiter 0 -> each scene's system size. */
int abd_bsr_solve_scenes(abd_stream_t stream, int32_t method, const AbdScene* scene, int32_t max_scene_bodies,
                         const int32_t* row_ptr, const int32_t* cols, const int32_t* slots, const double* blocks,
                         const double* b, double* x, double tol, int32_t maxiter, int32_t check_every,
                         const uint8_t* active, int32_t* iters, double* resid);

/* ---- native frame driver ---- */
/* Device buffers of one simulation, owned by the caller (all device pointers). */
typedef struct AbdSimState {
    double *A, *p, *Ak, *pk, *A0, *p0, *Adot, *pdot; /* [Nb][3][3] / [Nb][3] */
    double *x, *xk, *x_view;                       /* [V][3] */
    double *g, *dq;                                /* [4 Nb][3] */
} AbdSimState;
/* abd.py:21-95 AffineBodySimulator.__init__: *sim receives a driver bound to `scene`
 * (its pointers must stay valid), `params` and `state`. */
int abd_sim_create(const AbdScene* scene, const AbdParams* params, const AbdSimState* state, void** sim);
/* abd.py:212-276 step(): one frame (Newton loop with culling, gh, solve, CCD and line
 * search) on `stream`.  stats (host, nullable) [10] receives running totals: [0] Newton
 * iterations, [1..3] gh PT / EE / VG candidates, [4] active contacts, [5] energy
 * evaluations, [6] solver iterations, [7] frames, [8..9] CCD PT / EE candidates.
 * With n_scenes > 1 every scene runs its own Newton loop, step length and stopping
 * test (one launch per stage covers all scenes) and [0] sums the scenes' iterations. */
int abd_sim_step(void* sim, abd_stream_t stream, int64_t* stats);
/* Adjoint support (no reference counterpart): assembles the Newton matrix H of
 * abd.py:230-243 (inertia + orthogonality + contact terms, contacts culled at x(q))
 * at the state's current (A, p), and solves H y = rhs (device [4 Nb][3]; y from zero)
 * with `method` and `tol` as abd_bsr_solve does.  With params.psd = 2, H is the exact
 * Hessian of the frame energy, so y = H^-1 rhs is the implicit-function-theorem
 * adjoint of one converged step.  iters / resid: host, nullable.  One scene only. */
int abd_sim_hessian_solve(void* sim, abd_stream_t stream, int32_t method, double tol, int32_t maxiter,
                          const double* rhs, double* y, int32_t* iters, double* resid);
/* Autotuner (no reference counterpart): times the collision-set culling at the driver's
 * current state for a small set of AbdTuning shapes (chunk 1024..16384, grid 4096..65536,
 * best of three each), then with those tiles the contact gradient and Hessian of that set
 * for pair_grid 1024 and 2048; keeps the fastest with abd_set_tuning and returns it (nullable).
 * Results are unchanged by the choice; call it once the scene has representative contacts. */
int abd_sim_autotune(void* sim, abd_stream_t stream, AbdTuning* chosen);
/* Newton iterations of each scene in the last abd_sim_step (iters: host [n_scenes];
 * with one scene, stats[0] holds the same count summed over frames). */
int abd_sim_scene_iters(void* sim, int32_t* iters);
/* Graph replay of the batched driver's Newton iterations (no reference counterpart; on by
 * default): an iteration whose launches are those of an earlier one (same candidate
 * superset, scratch and culling tiles) replays a HIP graph recorded from it instead of
 * launching each kernel; results are bit for bit those of the launches.  on = 0 launches
 * every kernel.  Iterations run with kernel timing on (abd_profile_begin) launch eagerly. */
int abd_sim_set_graphs(void* sim, int32_t on);
/* counts (host [3]): batched Newton iterations replayed from a graph, recorded into one
 * (and replayed), and launched kernel by kernel, since the simulator was created. */
int abd_sim_graph_counts(void* sim, int64_t* counts);
int abd_sim_destroy(void* sim);

/* ---- measurement ---- */
/* Kernel timing with HIP events recorded on each launch's stream, and the
 * algorithmic bytes each launch moves (no reference counterpart: it exists so
 * the benchmark can report achieved bandwidth per kernel).  Kernel ids:
 * 0 gh classify, 1 gh contact terms, 2 gh pair reduce, 3 gh body finalize,
 * 4 energy, 5 CCD, 6 cull filter, 7 cull count, 8 cull fill, 9 body pairs,
 * 10 vertex-ground cull, 11 Krylov solve, 12 vertex projection, 13 boxes,
 * 14 block-Jacobi diagonal-block inverses, 15 CCD root searches, 16 superset filter
 * (count, scans and write). */
int abd_profile_begin(void);
int abd_profile_end(void);
/* Kernel ids timed while profiling (bit k = id k; all by default): timing only the
 * kernels of interest keeps the event overhead out of a timed region. */
int abd_profile_select(uint32_t mask);
/* Synchronizes outstanding events; totals since abd_profile_begin(). */
int abd_profile_query(int32_t kernel, double* total_ms, int64_t* launches, double* bytes);

#ifdef __cplusplus
}
#endif
#endif /* ABD_HIP_H */
