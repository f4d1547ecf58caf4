// abd_sim.hip — native frame driver: AffineBodySimulator.step (abd.py:212-276)
// run from C++ over the stage entry points of this library.  The candidate lists a
// Newton iteration filters from the candidate superset keep their sizes on the device
// (every launch sized by the superset's, which bound them), so an iteration waits for
// the device once: the plain driver for the scalars its line search and convergence
// test branch on (time of impact, q.g, g.g, the first trial's energies), the batched
// driver, whose line searches and convergence tests run on the device, for the lists'
// sizes and the number of scenes still iterating.  The arithmetic is the same sequence
// of kernels as the Python mirror, so both give identical states.
#include <hip/hip_runtime.h>

#include <math.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "abd_internal.h"
#include "abd_profile.h"

namespace abd {

namespace {

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t cap = 0;  // elements
    int reserve(size_t n) {
        if (n <= cap && p) return ABD_OK;
        if (capturing()) return ABD_ERR_CAPACITY;  // no allocation inside a recorded graph
        note_reallocation();
        size_t want = std::max<size_t>(n + n / 2, 64);
        if (p) {
            ABD_CHECK(hipDeviceSynchronize());
            ABD_CHECK(hipFree(p));
            p = nullptr;
        }
        const hipError_t e = hipMalloc((void**)&p, want * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            if (std::getenv("ABD_DEBUG_ERRORS"))
                std::fprintf(stderr, "abd: simulator buffer of %zu bytes not allocated\n", want * sizeof(T));
            return hip_failure(e, __FILE__, __LINE__, "hipMalloc");
        }
        cap = want;
        return ABD_OK;
    }
    void release() {  // teardown: a failure leaves nothing to recover
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

struct BoxBuf {
    DevBuf<float> lo, hi;
    int reserve(size_t n) {
        int rc = lo.reserve(3 * n);
        return rc ? rc : hi.reserve(3 * n);
    }
};

// Body pairs (ij sorted by (i, j), ji the same pairs sorted by (j, i)) and the PT / EE / VG
// candidate tuples of one set of boxes.  The counts are exact when the set was culled
// (exact: consumers take them from here); a set filtered on the device has its sizes at dn
// ([nij, nij, npt, nee, nvg], device) and the counts here bound them.
// With `vids`, the PT and EE tuples' scene vertex ids (4 each, VertIds) are kept beside them.
struct CandSet {
    DevBuf<int32_t> ij, ji, ptl, eel, vgl;
    DevBuf<int32_t> ptv, eev;
    int64_t nij = 0, npt = 0, nee = 0, nvg = 0;
    bool exact = true;
    bool vids = false;
    const int64_t* dn = nullptr;
    const int64_t* counts() const { return exact ? nullptr : dn; }
    VertIds vertex_ids() const { return vids ? VertIds{ptv.p, eev.p} : VertIds{}; }
    int reserve_like(const CandSet& o, bool with_vids = false) {
        int rc = ij.reserve(2 * (size_t)o.nij + 2);
        if (!rc) rc = ji.reserve(2 * (size_t)o.nij + 2);
        if (!rc) rc = ptl.reserve(5 * (size_t)o.npt + 5);
        if (!rc) rc = eel.reserve(5 * (size_t)o.nee + 5);
        if (!rc) rc = vgl.reserve(2 * (size_t)o.nvg + 2);
        if (!rc && with_vids) rc = ptv.reserve(4 * (size_t)o.npt + 4);
        if (!rc && with_vids) rc = eev.reserve(4 * (size_t)o.nee + 4);
        return rc;
    }
    void release() {
        ij.release();
        ji.release();
        ptl.release();
        eel.release();
        vgl.release();
        ptv.release();
        eev.release();
    }
};

struct Sim {
    Arena arena;  // the stages' scratch during this simulator's calls
    AbdScene sc;
    AbdParams P;
    AbdSimState st;
    int nb, nv, ne, nf;
    BoxBuf body, tri, edge, pt;         // boxes of the last collision_set
    BoxBuf tbody;                       // swept body boxes (the primitives' are evaluated per entry)
    // the collision set at xk (gh, BSR structure) and the swept set of the last CCD step
    // (CCD, line search; the next iteration's collision set is filtered from it)
    CandSet col, sw;
    bool sw_valid = false;  // sw covers the current state (set by the CCD, cleared per frame)
    // candidate superset: the culling of the swept boxes of vertex ranges (a CCD step's two
    // ends widened by a skin); every later step whose two ends stay inside the ranges
    // takes its swept set from it by filtering instead of culling
    CandSet sup;
    BoxBuf sbody, stri, sedge, spt;  // the superset's boxes
    DevBuf<double> vlo, vhi;         // its vertex ranges
    bool sup_valid = false;
    bool col_fresh = false;  // col (with its structure) is the collision set at the current x
    DevBuf<int32_t> row_ptr, cols, slots;
    DevBuf<double> blocks;
    DevBuf<double> scal;  // [0] E contact, [1] E inertia, [2] toi, [3] dot, [4] resid
    DevBuf<int32_t> iters;
    DevBuf<int64_t> dcount;     // culling totals gathered for one read-back
    DevBuf<int64_t> cnt;        // device sizes of the filtered sets (kCnt*), read back per iteration
    int64_t* hcnt = nullptr;    // pinned copy of cnt
    int64_t gen = 0;            // the current step's number: the range test's flag value
    double* host = nullptr;     // pinned scalars
    int64_t* hcount = nullptr;  // pinned list counts
    int64_t stats[10] = {};
    // independent scenes (n_scenes > 1): per-scene device values and pinned host copies
    int ns = 1, max_scene_bodies = 0;
    DevBuf<double> sv;          // [kSv][ns]
    // the batched driver's Newton / line-search state on the device (ls_decide_kernel): rows
    // of the current iteration (lcur) and of the next, so that a step redone from a
    // recomputed swept set repeats the same decisions
    DevBuf<uint8_t> lsflag;     // [kLfRows][ns]
    DevBuf<double> lsval;       // [kLvRows][ns]
    DevBuf<int32_t> lsit;       // [2][ns] Newton iterations of each scene in this frame
    // the batched Newton iteration recorded as a HIP graph and replayed while the launches it
    // holds stay valid (GraphKey); a key met for the first time runs eagerly, the second
    // time it is recorded
    bool graphs = true;
    int64_t gcount[3] = {};  // iterations replayed, recorded (then replayed), launched eagerly
    hipGraphExec_t gexec = nullptr;
    hipStream_t cap_stream = nullptr;  // the recording stream (the caller's may be the null stream)
    uint64_t gkey[4] = {}, gseen[4] = {}, gfailed[4] = {};
    int64_t sup_gen = 0;  // advanced whenever the candidate superset is culled anew
    int32_t* hiters = nullptr;  // pinned [ns]: the iteration counts read back at a frame's end
    std::vector<int32_t> scene_it;  // Newton iterations of each scene in the last frame
    // line-search candidates evaluated together: states, vertices and per-candidate scalars
    DevBuf<double> cand_E;          // [kLsSpec][2 ns] trial energies
    DevBuf<double> cand_alpha;      // [kLsSpec][ns]
    DevBuf<int64_t> plan;           // reduction plan of the swept candidate set (CCD + line search)
    double* hcand = nullptr;        // pinned [kLsSpec][ns] alphas, then [kLsSpec][2 ns] energies
};

// Step lengths evaluated per line-search pass after the first: the trials alpha,
// alpha/2, ..., alpha/2^(kLsSpec-1) of the backtracking loop (abd.py:146-160), whose
// accept test depends only on each trial's own energy, so taking the first accepted
// trial of a pass gives exactly the sequential loop's step.
constexpr int kLsSpec = 12;

// Trials a pass needs from step length a on: a, a/2, ... up to the first one below
// alpha_min, which the loop accepts whatever its energy (at most kLsSpec; a pass that
// ends above alpha_min is followed by another).  With alpha <= 1 and alpha_min = 1e-3
// one pass after the first always decides.
__host__ __device__ inline int trials_to_floor(double a, double alpha_min) {
    int k = 1;
    while (k < kLsSpec && !(a < alpha_min)) {
        a *= 0.5;
        ++k;
    }
    return k;
}

// per-scene value rows of Sim::sv
enum { kSvEc = 0, kSvEi, kSvToi, kSvQtg, kSvGtg, kSvAlpha, kSv };
// Sim::cnt: the collision set's sizes, the swept set's, the active contacts of the gh
// stage, the flag of a step leaving the superset's ranges, the batched line search's
// outputs (scenes iterating after this iteration, scenes still searching, trials of the
// second pass and of the next, the solver's iterations summed over the iterating scenes);
// the host's copy holds a staging row for the exact sizes of a culled collision set after it
enum {
    kCntCol = 0,
    kCntSw = 8,
    kCntNact = 16,
    kCntRange = 17,
    kCntNewton = 18,
    kCntSearch = 19,
    kCntK1 = 20,
    kCntK2 = 21,
    kCntSolver = 22,
    kCntLen = 24
};
// rows of Sim::lsflag (iterating: the current and the next iteration's; searching) and of
// Sim::lsval (E0: the current and the next; the accepted trial's energy); the next
// iteration's rows are copied over the current ones at the start of every iteration
// (ls_commit_kernel)
enum { kLfNewton = 0, kLfSearch = 2, kLfRows };
enum { kLvE0 = 0, kLvE1 = 2, kLvRows };

int get_scalars(hipStream_t s, Sim& S, int first, int count) {
    ABD_CHECK(hipMemcpyAsync(S.host + first, S.scal.p + first, sizeof(double) * count, hipMemcpyDeviceToHost, s));
    ABD_CHECK(hipStreamSynchronize(s));
    return ABD_OK;
}

// The EE, PT and VG candidate lists of one set of boxes (culling.py:32-76), culled side
// by side so that their counts come back in two host synchronizations instead of five;
// the stages copy their totals into S.dcount, one read-back each.
int cull_lists(hipStream_t s, Sim& S, CandSet& C, const BoxBuf& body, const BoxBuf& pts, const BoxBuf& tris,
               const BoxBuf& edges) {
    PrimCull ce, cp;
    VgCull cv;
    const bool pairs = C.nij > 0;
    int64_t* dc = S.dcount.p;
    int rc;
    if (pairs) {
        if ((rc = prim_cull_begin(s, &S.sc, ABD_CULL_EE, C.ij.p, C.nij, edges.lo.p, edges.hi.p, edges.lo.p,
                                  edges.hi.p, 0, ce, dc + 0)))
            return rc;
        if ((rc = prim_cull_begin(s, &S.sc, ABD_CULL_PT, C.ij.p, C.nij, pts.lo.p, pts.hi.p, tris.lo.p, tris.hi.p, 1,
                                  cp, dc + 1)))
            return rc;
    }
    if ((rc = vg_cull_begin(s, &S.sc, body.lo.p, pts.lo.p, pts.hi.p, S.P.dhat, S.P.ground, cv, dc + 2))) return rc;
    int64_t* hv = S.hcount;
    if (pairs || cv.total) {
        ABD_CHECK(hipMemcpyAsync(hv, dc, 3 * sizeof(int64_t), hipMemcpyDeviceToHost, s));
        ABD_CHECK(hipStreamSynchronize(s));
    }
    const int64_t T_ee = pairs ? hv[0] : 0, T_pt = pairs ? hv[1] : 0;
    C.nvg = cv.total ? hv[2] : 0;
    if ((rc = C.vgl.reserve(2 * (size_t)C.nvg + 2))) return rc;
    if ((rc = vg_cull_fill(s, cv, C.vgl.p, (int64_t)(C.vgl.cap / 2), C.nvg))) return rc;
    C.nee = C.npt = 0;
    if (!pairs) return ABD_OK;
    if ((rc = prim_cull_count(s, ce, T_ee, dc + 3))) return rc;
    if ((rc = prim_cull_count(s, cp, T_pt, dc + 4))) return rc;
    if (T_ee || T_pt) {
        ABD_CHECK(hipMemcpyAsync(hv + 3, dc + 3, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, s));
        ABD_CHECK(hipStreamSynchronize(s));
    }
    C.nee = T_ee ? hv[3] : 0;
    C.npt = T_pt ? hv[4] : 0;
    if ((rc = C.eel.reserve(5 * (size_t)C.nee + 5))) return rc;
    if ((rc = C.ptl.reserve(5 * (size_t)C.npt + 5))) return rc;
    if ((rc = prim_cull_fill(s, ce, C.eel.p, (int64_t)(C.eel.cap / 5), C.nee))) return rc;
    return prim_cull_fill(s, cp, C.ptl.p, (int64_t)(C.ptl.cap / 5), C.npt);
}

// Body pairs of C (with the (j, i) order when with_ji), then the BSR structure of the
// collision set when `structure`.
int body_pairs(hipStream_t s, Sim& S, CandSet& C, const BoxBuf& body, bool with_ji) {
    for (;;) {
        int64_t cnt = 0;
        int rc = cull_bodies(s, &S.sc, body.lo.p, body.hi.p, C.ij.p, with_ji ? C.ji.p : nullptr,
                             (int64_t)(C.ij.cap / 2), &cnt);
        if (rc == ABD_ERR_OVERFLOW) {
            if ((rc = C.ij.reserve(2 * (size_t)cnt))) return rc;
            if ((rc = C.ji.reserve(2 * (size_t)cnt))) return rc;
            continue;
        }
        if (rc) return rc;
        C.nij = cnt;
        break;
    }
    return ABD_OK;
}

// abd.py:240-243: the slot layout of the collision set's body pairs (its pair count on the
// device when the set was filtered)
int structure_of(hipStream_t s, Sim& S) {
    int rc;
    if ((rc = S.cols.reserve(S.nb + 2 * (size_t)S.col.nij))) return rc;
    if ((rc = S.slots.reserve(S.nb + 2 * (size_t)S.col.nij))) return rc;
    return bsr_structure(s, S.nb, S.col.ij.p, S.col.ji.p, S.col.nij, S.col.counts(), S.row_ptr.p, S.cols.p,
                         S.slots.p);
}

// abd.py:187-207 on the vertices xv (xk of the reference's compute_energy, or x = xk at the
// start of a Newton iteration).  The culled counts are exact; with `structure` they are also
// staged to the device row the filtered sets use, so that every consumer of the collision
// set reads its sizes there.
int collision_set(hipStream_t s, Sim& S, bool structure, const double* xv) {
    const double dh = S.P.dhat * 0.5;
    int rc;
    if ((rc = abd_body_boxes(s, &S.sc, xv, dh, S.body.lo.p, S.body.hi.p))) return rc;
    if ((rc = body_pairs(s, S, S.col, S.body, structure))) return rc;
    S.col.exact = true;
    if (structure && (rc = structure_of(s, S))) return rc;
    if ((rc = abd_prim_boxes(s, &S.sc, ABD_KIND_TRI, xv, 0.0, S.tri.lo.p, S.tri.hi.p))) return rc;
    if ((rc = abd_prim_boxes(s, &S.sc, ABD_KIND_EDGE, xv, dh, S.edge.lo.p, S.edge.hi.p))) return rc;
    if ((rc = abd_prim_boxes(s, &S.sc, ABD_KIND_POINT, xv, S.P.dhat, S.pt.lo.p, S.pt.hi.p))) return rc;
    if ((rc = cull_lists(s, S, S.col, S.body, S.pt, S.tri, S.edge))) return rc;
    if (structure) {
        // the pinned staging row is rewritten only after the next read-back of the iteration
        int64_t* h = S.hcnt + kCntLen;
        h[0] = h[1] = S.col.nij;
        h[2] = S.col.npt;
        h[3] = S.col.nee;
        h[4] = S.col.nvg;
        ABD_CHECK(hipMemcpyAsync(S.cnt.p + kCntCol, h, 5 * sizeof(int64_t), hipMemcpyHostToDevice, s));
        S.col.dn = S.cnt.p + kCntCol;
        S.col.exact = false;  // consumers read the staged sizes (equal to the counts here)
    }
    return ABD_OK;
}

// The collision set at the vertices xv = x(q) of the accepted step (with its structure)
// filtered from a set that holds every pair and candidate of it (the swept set of the last
// CCD step: q = q_prev - alpha dq lies on that step and the swept boxes contain the boxes at
// every point of it; or the candidate superset while xv lies inside its ranges).  The result
// is the culling's own (keep_from_superset); its sizes stay on the device.
int collision_from(hipStream_t s, Sim& S, const CandSet& W, const double* xv) {
    const double dh = S.P.dhat * 0.5;
    int rc;
    // the body boxes at x (also the start half of the next swept body boxes); the primitive
    // boxes of collision_set's dilations are evaluated per entry of W by the filter: W holds
    // far fewer entries than the scene has primitives
    if ((rc = abd_body_boxes(s, &S.sc, xv, dh, S.body.lo.p, S.body.hi.p))) return rc;
    if ((rc = S.col.reserve_like(W))) return rc;
    const KeepInput in{W.ij.p, W.ji.p, W.ptl.p, W.eel.p, W.vgl.p, W.nij, W.npt, W.nee, W.nvg, W.counts(),
                       W.vertex_ids()};
    const KeepBoxes bx{S.body.lo.p, S.body.hi.p, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    const KeepOutput out{S.col.ij.p, S.col.ji.p, S.col.ptl.p, S.col.eel.p, S.col.vgl.p};
    const KeepAtX at{xv, nullptr, S.P.dhat, dh, 0.0};
    if ((rc = keep_from_superset(s, &S.sc, in, bx, S.P.dhat, S.P.ground, out, S.cnt.p + kCntCol, nullptr, false,
                                 &at)))
        return rc;
    S.col.nij = W.nij;
    S.col.npt = W.npt;
    S.col.nee = W.nee;
    S.col.nvg = W.nvg;
    S.col.dn = S.cnt.p + kCntCol;
    S.col.exact = false;
    return structure_of(s, S);
}
int collision_from_swept(hipStream_t s, Sim& S, const double* xv) { return collision_from(s, S, S.sw, xv); }

// A frame's first collision set: filtered from the candidate superset when the state lies
// inside its vertex ranges (boxes at any positions inside them lie inside its boxes, so the
// culling output is a subset of it), else culled.
int collision_at_frame_start(hipStream_t s, Sim& S, const double* xv) {
    if (S.sup_valid) {
        int rc;
        ABD_CHECK(hipMemsetAsync(S.dcount.p + 7, 0, sizeof(int64_t), s));
        if ((rc = within_ranges(s, 3 * (int64_t)S.nv, xv, xv, S.vlo.p, S.vhi.p, S.dcount.p + 7))) return rc;
        ABD_CHECK(hipMemcpyAsync(S.hcount + 7, S.dcount.p + 7, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        ABD_CHECK(hipStreamSynchronize(s));
        if (S.hcount[7] == 0) return collision_from(s, S, S.sup, xv);
    }
    return collision_set(s, S, true, xv);
}

// Skin of the candidate superset's vertex ranges, in units of dhat: wide enough that the
// later Newton steps of a frame (and of the next frames while the scenes rest) stay inside,
// so that a batch of scenes, which reuses the superset only when all of them do, skips
// most cullings; 30 measured best among 2, 10, 30 and 100 on the 256-scene workload.
constexpr double kSkinOverDhat = 30.0;

// abd.py:170-185, with both ends of every box dilated as in collision_set: one candidate
// set per Newton iteration that holds both the CCD candidates of the step (a pair that
// collides at some t has overlapping boxes at t, inside these) and every candidate of a
// collision set at any point of the step (x - alpha dq, alpha in [0, 1]).  The CCD then
// finds the same time of impact as the reference's trajectory lists (extra pairs have no
// verified root), and the line search's energies need no re-culling: a pair missing from
// the collision set at x - alpha dq has boxes apart, hence distance >= dhat and a zero
// barrier, so only the order of the nonzero terms' partial sums changes.
// The set is the culling output of the step's swept boxes, taken by filtering the
// candidate superset (keep_from_superset: the same predicates), valid while both ends of
// the step lie inside the superset's vertex ranges (every vertex of every scene).  Launched
// only: the range test's flag (cnt[kCntRange]) comes back with the iteration's scalars, and
// a step that left the ranges is redone after swept_set_recull.  Without a superset the
// superset is culled first (swept_set_recull's path).  The range test itself runs with
// the step's end (step_end).
int swept_set_launch(hipStream_t s, Sim& S);

// The superset culled anew around the current step (its ends widened by the skin; host
// synchronizations for the culling counts), then the swept set filtered from it.
int swept_set_recull(hipStream_t s, Sim& S) {
    const double dh = S.P.dhat * 0.5;
    const AbdSimState& st = S.st;
    const int64_t n3 = 3 * (int64_t)S.nv;
    int rc;
    S.sup_valid = false;
    ++S.sup_gen;
    if ((rc = vertex_ranges(s, n3, st.x, st.xk, kSkinOverDhat * S.P.dhat, S.vlo.p, S.vhi.p))) return rc;
    if ((rc = swept_boxes(s, &S.sc, -1, S.vlo.p, S.vhi.p, dh, S.sbody.lo.p, S.sbody.hi.p))) return rc;
    // with the (j, i) order: the swept set (and the collision sets after it) keep it
    if ((rc = body_pairs(s, S, S.sup, S.sbody, true))) return rc;
    if ((rc = swept_boxes(s, &S.sc, ABD_KIND_TRI, S.vlo.p, S.vhi.p, 0.0, S.stri.lo.p, S.stri.hi.p))) return rc;
    if ((rc = swept_boxes(s, &S.sc, ABD_KIND_EDGE, S.vlo.p, S.vhi.p, dh, S.sedge.lo.p, S.sedge.hi.p))) return rc;
    if ((rc = swept_boxes(s, &S.sc, ABD_KIND_POINT, S.vlo.p, S.vhi.p, S.P.dhat, S.spt.lo.p, S.spt.hi.p))) return rc;
    if ((rc = cull_lists(s, S, S.sup, S.sbody, S.spt, S.stri, S.sedge))) return rc;
    S.sup.exact = true;
    // the tuples' vertex ids, read by the filters and, carried into the swept set, by the CCD
    // and the line search (the superset serves many steps)
    if ((rc = S.sup.ptv.reserve(4 * (size_t)S.sup.npt + 4))) return rc;
    if ((rc = S.sup.eev.reserve(4 * (size_t)S.sup.nee + 4))) return rc;
    if ((rc = cand_vertex_ids(s, &S.sc, 0, S.sup.ptl.p, S.sup.npt, S.sup.ptv.p))) return rc;
    if ((rc = cand_vertex_ids(s, &S.sc, 1, S.sup.eel.p, S.sup.nee, S.sup.eev.p))) return rc;
    S.sup.vids = true;
    S.sup_valid = true;
    return swept_set_launch(s, S);
}

int swept_set_launch(hipStream_t s, Sim& S) {
    const double dh = S.P.dhat * 0.5;
    const AbdSimState& st = S.st;
    int rc;
    if (!S.sup_valid) return swept_set_recull(s, S);
    // body boxes: the half at x is this iteration's collision set's (same dilation; reading
    // it beats a second pass over every vertex).  Primitive swept boxes: evaluated by the
    // filter per entry of the superset from both ends (far fewer entries than primitives)
    if ((rc = swept_boxes(s, &S.sc, -1, st.x, st.xk, dh, S.tbody.lo.p, S.tbody.hi.p, S.body.lo.p, S.body.hi.p)))
        return rc;
    if ((rc = S.sw.reserve_like(S.sup, S.sup.vids))) return rc;
    const CandSet& W = S.sup;
    const KeepInput in{W.ij.p, W.ji.p, W.ptl.p, W.eel.p, W.vgl.p, W.nij, W.npt, W.nee, W.nvg, W.counts(),
                       W.vertex_ids()};
    const KeepBoxes bx{S.tbody.lo.p, S.tbody.hi.p, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    KeepOutput o{S.sw.ij.p, S.sw.ji.p, S.sw.ptl.p, S.sw.eel.p, S.sw.vgl.p};
    if (W.vids) {
        o.ptv = S.sw.ptv.p;
        o.eev = S.sw.eev.p;
    }
    const KeepAtX at{st.x, st.xk, S.P.dhat, dh, 0.0};
    if ((rc = keep_from_superset(s, &S.sc, in, bx, S.P.dhat, S.P.ground, o, S.cnt.p + kCntSw, nullptr, false, &at)))
        return rc;
    S.sw.nij = W.nij;
    S.sw.npt = W.npt;
    S.sw.nee = W.nee;
    S.sw.nvg = W.nvg;
    S.sw.dn = S.cnt.p + kCntSw;
    S.sw.exact = false;
    S.sw.vids = W.vids;
    S.sw_valid = true;
    return ABD_OK;
}

// The end of the step (xk = x(q - dq) for the scenes in `active`, or every scene), with the
// range test of both ends against the superset's ranges when there is a superset.
int step_end(hipStream_t s, Sim& S, const uint8_t* active) {
    const AbdSimState& st = S.st;
    ++S.gen;
    return project_step(s, &S.sc, st.A, st.p, st.dq, nullptr, 1.0, active, st.xk, st.x,
                        S.sup_valid ? S.vlo.p : nullptr, S.sup_valid ? S.vhi.p : nullptr, S.cnt.p + kCntRange,
                        S.gen);
}

// After the iteration's synchronization: true when the step left the superset's ranges, so
// that the swept set must be recomputed.
bool swept_left_ranges(const Sim& S) { return S.hcnt[kCntRange] == S.gen; }

// The lists' sizes for the statistics (after the iteration's synchronization).
void count_stats(Sim& S) {
    const int64_t* h = S.hcnt;
    S.stats[1] += h[kCntCol + 2];
    S.stats[2] += h[kCntCol + 3];
    S.stats[3] += h[kCntCol + 4];
    S.stats[4] += h[kCntNact];
    S.stats[8] += h[kCntSw + 2];
    S.stats[9] += h[kCntSw + 3];
}

// abd.py:133-141 at a frame's start (alpha = 0: xk = x), which also gives the first Newton
// iteration's collision set with its structure
int compute_energy(hipStream_t s, Sim& S, double alpha, double* E) {
    int rc;
    const AbdSimState& st = S.st;
    if ((rc = abd_update_q(s, S.nb, st.A, st.p, st.dq, alpha, st.Ak, st.pk))) return rc;
    if ((rc = abd_project_vertices(s, &S.sc, st.Ak, st.pk, st.xk))) return rc;
    if ((rc = collision_at_frame_start(s, S, st.xk))) return rc;
    S.col_fresh = alpha == 0.0;
    const CandSet& C = S.col;
    if ((rc = contact_energy(s, &S.sc, &S.P, 1, st.xk, C.ptl.p, C.npt, C.eel.p, C.nee, C.vgl.p, C.nvg, C.counts(),
                             S.scal.p)))
        return rc;
    if ((rc = abd_inertia_energy(s, &S.sc, &S.P, st.Ak, st.pk, st.A0, st.p0, st.Adot, st.pdot, S.scal.p + 1)))
        return rc;
    if ((rc = get_scalars(s, S, 0, 2))) return rc;
    S.stats[5] += 1;
    *E = S.host[0] + S.host[1];
    return ABD_OK;
}

// abd.py:133-141 inside the line search, for K trial step lengths at once (energies to
// E[K], one host round trip): the energy at q - alpha dq over the swept candidate set of
// this iteration (swept_set_launch), without re-culling
int line_energy_multi(hipStream_t s, Sim& S, const double* alphas, int K, double* E) {
    int rc;
    const AbdSimState& st = S.st;
    double* ha = S.hcand;
    double* he = S.hcand + kLsSpec;
    for (int k = 0; k < K; ++k) ha[k] = alphas[k];
    ABD_CHECK(hipMemcpyAsync(S.cand_alpha.p, ha, sizeof(double) * K, hipMemcpyHostToDevice, s));
    if ((rc = line_trials(s, &S.sc, &S.P, 1, K, S.cand_alpha.p, st.A, st.p, st.dq, st.A0, st.p0, st.Adot, st.pdot,
                          S.sw.ptl.p, S.sw.npt, S.sw.eel.p, S.sw.nee, S.sw.vgl.p, S.sw.nvg, S.plan.p, S.cand_E.p,
                          S.sw.counts(), S.sw.vertex_ids())))
        return rc;
    ABD_CHECK(hipMemcpyAsync(he, S.cand_E.p, sizeof(double) * 2 * K, hipMemcpyDeviceToHost, s));
    ABD_CHECK(hipStreamSynchronize(s));
    for (int k = 0; k < K; ++k) E[k] = he[2 * k] + he[2 * k + 1];
    S.stats[5] += K;
    return ABD_OK;
}

// abd.py:278-303 and the line search's first trial, launched only, then one read-back: the
// step's swept set (ccd: xk = the end of the full step), its reduction plan, the TOI into
// scal[2], q.g and g.g, the first trial's energies (its step length from the TOI, on the
// device), the solver's iteration count and the sets' sizes.  A step that left the
// superset's ranges is recomputed from a superset culled around it.
int ccd_and_first_trial(hipStream_t s, Sim& S) {
    int rc;
    const AbdSimState& st = S.st;
    const int nb = S.nb;
    for (int attempt = 0;; ++attempt) {
        if ((rc = attempt == 0 ? swept_set_launch(s, S) : swept_set_recull(s, S))) return rc;
        const CandSet& W = S.sw;
        // one reduction plan of the swept set serves the TOI and every line-search trial
        if ((rc = reduce_plan(s, &S.sc, &S.P, 1, W.ptl.p, W.npt, W.eel.p, W.nee, W.vgl.p, W.nvg, S.plan.p,
                              W.counts())))
            return rc;
        // the TOI, and from it on the device the first trial's step length
        if ((rc = scene_toi(s, &S.sc, &S.P, 1, st.x, st.xk, W.ptl.p, W.npt, W.eel.p, W.nee, W.vgl.p, W.nvg,
                            S.plan.p, S.scal.p + 2, nullptr, W.counts(), S.cand_alpha.p, W.vertex_ids())))
            return rc;
        // q.g (abd.py:148) and g.g (abd.py:265): g and dq are final here
        if ((rc = dots(s, &S.sc, 1, st.dq, st.g, st.g, st.g, S.scal.p + 3, S.scal.p + 5))) return rc;
        if ((rc = line_trials(s, &S.sc, &S.P, 1, 1, S.cand_alpha.p, st.A, st.p, st.dq, st.A0, st.p0, st.Adot,
                              st.pdot, W.ptl.p, W.npt, W.eel.p, W.nee, W.vgl.p, W.nvg, S.plan.p, S.cand_E.p,
                              W.counts(), W.vertex_ids())))
            return rc;
        ABD_CHECK(hipMemcpyAsync(S.hcand + kLsSpec, S.cand_E.p, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
        ABD_CHECK(hipMemcpyAsync(S.host + 6, S.iters.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        ABD_CHECK(hipMemcpyAsync(S.hcnt, S.cnt.p, kCntLen * sizeof(int64_t), hipMemcpyDeviceToHost, s));
        if ((rc = get_scalars(s, S, 2, 4))) return rc;
        // a recomputed swept set lies inside the superset culled around this very step
        if (attempt > 0 || !swept_left_ranges(S)) return ABD_OK;
    }
}

int step(hipStream_t s, Sim& S) {
    const AbdSimState& st = S.st;
    const AbdParams& P = S.P;
    const int nb = S.nb;
    int rc;
    // q_gets_q0 (abd.py:124-128)
    ABD_CHECK(hipMemcpyAsync(st.A, st.A0, 72 * (size_t)nb, hipMemcpyDeviceToDevice, s));
    ABD_CHECK(hipMemcpyAsync(st.p, st.p0, 24 * (size_t)nb, hipMemcpyDeviceToDevice, s));
    ABD_CHECK(hipMemcpyAsync(st.Ak, st.A0, 72 * (size_t)nb, hipMemcpyDeviceToDevice, s));
    ABD_CHECK(hipMemcpyAsync(st.pk, st.p0, 24 * (size_t)nb, hipMemcpyDeviceToDevice, s));
    double E0;
    S.sw_valid = false;  // a frame starts from the state the caller holds
    if ((rc = compute_energy(s, S, 0.0, &E0))) return rc;
    int it = 0;
    for (;;) {
        // (Ak, pk) = (A, p) here: the collision set is culled on x
        if ((rc = abd_project_vertices(s, &S.sc, st.A, st.p, st.x))) return rc;
        if (S.col_fresh)
            S.col_fresh = false;  // the frame-start energy's set: the same positions
        else if ((rc = S.sw_valid ? collision_from_swept(s, S, st.x) : collision_at_frame_start(s, S, st.x)))
            return rc;
        const size_t nslot = nb + 2 * (size_t)S.col.nij;
        if ((rc = S.blocks.reserve(nslot * 16 * 9))) return rc;
        ABD_CHECK(hipMemsetAsync(S.blocks.p, 0, sizeof(double) * nslot * 16 * 9, s));
        // g: overwritten by the inertia gradient (abd.py:236-237 zeroes it first)
        if ((rc = inertia_gh(s, &S.sc, &P, st.A, st.p, st.A0, st.p0, st.Adot, st.pdot, st.g, S.blocks.p)))
            return rc;
        const CandSet& C = S.col;
        const GhLists L{C.ptl.p, C.npt, C.eel.p, C.nee, C.vgl.p, C.nvg, C.nij, C.counts()};
        if ((rc = contact_gh(s, &S.sc, &P, st.x, L, S.row_ptr.p, S.slots.p, st.g, S.blocks.p, nullptr, nullptr,
                             S.cnt.p + kCntNact)))
            return rc;
        if ((rc = abd_bsr_solve(s, P.solver, nb, S.row_ptr.p, S.cols.p, S.slots.p, S.blocks.p, st.g, st.dq,
                                P.solver_tol, P.solver_maxiter, 10, S.iters.p, nullptr)))
            return rc;
        // CCD over the full step (xk = its end) and the first trial (abd.py:278-303)
        if ((rc = step_end(s, S, nullptr))) return rc;
        if ((rc = ccd_and_first_trial(s, S))) return rc;
        count_stats(S);
        S.stats[5] += 1;
        const double qTg = S.host[3], gTg = S.host[5];
        int32_t sit;
        memcpy(&sit, S.host + 6, sizeof(int32_t));
        S.stats[6] += sit;
        // abd.py:146-160 line search (Armijo test as written)
        double alpha = first_step_length(S.host[2]), E1 = 0.0;
        for (int pass = 0;; ++pass) {
            // the first trial alone (usually accepted), then the rest down to alpha_min
            const int K = pass == 0 ? 1 : trials_to_floor(alpha, P.alpha_min);
            double trial[kLsSpec], Ek[kLsSpec];
            for (int k = 0; k < K; ++k) trial[k] = k == 0 ? alpha : trial[k - 1] * 0.5;
            if (pass == 0)
                Ek[0] = S.hcand[kLsSpec] + S.hcand[kLsSpec + 1];
            else if ((rc = line_energy_multi(s, S, trial, K, Ek)))
                return rc;
            int acc = -1;
            for (int k = 0; k < K && acc < 0; ++k) {
                const bool wolfe = Ek[k] < E0 + P.c1 * trial[k] * qTg;
                if (wolfe || trial[k] < P.alpha_min) acc = k;
            }
            if (acc >= 0) {
                alpha = trial[acc];
                E1 = Ek[acc];
                break;
            }
            alpha = trial[K - 1] * 0.5;
        }
        E0 = E1;
        // q and (Ak, pk) = q - alpha dq, the accepted trial
        if ((rc = update_q_dual(s, nb, st.A, st.p, st.dq, alpha, st.A, st.p, st.Ak, st.pk))) return rc;
        ++it;
        if (gTg < P.tol || it >= P.max_iter) break;
    }
    if ((rc = abd_update_q0qdot(s, nb, P.dt, st.A, st.p, st.A0, st.p0, st.Adot, st.pdot))) return rc;
    if ((rc = abd_project_vertices(s, &S.sc, st.A0, st.p0, st.x_view))) return rc;
    S.stats[0] += it;
    S.stats[7] += 1;
    S.scene_it.assign(1, it);
    return ABD_OK;
}

// ---------------------------------------------------------------- batched
// Many independent scenes advanced together: one launch per stage covers every
// scene, and each scene keeps its own Newton loop, step length and stopping
// test, exactly the per-scene sequence of step() above.  A scene that has
// converged is masked out of the solve and takes zero-length steps.  The line
// searches and the stopping tests run on the device (ls_decide_kernel), so an
// iteration waits for the device once.

// The device rows of the batched Newton / line-search state for the current iteration.
struct LsDev {
    int ns;
    const uint8_t* newton;  // scenes iterating (the solver's mask of this iteration)
    uint8_t* newton_nxt;
    uint8_t* search;        // scenes whose line search has not accepted a trial
    const double* E0;       // energy at the iteration's start (abd.py:225)
    double* E0_nxt;
    double* E1;             // the accepted trial's energy
    double* dalpha;         // the accepted step length (0: a scene not iterating)
    const int32_t* it;      // Newton iterations of the frame so far
    int32_t* it_nxt;
    const double* qtg;      // q.g and g.g of the iteration (abd.py:148, 265)
    const double* gtg;
    const int32_t* iters;   // the solver's iterations
    int64_t* cnt;           // Sim::cnt
};

LsDev ls_dev(Sim& S) {
    const int ns = S.ns, c = 0, n = 1;
    LsDev L;
    L.ns = ns;
    L.newton = S.lsflag.p + (size_t)(kLfNewton + c) * ns;
    L.newton_nxt = S.lsflag.p + (size_t)(kLfNewton + n) * ns;
    L.search = S.lsflag.p + (size_t)kLfSearch * ns;
    L.E0 = S.lsval.p + (size_t)(kLvE0 + c) * ns;
    L.E0_nxt = S.lsval.p + (size_t)(kLvE0 + n) * ns;
    L.E1 = S.lsval.p + (size_t)kLvE1 * ns;
    L.dalpha = S.sv.p + (size_t)kSvAlpha * ns;
    L.it = S.lsit.p + (size_t)c * ns;
    L.it_nxt = S.lsit.p + (size_t)n * ns;
    L.qtg = S.sv.p + (size_t)kSvQtg * ns;
    L.gtg = S.sv.p + (size_t)kSvGtg * ns;
    L.iters = S.iters.p;
    L.cnt = S.cnt.p;
    return L;
}

constexpr int kLsThreads = 256;

__device__ __forceinline__ int64_t ls_block_reduce(int64_t v, bool take_max, int64_t* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int w = kLsThreads / 2; w > 0; w >>= 1) {
        if (t < w) sh[t] = take_max ? (sh[t] > sh[t + w] ? sh[t] : sh[t + w]) : sh[t] + sh[t + w];
        __syncthreads();
    }
    const int64_t r = sh[0];
    __syncthreads();
    return r;
}

// One pass of every scene's line search (abd.py:146-160; the Armijo test as written, a
// trial below alpha_min accepted whatever its energy), then the next pass's trials:
// alpha [K][ns] holds this pass's step lengths (negative: not evaluated) and E [K][2 ns]
// their contact and inertia energies (line_trials); a scene takes the first trial that
// passes.  first: the pass of the first trial, over the iterating scenes (the others get
// step length 0).  Then alpha [kLsSpec][ns] receives the next pass's trials of the scenes
// still searching: a, a/2, ... from half the last rejected trial down to the first below
// alpha_min (trials_to_floor), -1 elsewhere.  When no scene is searching any more, the
// stopping test of every iterating scene (abd.py:265-268: g.g < tol or the iteration cap)
// writes the next iteration's rows.  One workgroup; the host's step() makes the same
// decisions with the same arithmetic.
__global__ __launch_bounds__(kLsThreads) void ls_decide_kernel(LsDev L, double c1, double alpha_min, double tol,
                                                                int max_iter, int first, int K,
                                                                double* __restrict__ alpha,
                                                                const double* __restrict__ E) {
    __shared__ int64_t sh[kLsThreads];
    const int ns = L.ns;
    int64_t nsearch = 0, kmax = 0;
    for (int s = threadIdx.x; s < ns; s += kLsThreads) {
        bool searching = first ? L.newton[s] != 0 : L.search[s] != 0;
        if (first && !searching) L.dalpha[s] = 0.0;
        double a0 = 0.0;  // the next pass's first trial
        if (searching) {
            bool acc = false;
            double last = 0.0;
            const double e0 = L.E0[s], qtg = L.qtg[s];
            for (int k = 0; k < K && !acc; ++k) {
                const double a = alpha[(int64_t)k * ns + s];
                if (a < 0.0) break;  // past this scene's trials
                last = a;
                const double e1 = E[2 * (int64_t)ns * k + s] + E[2 * (int64_t)ns * k + ns + s];
                const bool wolfe = e1 < e0 + c1 * a * qtg;
                if (wolfe || a < alpha_min) {
                    acc = true;
                    L.dalpha[s] = a;
                    L.E1[s] = e1;
                }
            }
            if (acc) {
                searching = false;
            } else {
                a0 = last * 0.5;
            }
        }
        L.search[s] = searching ? 1 : 0;
        const int need = searching ? trials_to_floor(a0, alpha_min) : 0;
        double ak = a0;  // a0, a0 / 2, ... in a register (no re-read of what was just written)
        for (int k = 0; k < kLsSpec; ++k) {
            alpha[(int64_t)k * ns + s] = k >= need ? -1.0 : ak;
            ak *= 0.5;
        }
        nsearch += searching ? 1 : 0;
        kmax = need > kmax ? need : kmax;
    }
    nsearch = ls_block_reduce(nsearch, false, sh);
    kmax = ls_block_reduce(kmax, true, sh);
    int64_t nact = 0, solver = 0;
    if (nsearch == 0) {  // uniform: every line search has ended
        for (int s = threadIdx.x; s < ns; s += kLsThreads) {
            uint8_t nn = 0;
            double e0 = L.E0[s];
            int32_t itn = L.it[s];
            if (L.newton[s]) {
                e0 = L.E1[s];
                itn += 1;
                solver += L.iters[s] > 0 ? L.iters[s] : 0;
                nn = (L.gtg[s] < tol || itn >= max_iter) ? 0 : 1;
            }
            L.E0_nxt[s] = e0;
            L.it_nxt[s] = itn;
            L.newton_nxt[s] = nn;
            nact += nn;
        }
        nact = ls_block_reduce(nact, false, sh);
        solver = ls_block_reduce(solver, false, sh);
    }
    if (threadIdx.x == 0) {
        L.cnt[kCntSearch] = nsearch;
        L.cnt[first ? kCntK1 : kCntK2] = kmax;
        if (nsearch == 0) {
            L.cnt[kCntNewton] = nact;
            L.cnt[kCntSolver] = solver;
        }
    }
}

// A frame's start: E0 = the energy at q0 (abd.py:224-225), every scene iterating, no
// iterations yet.
__global__ __launch_bounds__(kLsThreads) void ls_frame_start_kernel(int ns, const double* __restrict__ Ec,
                                                                     const double* __restrict__ Ei,
                                                                     double* __restrict__ E0,
                                                                     uint8_t* __restrict__ newton,
                                                                     int32_t* __restrict__ it) {
    for (int s = blockIdx.x * kLsThreads + threadIdx.x; s < ns; s += gridDim.x * kLsThreads) {
        E0[s] = Ec[s] + Ei[s];
        newton[s] = 1;
        it[s] = 0;
    }
}

// The start of a Newton iteration: the rows the last iteration's decisions wrote become the
// current ones, and the range test's flag is cleared for this step.
__global__ __launch_bounds__(kLsThreads) void ls_commit_kernel(int ns, uint8_t* __restrict__ flag,
                                                                double* __restrict__ val,
                                                                int32_t* __restrict__ it,
                                                                int64_t* __restrict__ range_flag) {
    for (int s = blockIdx.x * kLsThreads + threadIdx.x; s < ns; s += gridDim.x * kLsThreads) {
        flag[(int64_t)kLfNewton * ns + s] = flag[(int64_t)(kLfNewton + 1) * ns + s];
        val[(int64_t)kLvE0 * ns + s] = val[(int64_t)(kLvE0 + 1) * ns + s];
        it[s] = it[ns + s];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *range_flag = 0;
}

// per-scene energies at q = q_base - alpha_s dq (abd.py:133-141): sv rows Ec, Ei
int compute_energy_scenes(hipStream_t s, Sim& S, const double* dalpha) {
    int rc;
    const AbdSimState& st = S.st;
    const int ns = S.ns;
    if ((rc = abd_update_q_scenes(s, &S.sc, st.A, st.p, st.dq, dalpha, st.Ak, st.pk))) return rc;
    if ((rc = abd_project_vertices(s, &S.sc, st.Ak, st.pk, st.xk))) return rc;
    // at a frame's start (every alpha 0: zero-length steps are exact copies, xk = x): the
    // first Newton iteration's collision set too, with its structure
    if ((rc = collision_at_frame_start(s, S, st.xk))) return rc;
    S.col_fresh = true;
    const CandSet& C = S.col;
    if ((rc = contact_energy(s, &S.sc, &S.P, ns, st.xk, C.ptl.p, C.npt, C.eel.p, C.nee, C.vgl.p, C.nvg, C.counts(),
                             S.sv.p + kSvEc * ns)))
        return rc;
    if ((rc = abd_scene_inertia_energy(s, &S.sc, &S.P, st.Ak, st.pk, st.A0, st.p0, st.Adot, st.pdot,
                                       S.sv.p + kSvEi * ns)))
        return rc;
    S.stats[5] += 1;
    return ABD_OK;
}

// One line-search pass over every scene: the energies at q - alpha dq of the trials alpha
// [K][ns] (device) over the swept candidate set of this iteration, into cand_E [K][2 ns]
int line_pass(hipStream_t s, Sim& S, int K) {
    const AbdSimState& st = S.st;
    return line_trials(s, &S.sc, &S.P, S.ns, K, S.cand_alpha.p, st.A, st.p, st.dq, st.A0, st.p0, st.Adot, st.pdot,
                       S.sw.ptl.p, S.sw.npt, S.sw.eel.p, S.sw.nee, S.sw.vgl.p, S.sw.nvg, S.plan.p, S.cand_E.p,
                       S.sw.counts(), S.sw.vertex_ids());
}

int ls_decide(hipStream_t s, Sim& S, const LsDev& L, bool first, int K) {
    hipLaunchKernelGGL(ls_decide_kernel, dim3(1), dim3(kLsThreads), 0, s, L, S.P.c1, S.P.alpha_min, S.P.tol,
                       (int)S.P.max_iter, first ? 1 : 0, K, S.cand_alpha.p, (const double*)S.cand_E.p);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

int read_counts(hipStream_t s, Sim& S) {
    ABD_CHECK(hipMemcpyAsync(S.hcnt, S.cnt.p, kCntLen * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    ABD_CHECK(hipStreamSynchronize(s));
    return ABD_OK;
}

// The range test of the batched step's end (a vertex outside the superset's ranges writes 1
// over the 0 of ls_commit_kernel).
int step_end_scenes(hipStream_t s, Sim& S, const uint8_t* active) {
    const AbdSimState& st = S.st;
    return project_step(s, &S.sc, st.A, st.p, st.dq, nullptr, 1.0, active, st.xk, st.x,
                        S.sup_valid ? S.vlo.p : nullptr, S.sup_valid ? S.vhi.p : nullptr, S.cnt.p + kCntRange, 1);
}

// The swept set's TOI, q.g and g.g, and both line-search passes with their decisions, then
// the counts' read-back (launched; the caller waits).
int ccd_and_search_scenes(hipStream_t s, Sim& S, const LsDev& L) {
    const AbdSimState& st = S.st;
    const AbdParams& P = S.P;
    const int ns = S.ns;
    const CandSet& W = S.sw;
    int rc;
    if ((rc = reduce_plan(s, &S.sc, &P, ns, W.ptl.p, W.npt, W.eel.p, W.nee, W.vgl.p, W.nvg, S.plan.p, W.counts())))
        return rc;
    // the TOI of the scenes still iterating and from it the first trial's step lengths; in the
    // same final pass q.g and g.g of the Armijo and stopping tests (abd.py:148,265: g and dq
    // are final here)
    SceneDots qg;
    qg.sboff = ns > 1 ? S.sc.scene_boff : nullptr;
    qg.n = 12 * (int64_t)S.sc.n_bodies;
    qg.a = st.dq;
    qg.b = st.g;
    qg.c = st.g;
    qg.d = st.g;
    qg.out1 = S.sv.p + kSvQtg * ns;
    qg.out2 = S.sv.p + kSvGtg * ns;
    if ((rc = scene_toi(s, &S.sc, &P, ns, st.x, st.xk, W.ptl.p, W.npt, W.eel.p, W.nee, W.vgl.p, W.nvg, S.plan.p,
                        S.sv.p + kSvToi * ns, L.newton, W.counts(), S.cand_alpha.p, W.vertex_ids(), &qg)))
        return rc;
    // the line search: the first trial, then every remaining halving of the scenes that
    // rejected it, each decided on the device
    if ((rc = line_pass(s, S, 1))) return rc;
    if ((rc = ls_decide(s, S, L, true, 1))) return rc;
    if ((rc = line_pass(s, S, kLsSpec))) return rc;
    if ((rc = ls_decide(s, S, L, false, kLsSpec))) return rc;
    ABD_CHECK(hipMemcpyAsync(S.hcnt, S.cnt.p, kCntLen * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    return ABD_OK;
}

// One batched Newton iteration up to the read-back of its counts (abd.py:224-271): the last
// iteration's accepted steps applied, then collision set, gradient and Hessian, solve, step
// end with the range test, swept set, TOI and line search.  With no host branch on device
// values, the same launches every time for the same sizes: the form a HIP graph records.
int iteration_scenes(hipStream_t s, Sim& S) {
    const AbdSimState& st = S.st;
    const AbdParams& P = S.P;
    const int nb = S.nb, ns = S.ns;
    const LsDev L = ls_dev(S);
    int rc;
    // q and (Ak, pk) = the accepted trial of every scene (step length 0: an exact copy)
    if ((rc = update_q_scenes_dual(s, &S.sc, st.A, st.p, st.dq, L.dalpha, st.A, st.p, st.Ak, st.pk))) return rc;
    hipLaunchKernelGGL(ls_commit_kernel, dim3((unsigned)ceil_div(ns, kLsThreads)), dim3(kLsThreads), 0, s, ns,
                       S.lsflag.p, S.lsval.p, S.lsit.p, S.cnt.p + kCntRange);
    ABD_CHECK_LAUNCH();
    // (Ak, pk) = (A, p) here: the collision set is culled on x
    if ((rc = abd_project_vertices(s, &S.sc, st.A, st.p, st.x))) return rc;
    if (S.col_fresh)
        S.col_fresh = false;  // the frame-start energy's set: the same positions
    else if ((rc = S.sw_valid ? collision_from_swept(s, S, st.x) : collision_at_frame_start(s, S, st.x)))
        return rc;
    const size_t nslot = nb + 2 * (size_t)S.col.nij;
    if ((rc = S.blocks.reserve(nslot * 16 * 9))) return rc;
    if ((rc = zero_fill(s, S.blocks.p, sizeof(double) * nslot * 16 * 9))) return rc;
    // g: overwritten by the inertia gradient (abd.py:236-237 zeroes it first)
    if ((rc = inertia_gh(s, &S.sc, &P, st.A, st.p, st.A0, st.p0, st.Adot, st.pdot, st.g, S.blocks.p))) return rc;
    const CandSet& C = S.col;
    const GhLists GL{C.ptl.p, C.npt, C.eel.p, C.nee, C.vgl.p, C.nvg, C.nij, C.counts()};
    if ((rc = contact_gh(s, &S.sc, &P, st.x, GL, S.row_ptr.p, S.slots.p, st.g, S.blocks.p, nullptr, nullptr,
                         S.cnt.p + kCntNact)))
        return rc;
    if ((rc = abd_bsr_solve_scenes(s, P.solver, &S.sc, S.max_scene_bodies, S.row_ptr.p, S.cols.p, S.slots.p,
                                   S.blocks.p, st.g, st.dq, P.solver_tol, P.solver_maxiter, 10, L.newton, S.iters.p,
                                   nullptr)))
        return rc;
    // CCD over the full steps of the scenes still iterating (abd.py:278-303)
    if ((rc = step_end_scenes(s, S, L.newton))) return rc;
    if ((rc = swept_set_launch(s, S))) return rc;
    return ccd_and_search_scenes(s, S, L);
}

// What a recorded iteration depends on besides device values: the candidate superset (its
// sizes bound every launch), the scratch pointers and the culling tiles.
void graph_key(const Sim& S, uint64_t key[4]) {
    const AbdTuning t = tuning();
    key[0] = (uint64_t)S.sup_gen + 1;
    key[1] = alloc_generation();
    key[2] = ((uint64_t)(uint32_t)t.cull_chunk << 32) | (uint32_t)t.cull_grid;
    key[3] = (uint64_t)S.sup.nij ^ ((uint64_t)S.sup.npt << 20) ^ ((uint64_t)S.sup.nee << 40) ^ (uint64_t)S.sup.nvg;
}
bool same_key(const uint64_t a[4], const uint64_t b[4]) { return !memcmp(a, b, 4 * sizeof(uint64_t)); }

// Records iteration_scenes into S.gexec (updated in place when the topology allows), on a
// stream of the simulator's own (the legacy null stream cannot record); nothing runs.  A
// failure leaves no error state behind, and the caller launches the iteration instead.
int record_iteration(Sim& S) {
    if (!S.cap_stream) {
        const hipError_t ce = hipStreamCreateWithFlags(&S.cap_stream, hipStreamNonBlocking);
        if (ce != hipSuccess) {
            S.cap_stream = nullptr;
            (void)hipGetLastError();
            return -static_cast<int>(ce);
        }
    }
    hipStream_t s = S.cap_stream;
    hipGraph_t g = nullptr;
    const hipError_t be = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (be != hipSuccess) {
        (void)hipGetLastError();
        return -static_cast<int>(be);
    }
    set_capturing(true);
    int rc = iteration_scenes(s, S);
    set_capturing(false);
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (rc || e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return rc ? rc : -static_cast<int>(e == hipSuccess ? hipErrorUnknown : e);
    }
    if (S.gexec) {
        hipGraphNode_t err_node = nullptr;
        hipGraphExecUpdateResult res;
        if (hipGraphExecUpdate(S.gexec, g, &err_node, &res) != hipSuccess || res != hipGraphExecUpdateSuccess) {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(S.gexec);
            S.gexec = nullptr;
        }
    }
    hipError_t ie = hipSuccess;
    if (!S.gexec) ie = hipGraphInstantiate(&S.gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
        S.gexec = nullptr;
        (void)hipGetLastError();
        return -static_cast<int>(ie);
    }
    return ABD_OK;
}

int step_scenes(hipStream_t s, Sim& S) {
    const AbdSimState& st = S.st;
    const AbdParams& P = S.P;
    const int nb = S.nb, ns = S.ns;
    int rc;
    double* dalpha = S.sv.p + kSvAlpha * ns;
    ABD_CHECK(hipMemcpyAsync(st.A, st.A0, 72 * (size_t)nb, hipMemcpyDeviceToDevice, s));
    ABD_CHECK(hipMemcpyAsync(st.p, st.p0, 24 * (size_t)nb, hipMemcpyDeviceToDevice, s));
    ABD_CHECK(hipMemcpyAsync(st.Ak, st.A0, 72 * (size_t)nb, hipMemcpyDeviceToDevice, s));
    ABD_CHECK(hipMemcpyAsync(st.pk, st.p0, 24 * (size_t)nb, hipMemcpyDeviceToDevice, s));
    if ((rc = zero_fill(s, dalpha, sizeof(double) * ns))) return rc;
    S.sw_valid = false;  // a frame starts from the state the caller holds
    if ((rc = compute_energy_scenes(s, S, dalpha))) return rc;
    {
        // the first iteration's rows, taken over by its ls_commit_kernel
        const LsDev L = ls_dev(S);
        hipLaunchKernelGGL(ls_frame_start_kernel, dim3((unsigned)ceil_div(ns, kLsThreads)), dim3(kLsThreads), 0, s,
                           ns, (const double*)(S.sv.p + kSvEc * ns), (const double*)(S.sv.p + kSvEi * ns), L.E0_nxt,
                           L.newton_nxt, L.it_nxt);
        ABD_CHECK_LAUNCH();
    }
    int64_t n_active = ns;
    while (n_active > 0) {
        const LsDev L = ls_dev(S);
        // a steady iteration (collision set from the last swept set, superset kept) replays
        // its recorded graph; the first time a key is met the iteration runs eagerly, the
        // second time it is recorded (every buffer then holds its size for these launches)
        const bool steady = S.graphs && S.sw_valid && !S.col_fresh && S.sup_valid && !prof_enabled();
        uint64_t key[4];
        graph_key(S, key);
        bool ran = false;
        if (steady && S.gexec && same_key(key, S.gkey)) {
            ABD_CHECK(hipGraphLaunch(S.gexec, s));
            ran = true;
            ++S.gcount[0];
        } else if (steady && same_key(key, S.gseen) && !same_key(key, S.gfailed)) {
            if (record_iteration(S) == ABD_OK) {
                ABD_CHECK(hipGraphLaunch(S.gexec, s));
                memcpy(S.gkey, key, sizeof(key));
                ran = true;
                ++S.gcount[1];
            } else {
                memcpy(S.gfailed, key, sizeof(key));  // this key runs eagerly
            }
        }
        if (!ran) {
            ++S.gcount[2];
            if ((rc = iteration_scenes(s, S))) return rc;
            if (steady) {
                graph_key(S, key);  // the eager run may have grown buffers
                memcpy(S.gseen, key, sizeof(key));
            }
        }
        ABD_CHECK(hipStreamSynchronize(s));
        if (S.hcnt[kCntRange] != 0) {
            // the step left the superset's ranges: the swept set (and all after it) again from
            // a superset culled around this very step, which holds it
            if ((rc = swept_set_recull(s, S))) return rc;
            if ((rc = ccd_and_search_scenes(s, S, L))) return rc;
            ABD_CHECK(hipStreamSynchronize(s));
        }
        count_stats(S);
        S.stats[5] += 1 + S.hcnt[kCntK1];
        // further passes: only when alpha_min lies more than kLsSpec halvings below a first trial
        while (S.hcnt[kCntSearch] > 0) {
            const int64_t K = S.hcnt[kCntK2];
            if ((rc = line_pass(s, S, kLsSpec))) return rc;
            if ((rc = ls_decide(s, S, L, false, kLsSpec))) return rc;
            if ((rc = read_counts(s, S))) return rc;
            S.stats[5] += K;
        }
        S.stats[0] += n_active;
        S.stats[6] += S.hcnt[kCntSolver];
        n_active = S.hcnt[kCntNewton];
    }
    if ((rc = update_q_scenes_dual(s, &S.sc, st.A, st.p, st.dq, dalpha, st.A, st.p, st.Ak, st.pk))) return rc;
    if ((rc = abd_update_q0qdot(s, nb, P.dt, st.A, st.p, st.A0, st.p0, st.Adot, st.pdot))) return rc;
    if ((rc = abd_project_vertices(s, &S.sc, st.A0, st.p0, st.x_view))) return rc;
    ABD_CHECK(hipMemcpyAsync(S.hiters, S.lsit.p + ns, sizeof(int32_t) * ns, hipMemcpyDeviceToHost, s));
    ABD_CHECK(hipStreamSynchronize(s));
    S.scene_it.assign(S.hiters, S.hiters + ns);
    S.stats[7] += 1;
    return ABD_OK;
}

}  // namespace

}  // namespace abd

using namespace abd;

extern "C" {

int abd_sim_create(const AbdScene* scene, const AbdParams* params, const AbdSimState* state, void** out) {
    if (!scene || !params || !state || !out || scene->n_bodies <= 0) return ABD_ERR_ARG;
    if (params->psd < 0 || params->psd > 3) return ABD_ERR_ARG;
    Sim* S = new (std::nothrow) Sim();
    if (!S) return ABD_ERR_ARG;
    S->sc = *scene;
    S->P = *params;
    S->st = *state;
    S->nb = scene->n_bodies;
    S->nv = scene->n_verts;
    S->ne = scene->n_edges;
    S->nf = scene->n_tris;
    int rc = 0;
    const size_t nb = S->nb;
    BoxBuf* b4[] = {&S->body, &S->tbody};
    for (BoxBuf* b : b4)
        if (!rc) rc = b->reserve(nb);
    if (!rc) rc = S->tri.reserve(S->nf + 1);
    if (!rc) rc = S->edge.reserve(S->ne + 1);
    if (!rc) rc = S->pt.reserve(S->nv);
    if (!rc) rc = S->sbody.reserve(nb);
    if (!rc) rc = S->stri.reserve(S->nf + 1);
    if (!rc) rc = S->sedge.reserve(S->ne + 1);
    if (!rc) rc = S->spt.reserve(S->nv);
    if (!rc) rc = S->vlo.reserve(3 * (size_t)S->nv + 3);
    if (!rc) rc = S->vhi.reserve(3 * (size_t)S->nv + 3);
    for (CandSet* C : {&S->col, &S->sw, &S->sup}) {
        if (!rc) rc = C->ij.reserve(2 * 256);
        if (!rc) rc = C->ji.reserve(2 * 256);
        if (!rc) rc = C->ptl.reserve(5 * 4096);
        if (!rc) rc = C->eel.reserve(5 * 4096);
        if (!rc) rc = C->vgl.reserve(2 * 1024);
    }
    if (!rc) rc = S->row_ptr.reserve(nb + 1);
    if (!rc) rc = S->cols.reserve(nb + 512);
    if (!rc) rc = S->slots.reserve(nb + 512);
    if (!rc) rc = S->blocks.reserve((nb + 512) * 144);
    if (!rc) rc = S->scal.reserve(8);
    if (!rc) rc = S->iters.reserve(4);
    if (!rc && hipHostMalloc((void**)&S->host, 8 * sizeof(double), hipHostMallocDefault) != hipSuccess) rc = ABD_ERR_ARG;
    if (!rc && hipHostMalloc((void**)&S->hcount, 8 * sizeof(int64_t), hipHostMallocDefault) != hipSuccess)
        rc = ABD_ERR_ARG;
    {
        const int nsc = scene->n_scenes > 1 ? scene->n_scenes : 1;
        if (!rc) rc = S->cand_E.reserve(2 * (size_t)nsc * kLsSpec);
        if (!rc) rc = S->cand_alpha.reserve((size_t)nsc * kLsSpec);
        if (!rc) rc = S->plan.reserve(reduce_plan_elems(nsc));
        if (!rc) rc = S->dcount.reserve(8);
        if (!rc) rc = S->cnt.reserve(kCntLen);
        if (!rc && hipMemset(S->cnt.p, 0, sizeof(int64_t) * S->cnt.cap) != hipSuccess) rc = ABD_ERR_ARG;
        if (!rc && hipHostMalloc((void**)&S->hcnt, 2 * kCntLen * sizeof(int64_t), hipHostMallocDefault) != hipSuccess)
            rc = ABD_ERR_ARG;
        if (!rc && hipHostMalloc((void**)&S->hcand, sizeof(double) * 3 * nsc * kLsSpec, hipHostMallocDefault) !=
                       hipSuccess)
            rc = ABD_ERR_ARG;
    }
    if (!rc && scene->n_scenes > 1) {
        // scene ranges from the host copy of scene_boff
        const int ns = scene->n_scenes;
        std::vector<int32_t> sb(ns + 1);
        if (!scene->scene_boff || !scene->body_scene ||
            hipMemcpy(sb.data(), scene->scene_boff, sizeof(int32_t) * (ns + 1), hipMemcpyDeviceToHost) != hipSuccess)
            rc = ABD_ERR_ARG;
        for (int k = 0; !rc && k < ns; ++k) {
            if (sb[k + 1] < sb[k]) rc = ABD_ERR_ARG;
            S->max_scene_bodies = std::max(S->max_scene_bodies, sb[k + 1] - sb[k]);
        }
        if (!rc && (sb[0] != 0 || sb[ns] != scene->n_bodies)) rc = ABD_ERR_ARG;
        S->ns = ns;
        if (!rc) rc = S->sv.reserve((size_t)kSv * ns);
        if (!rc) rc = S->lsflag.reserve((size_t)kLfRows * ns);
        if (!rc) rc = S->lsval.reserve((size_t)kLvRows * ns);
        if (!rc) rc = S->lsit.reserve(2 * (size_t)ns);
        if (!rc) rc = S->iters.reserve(ns);
        if (!rc) rc = S->scal.reserve(8 + (size_t)ns);
        if (!rc && hipHostMalloc((void**)&S->hiters, sizeof(int32_t) * ns, hipHostMallocDefault) != hipSuccess)
            rc = ABD_ERR_ARG;
    }
    if (rc) {
        abd_sim_destroy(S);
        return rc;
    }
    *out = S;
    return ABD_OK;
}

int abd_sim_step(void* sim, abd_stream_t stream, int64_t* stats) {
    if (!sim) return ABD_ERR_ARG;
    Sim& S = *static_cast<Sim*>(sim);
    ArenaScope scope(&S.arena);
    int rc = S.ns > 1 ? step_scenes(to_stream(stream), S) : step(to_stream(stream), S);
    if (stats)
        for (int k = 0; k < 10; ++k) stats[k] = S.stats[k];
    return rc;
}

int abd_sim_hessian_solve(void* sim, abd_stream_t stream, int32_t method, double tol, int32_t maxiter,
                          const double* rhs, double* y, int32_t* iters, double* resid) {
    if (!sim || !rhs || !y) return ABD_ERR_ARG;
    if (method != ABD_SOLVER_BICGSTAB && method != ABD_SOLVER_CG_JACOBI && method != ABD_SOLVER_CG_BLOCK)
        return ABD_ERR_ARG;
    Sim& S = *static_cast<Sim*>(sim);
    if (S.ns > 1) return ABD_ERR_ARG;
    ArenaScope scope(&S.arena);
    hipStream_t s = to_stream(stream);
    const AbdSimState& st = S.st;
    int rc;
    // the Newton matrix of abd.py:230-243 at q = (A, p), contacts from x(q)
    if ((rc = abd_project_vertices(s, &S.sc, st.A, st.p, st.x))) return rc;
    if ((rc = collision_set(s, S, true, st.x))) return rc;
    const size_t nslot = S.nb + 2 * (size_t)S.col.nij;
    if ((rc = S.blocks.reserve(nslot * 16 * 9))) return rc;
    ABD_CHECK(hipMemsetAsync(S.blocks.p, 0, sizeof(double) * nslot * 16 * 9, s));
    ABD_CHECK(hipMemsetAsync(st.g, 0, sizeof(double) * 12 * S.nb, s));
    if ((rc = abd_inertia_hessian(s, &S.sc, &S.P, st.A, S.blocks.p))) return rc;
    int64_t nact = 0;
    if ((rc = abd_contact_gh(s, &S.sc, &S.P, st.x, S.col.ptl.p, S.col.npt, S.col.eel.p, S.col.nee, S.col.vgl.p, S.col.nvg, S.col.nij,
                             S.row_ptr.p, S.slots.p, st.g, S.blocks.p, nullptr, &nact)))
        return rc;
    ABD_CHECK(hipMemsetAsync(y, 0, sizeof(double) * 12 * S.nb, s));
    if ((rc = abd_bsr_solve(s, method, S.nb, S.row_ptr.p, S.cols.p, S.slots.p, S.blocks.p, rhs, y, tol, maxiter, 10,
                            S.iters.p, S.scal.p + 4)))
        return rc;
    if (iters) ABD_CHECK(hipMemcpyAsync(iters, S.iters.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    if (resid) ABD_CHECK(hipMemcpyAsync(resid, S.scal.p + 4, sizeof(double), hipMemcpyDeviceToHost, s));
    ABD_CHECK(hipStreamSynchronize(s));
    return ABD_OK;
}

int abd_sim_autotune(void* sim, abd_stream_t stream, AbdTuning* chosen) {
    if (!sim) return ABD_ERR_ARG;
    Sim& S = *static_cast<Sim*>(sim);
    ArenaScope scope(&S.arena);
    hipStream_t s = to_stream(stream);
    const AbdTuning saved = tuning();
    int rc;
    // the candidate lists at the current state (xk) are what every collision_set of a frame culls
    if ((rc = abd_project_vertices(s, &S.sc, S.st.Ak, S.st.pk, S.st.xk))) return rc;
    hipEvent_t e0, e1;
    ABD_CHECK(hipEventCreate(&e0));
    ABD_CHECK(hipEventCreate(&e1));
    AbdTuning best = saved;
    float best_ms = 1e30f;
    const int chunks[] = {1024, 2048, 4096, 8192, 16384};
    const int grids[] = {4096, 16384, 65536};
    for (int c : chunks)
        for (int g : grids) {
            AbdTuning t{c, g, saved.contact_bound_mib, saved.pair_grid, saved.toi_grid};
            if ((rc = abd_set_tuning(&t))) break;
            float ms_min = 1e30f;
            for (int rep = 0; rep < 3 && !rc; ++rep) {
                hipError_t he = hipEventRecord(e0, s);
                if (he != hipSuccess) {
                    rc = -static_cast<int>(he);
                    break;
                }
                rc = collision_set(s, S, false, S.st.xk);
                if (!rc && (he = hipEventRecord(e1, s)) != hipSuccess) rc = -static_cast<int>(he);
                float ms = 0.f;
                if (!rc && hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess)
                    ms_min = ms < ms_min ? ms : ms_min;
            }
            if (rc) break;
            if (ms_min < best_ms) {
                best_ms = ms_min;
                best = t;
            }
        }
    // stage D's grid, with the chosen tiles: the contact Hessian of the collision set at this
    // state (its structure built once), best of three per grid; g and the Hessian slots are
    // scratch that the next iteration rewrites
    if (!rc && (rc = abd_set_tuning(&best)) == ABD_OK && (rc = collision_set(s, S, true, S.st.xk)) == ABD_OK) {
        const size_t nslot = S.sc.n_bodies + 2 * (size_t)S.col.nij;
        rc = S.blocks.reserve(nslot * 16 * 9);
        const CandSet& C = S.col;
        const GhLists L{C.ptl.p, C.npt, C.eel.p, C.nee, C.vgl.p, C.nvg, C.nij, C.counts()};
        // one or two rounds of the workgroups the device holds at once (four a CU): larger
        // grids, and smaller ones, measured faster on an early frame's few contacts and slower
        // on the later frames' many
        const int pgrids[] = {1024, 2048};
        float bestp = 1e30f;
        int32_t pbest = best.pair_grid;
        for (int pg : pgrids) {
            if (rc) break;
            AbdTuning t = best;
            t.pair_grid = pg;
            if ((rc = abd_set_tuning(&t))) break;
            float ms_min = 1e30f;
            for (int rep = 0; rep < 3 && !rc; ++rep) {
                hipError_t he = hipEventRecord(e0, s);
                if (he != hipSuccess) {
                    rc = -static_cast<int>(he);
                    break;
                }
                rc = contact_gh(s, &S.sc, &S.P, S.st.xk, L, S.row_ptr.p, S.slots.p, S.st.g, S.blocks.p, nullptr,
                                nullptr, S.cnt.p + kCntNact);
                if (!rc && (he = hipEventRecord(e1, s)) != hipSuccess) rc = -static_cast<int>(he);
                float ms = 0.f;
                if (!rc && hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess)
                    ms_min = ms < ms_min ? ms : ms_min;
            }
            if (!rc && ms_min < bestp) {
                bestp = ms_min;
                pbest = pg;
            }
        }
        best.pair_grid = pbest;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    abd_set_tuning(rc ? &saved : &best);
    if (rc) return rc;
    if (chosen) *chosen = best;
    return ABD_OK;
}

int abd_sim_set_graphs(void* sim, int32_t on) {
    if (!sim) return ABD_ERR_ARG;
    static_cast<Sim*>(sim)->graphs = on != 0;
    return ABD_OK;
}

int abd_sim_graph_counts(void* sim, int64_t* counts) {
    if (!sim || !counts) return ABD_ERR_ARG;
    const Sim& S = *static_cast<Sim*>(sim);
    for (int k = 0; k < 3; ++k) counts[k] = S.gcount[k];
    return ABD_OK;
}

int abd_sim_scene_iters(void* sim, int32_t* iters) {
    if (!sim || !iters) return ABD_ERR_ARG;
    const Sim& S = *static_cast<Sim*>(sim);
    for (int k = 0; k < S.ns; ++k) iters[k] = k < (int)S.scene_it.size() ? S.scene_it[k] : 0;
    return ABD_OK;
}

int abd_sim_destroy(void* sim) {
    if (!sim) return ABD_OK;
    Sim* S = static_cast<Sim*>(sim);
    (void)hipDeviceSynchronize();  // queued work may still use the buffers released below
    if (S->gexec) (void)hipGraphExecDestroy(S->gexec);
    if (S->cap_stream) (void)hipStreamDestroy(S->cap_stream);
    (void)arena_free(S->arena);
    for (BoxBuf* b : {&S->body, &S->tri, &S->edge, &S->pt, &S->tbody, &S->sbody,
                      &S->stri, &S->sedge, &S->spt}) {
        b->lo.release();
        b->hi.release();
    }
    S->col.release();
    S->sw.release();
    S->sup.release();
    S->vlo.release();
    S->vhi.release();
    S->row_ptr.release();
    S->cols.release();
    S->slots.release();
    S->blocks.release();
    S->scal.release();
    S->iters.release();
    S->sv.release();
    S->lsflag.release();
    S->lsval.release();
    S->lsit.release();
    S->cand_E.release();
    S->cand_alpha.release();
    S->plan.release();
    S->dcount.release();
    S->cnt.release();
    if (S->hcand) (void)hipHostFree(S->hcand);
    for (void* h : {(void*)S->host, (void*)S->hcount, (void*)S->hcnt, (void*)S->hiters})
        if (h) (void)hipHostFree(h);
    delete S;
    return ABD_OK;
}

}  // extern "C"
