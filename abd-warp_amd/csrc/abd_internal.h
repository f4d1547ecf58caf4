// abd_internal.h — host-side internals shared by the libabd_hip.so translation
// units: error handling, the per-device scratch arena, and launchers of the
// generic scan / reduction kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/abd_hip.h"

namespace abd {

// The error code of a failed HIP call (-hipError_t); with ABD_DEBUG_ERRORS set in the
// environment it also prints the call's site and the runtime's message to stderr.
int hip_failure(hipError_t e, const char* file, int line, const char* expr);
#define ABD_CHECK(expr)                                                                  \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess) return ::abd::hip_failure(_e, __FILE__, __LINE__, #expr); \
    } while (0)
#define ABD_CHECK_LAUNCH() ABD_CHECK(hipGetLastError())

inline hipStream_t to_stream(abd_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// Per-device scratch arena: `slots` independent growable device buffers plus a
// small pinned host buffer for count read-back.  Growing synchronizes the
// device first, so a buffer still in use by queued work is never freed early.
constexpr int kArenaSlots = 41;
enum ArenaSlot {
    kSlotScanA = 0,
    kSlotScanB,
    kSlotReduce,
    kSlotFlags,
    kSlotList,
    kSlotHull1,
    kSlotHull2,
    kSlotFilter,
    kSlotFilterCount,
    kSlotChunk,
    kSlotChunkCount,
    kSlotContact,
    kSlotPair,
    kSlotMisc,
    kSlotSolver,
    kSlotSpare,
    kSlotSolveBytes,
    kSlotPlan,
    kSlotPrecond,
    // a second primitive-culling workspace (culls run side by side) and the VG culling's
    kSlotHull1B,
    kSlotHull2B,
    kSlotFilterB,
    kSlotFilterCountB,
    kSlotScanAB,
    kSlotScanBB,
    kSlotChunkCountB,
    kSlotChunkB,
    kSlotVgA,
    kSlotVgB,
    kSlotPairSeg,  // per body pair and contact type: the active contacts' [begin, end)
    kSlotKeep,     // keep_from_superset: block counts and offsets
    kSlotKeepMap,  // keep_from_superset: old -> new pair index
    kSlotKeepCount,
    kSlotLdlt,        // direct solve: the factors of the systems
    kSlotLdltStatus,  // direct solve: per-system status when the caller takes no iteration counts
    kSlotPairChunk,   // contact stage D: per-pair work-item counts and offsets
    kSlotPairPart,    // contact stage D: the work items' partial sums
    kSlotKeepFlags,   // keep_from_superset: each entry's predicate
    kSlotToiQueue,    // CCD: the candidates whose cubic has a bracketed root
    kSlotToiSlots,    // CCD: per-scene minima of the queued candidates and the queue count
    kSlotProfBytes,   // profiling: algorithmic byte counts computed on the device
};
// Launch shapes that change no result (abd_set_tuning); process-wide.
AbdTuning tuning();
// Staged culling (abd_cull.hip): begin / count / fill leave their counts on the device
// (total_chunks, total_pairs, total) so several cullings share one host synchronization;
// `set` 0 or 1 picks one of two independent workspaces.
struct PrimSets {
    const int32_t* off1;  // per-body offsets of set 1 (PT points / EE edges)
    const int32_t* off2;  // set 2 (PT triangles / EE edges)
    const float *lo1, *hi1, *lo2, *hi2;
    const float *hlo1, *hhi1, *hlo2, *hhi2;  // per-body hulls
};
struct PrimCull {
    int set = 0, pt = 0;
    const int32_t* ij = nullptr;
    PrimSets S{};
    int64_t nunits = 0, maxs = 0, T = 0;
    int32_t* flist = nullptr;
    int64_t *fcnt = nullptr, *nchunk = nullptr, *choff = nullptr, *ccnt = nullptr, *ccoff = nullptr;
    const int64_t* total_chunks = nullptr;  // device, after begin
    const int64_t* total_pairs = nullptr;   // device, after count
    AbdTuning tune{};
    unsigned grid = 0;
};
struct VgCull {
    const AbdScene* sc = nullptr;
    const float *blo = nullptr, *plo = nullptr, *phi = nullptr;
    double yth = 0.0;
    float fg = 0.f;
    int64_t nblk = 0;
    int64_t *bcnt = nullptr, *boff = nullptr;
    const int64_t* total = nullptr;  // device, after begin
};
// total_copy (device, nullable): a second place for the stage's total, so that a caller
// can gather the totals of several cullings into one buffer and read them with one copy.
int prim_cull_begin(hipStream_t s, const AbdScene* sc, int mode, const int32_t* ij, int64_t nij, const float* lo1,
                    const float* hi1, const float* lo2, const float* hi2, int set, PrimCull& C,
                    int64_t* total_copy = nullptr);
int prim_cull_count(hipStream_t s, PrimCull& C, int64_t T, int64_t* total_copy = nullptr);
int prim_cull_fill(hipStream_t s, PrimCull& C, int32_t* out, int64_t cap, int64_t n);
int vg_cull_begin(hipStream_t s, const AbdScene* sc, const float* blo, const float* plo, const float* phi, double dhat,
                  double ground, VgCull& C,
                  int64_t* total_copy = nullptr);
int vg_cull_fill(hipStream_t s, VgCull& C, int32_t* out, int64_t cap, int64_t n);

// Swept boxes dilated at both ends (kind ABD_KIND_*, or -1 for body boxes): the union of
// the dilated boxes at x0 and at x1.  plo / phi (body boxes only; both or neither): the
// body boxes at x0 with the same dilation, already computed (read instead of recomputed;
// the same values).
int swept_boxes(hipStream_t s, const AbdScene* sc, int kind, const double* x0, const double* x1, double dil,
                float* lo, float* hi, const float* plo = nullptr, const float* phi = nullptr);
// Vertex ranges of a candidate superset (n3 = 3 * vertices): lo = min(x0, x1) - skin,
// hi = max(x0, x1) + skin per coordinate.  Boxes of any positions inside these ranges
// lie inside the swept boxes of (lo, hi) (every box formula is monotone in the
// positions), so their culling output is a subset of that superset's.
int vertex_ranges(hipStream_t s, int64_t n3, const double* x0, const double* x1, double skin, double* lo,
                  double* hi);
// *flag (device, 0 before) becomes 1 when some coordinate of x0 or x1 leaves [lo, hi].
int within_ranges(hipStream_t s, int64_t n3, const double* x0, const double* x1, const double* lo, const double* hi,
                  int64_t* flag);
// K line-search trials q - alpha dq at once (alpha [K][S] on the device, S = 1: the
// whole system as one scene; a negative alpha skips that scene's trial and gives 0):
// per-scene contact energies to E[2 S k + s] and inertia energies to E[2 S k + S + s].
// dn (here and below, nullable): the lists' sizes on the device, [nij, nij, npt, nee, nvg],
// the counts passed then bounding them.
// The scene vertex ids of each PT / EE candidate of a list (4 per tuple, in the order
// fetch_pt_idx / fetch_ee_idx give them), stored beside it so that a kernel reads them
// instead of gathering them through the topology; either may be null.
struct VertIds {
    const int32_t* pt = nullptr;
    const int32_t* ee = nullptr;
};
int line_trials(hipStream_t s, const AbdScene* sc, const AbdParams* P, int S, int K, const double* alpha,
                const double* A, const double* p, const double* dq, const double* A0, const double* p0,
                const double* Adot, const double* pdot, const int32_t* pt, int64_t npt, const int32_t* ee, int64_t nee,
                const int32_t* vg, int64_t nvg, const int64_t* plan, double* E, const int64_t* dn = nullptr,
                VertIds vids = {});
// The per-scene chunk plan of the energy / TOI reductions over one set of candidate lists
// (plan: reduce_plan_elems(S) int64 on the device), made once and passed as `plan` to every
// reduction over those lists (nullptr: each makes its own).
size_t reduce_plan_elems(int S);
int reduce_plan(hipStream_t s, const AbdScene* sc, const AbdParams* P, int S, const int32_t* pt, int64_t npt,
                const int32_t* ee, int64_t nee, const int32_t* vg, int64_t nvg, int64_t* plan,
                const int64_t* dn = nullptr);
// Per-scene TOI (abd_scene_toi; S = 1: the whole system as one scene) with an optional plan;
// active (nullable, [S]): scenes whose TOI is wanted, the others report 1 unevaluated.
// Per-scene dot products a.b and c.d (seg_dot2_kernel's) that the TOI's final per-scene pass computes
// after the scene's TOI when out1 is set (the driver's q.g and g.g, one launch less).
struct SceneDots {
    const int32_t* sboff = nullptr;
    int64_t n = 0;
    const double *a = nullptr, *b = nullptr, *c = nullptr, *d = nullptr;
    double *out1 = nullptr, *out2 = nullptr;
};
int scene_toi(hipStream_t s, const AbdScene* sc, const AbdParams* P, int S, const double* x_t0, const double* x_t1,
              const int32_t* pt, int64_t npt, const int32_t* ee, int64_t nee, const int32_t* vg, int64_t nvg,
              const int64_t* plan, double* toi, const uint8_t* active, const int64_t* dn = nullptr,
              double* first_alpha = nullptr, VertIds vids = {}, const SceneDots* dots = nullptr);
// The line search's first step length from a TOI (abd.py:300-302): clamped to [0, 1], x 0.9
// below 1.  scene_toi's first_alpha (device, nullable) receives it per scene (-1 for the
// scenes not active).
__host__ __device__ inline double first_step_length(double t) {
    t = t < 1.0 ? t : 1.0;  // std::min(1.0, t)
    t = 0.0 < t ? t : 0.0;  // std::max(0.0, t)
    if (t < 1.0) t *= 0.9;
    return t;
}
// out1 = a . b and out2 = c . d per scene (S = 1: over all 12 nb rows), one launch
int dots(hipStream_t s, const AbdScene* sc, int S, const double* a, const double* b, const double* c,
         const double* d, double* out1, double* out2);
// The end of a step: x_out = x(q - alpha dq), the arithmetic of update_q + project_vertices
// with alpha_host (alpha_dev, active null), or of update_q_scenes + project_vertices with
// alpha_dev [n_scenes] or alpha_host for the scenes with active[s] (0 for the others).  With
// lo / hi, *flag (device) becomes gen when a coordinate of x_out or of x_other leaves them.
int project_step(hipStream_t s, const AbdScene* sc, const double* A, const double* p, const double* dq,
                 const double* alpha_dev, double alpha_host, const uint8_t* active, double* x_out,
                 const double* x_other, const double* lo, const double* hi, int64_t* flag, int64_t gen);
// abd_contact_energy (S = 1) / abd_scene_contact_energy with the lists' sizes on the device.
int contact_energy(hipStream_t s, const AbdScene* sc, const AbdParams* P, int S, const double* xk, const int32_t* pt,
                   int64_t npt, const int32_t* ee, int64_t nee, const int32_t* vg, int64_t nvg, const int64_t* dn,
                   double* E);
// abd_update_q / abd_update_q_scenes writing the result to a second state as well (when
// A_out2, p_out2 are given).
int update_q_dual(hipStream_t s, int nb, const double* A, const double* p, const double* dq, double alpha,
                  double* A_out, double* p_out, double* A_out2, double* p_out2);
int update_q_scenes_dual(hipStream_t s, const AbdScene* sc, const double* A, const double* p, const double* dq,
                         const double* alpha, double* A_out, double* p_out, double* A_out2, double* p_out2);
// Collision set from a superset of it (keep_from_superset, abd_cull.hip): the superset's
// body pairs (ij, and ji when given), PT / EE / VG tuples, the current body / point /
// edge / triangle boxes; outputs sized like the inputs (no aliasing).  The five totals
// (nij, nij of ji, npt, nee, nvg) go to dcount (device) and, after one synchronization,
// to hcount (host, pinned).
// dn (nullable): the input sizes on the device in the output layout [nij, nij, npt, nee,
// nvg] (the counts here then bound them).
struct KeepInput {
    const int32_t *ij, *ji, *pt, *ee, *vg;
    int64_t nij, npt, nee, nvg;
    const int64_t* dn = nullptr;
    VertIds vids = {};  // the PT / EE tuples' vertex ids (nullable)
};
struct KeepBoxes {
    const float *blo, *bhi, *plo, *phi, *elo, *ehi, *tlo, *thi;
};
struct KeepOutput {
    int32_t *ij, *ji, *pt, *ee, *vg;
    int32_t *ptv = nullptr, *eev = nullptr;  // the kept tuples' vertex ids (when the input has them)
};
// The vertex ids of n PT (kind 0) or EE (kind 1) tuples into vid [n][4] (VertIds).
int cand_vertex_ids(hipStream_t s, const AbdScene* sc, int kind, const int32_t* tuples, int64_t n, int32_t* vid);
// The primitive boxes of the predicates computed per entry from the vertices x instead of
// read from KeepBoxes (whose point, edge and triangle boxes are then unused): the box
// kernels' own expressions, points (float)x -+ (float)pdil, edges and triangles the FP64
// min / max of their vertices -+ edil / tdil, cast to FP32; with x1 set, the swept boxes'
// (the union of the boxes at x and at x1).
struct KeepAtX {
    const double *x, *x1;
    double pdil, edil, tdil;
};
// sync false: the totals reach hcount with the stream's next synchronization (the caller's);
// hcount null: they stay on the device.
int keep_from_superset(hipStream_t s, const AbdScene* sc, const KeepInput& in, const KeepBoxes& bx, double dhat,
                       double ground, const KeepOutput& out, int64_t* dcount, int64_t* hcount, bool sync = true,
                       const KeepAtX* at = nullptr);
// abd_bsr_structure; dnij (nullable): the pair count on the device (nij bounds it; cols and
// slots hold nb + 2 nij entries).
int bsr_structure(hipStream_t s, int nb, const int32_t* ij, const int32_t* ji, int64_t nij, const int64_t* dnij,
                  int32_t* row_ptr, int32_t* cols, int32_t* slots);
// abd_cull_bodies without the (j, i) order when ji == nullptr.
int cull_bodies(hipStream_t s, const AbdScene* sc, const float* lo, const float* hi, int32_t* ij, int32_t* ji,
                int64_t cap, int64_t* count);
// The gh stage's candidate lists.  dn (nullable): their sizes on the device, [nij, nij,
// npt, nee, nvg]; the counts here are then upper bounds.
struct GhLists {
    const int32_t* pt;
    int64_t npt;
    const int32_t* ee;
    int64_t nee;
    const int32_t* vg;
    int64_t nvg;
    int64_t nij;
    const int64_t* dn;
};
// The inertia gradient (overwriting g) and Hessian (added into the diagonal slots) in one launch.
int inertia_gh(hipStream_t s, const AbdScene* sc, const AbdParams* P, const double* A, const double* p,
               const double* A0, const double* p0, const double* Adot, const double* pdot, double* g, double* blocks);
int contact_gh(hipStream_t s, const AbdScene* sc, const AbdParams* P, const double* x, const GhLists& L,
               const int32_t* row_ptr, const int32_t* slots, double* g, double* blocks, uint8_t* highlight,
               int64_t* n_active, int64_t* n_active_dev);
// A set of arena buffers.  The stateless entry points share one per device (process-wide);
// each simulator owns one, made the calling thread's arena by an ArenaScope for the length
// of a simulator call, so simulators stepped from different threads or streams never share
// scratch.
struct Arena {
    void* ptr[kArenaSlots] = {};
    size_t size[kArenaSlots] = {};
    void* host = nullptr;
    size_t host_size = 0;
};
class ArenaScope {
  public:
    explicit ArenaScope(Arena* a);
    ~ArenaScope();
    ArenaScope(const ArenaScope&) = delete;
    ArenaScope& operator=(const ArenaScope&) = delete;

  private:
    Arena* prev_;
};
int arena_get(int slot, size_t bytes, void** ptr);
int arena_host(size_t bytes, void** ptr);  // pinned host memory
int arena_release_current();
int arena_free(Arena& a);  // synchronizes the device, then releases every buffer of `a`
// A counter that every (re)allocation of library scratch advances: device pointers baked into
// a recorded HIP graph stay valid while it is unchanged.
uint64_t alloc_generation();
void note_reallocation();
// While set on this thread (graph capture), scratch that would have to grow returns
// ABD_ERR_CAPACITY instead of allocating.
void set_capturing(bool on);
bool capturing();
// Zeroes `bytes` (a multiple of 8, 8-byte aligned) with a kernel: a fill that may be recorded in
// a HIP graph (memset nodes are avoided there).
int zero_fill(hipStream_t s, void* p, size_t bytes);

// Exclusive scan of n int64 counts (device) in place into `out` (may alias),
// writing the total to *total (device).  Fixed-order, single workgroup passes.
int scan_exclusive_i64(hipStream_t s, const int64_t* in, int64_t* out, int64_t n, int64_t* total,
                       int64_t* total2 = nullptr);
// Several independent exclusive scans in one launch (one workgroup each, same order as
// scan_exclusive_i64); total2 nullable.
struct ScanSeg {
    const int64_t* in;
    int64_t* out;
    int64_t n;
    int64_t *total, *total2;
};
constexpr int kMaxScanSegs = 8;
int scan_exclusive_i64_multi(hipStream_t s, const ScanSeg* segs, int nseg);
// Deterministic sum of n doubles: *out (device) = sum, fixed order.
// Copies one int64 from device to host (synchronizes `s`).
int read_i64(hipStream_t s, const int64_t* dev, int64_t* host);

constexpr int kBlock = 256;
inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline unsigned grid_for(int64_t n, int block = kBlock, int64_t cap = 65535) {
    int64_t g = ceil_div(n, block);
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return static_cast<unsigned>(g);
}

}  // namespace abd
