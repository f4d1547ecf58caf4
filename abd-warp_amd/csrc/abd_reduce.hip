// abd_reduce.hip — the scalar reductions of the Newton loop (barrier energy,
// inertia energy, CCD time of impact, dot products), per independent scene.
//
// Every reduction is evaluated in an order fixed by the scene's own data: a
// scene's candidate lists are cut into chunks of kChunkItems counted from the
// scene's first candidate, each chunk is summed (or min-reduced) by one
// workgroup in a fixed pattern, and the chunk results are reduced per scene in
// chunk order.  The same scene therefore gives bit-identical results whether it
// is evaluated alone (n_scenes = 1: the plain entry points) or inside a batch of
// any other scenes, and the plain and batched drivers follow the same Newton
// path to the last bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "abd_block.h"
#include "abd_device.h"
#include "abd_inertia.h"
#include "abd_internal.h"
#include "abd_profile.h"

namespace abd {

constexpr int kRedThreads = 256;
constexpr int kChunkItems = 2048;
constexpr int kLists = 3;

enum { kListPT = 0, kListEE = 1, kListVG = 2 };

struct ListRef {
    const int32_t* p;
    int64_t n;
    int kind;                   // kListPT / kListEE / kListVG
    const int32_t* v = nullptr;  // PT / EE: the candidates' scene vertex ids, 4 per tuple (nullable)
};
__device__ __forceinline__ const int32_t* vids_of(const ListRef& r, int64_t i) { return r.v ? r.v + 4 * i : nullptr; }
// dn (nullable): the lists' sizes on the device, [nij, nij, npt, nee, nvg] (the keep
// filter's layout); the n of each ListRef then bounds its size (a list with n = 0 is empty).
struct Lists {
    ListRef l[kLists];
    const int64_t* dn = nullptr;
};

__device__ __forceinline__ int64_t list_n(const Lists& L, int l) {
    const ListRef& r = L.l[l];
    if (!L.dn || r.n == 0) return r.n;
    return L.dn[r.kind == kListPT ? 2 : r.kind == kListEE ? 3 : 4];
}

__device__ __forceinline__ int width_of(int kind) { return kind == kListVG ? 2 : 5; }

// first k in [0, n) whose tuple's first body belongs to a scene >= s
__device__ __forceinline__ int64_t scene_lower_bound(const ListRef& L, int64_t n,
                                                     const int32_t* __restrict__ body_scene, int s) {
    const int w = width_of(L.kind);
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (body_scene[L.p[w * mid]] < s)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

__device__ __forceinline__ int64_t n_chunks(int64_t len) { return (len + kChunkItems - 1) / kChunkItems; }

// Per scene: the segment [a, b) of each list and the offset of its first chunk.  One
// workgroup: first the S + 1 scene boundaries of every list, one binary search per
// thread (scene s ends where scene s + 1 begins), then the chunk counts of the scenes in
// tiles of blockDim with a running offset.
constexpr int kPlanThreads = 1024;
__global__ __launch_bounds__(kPlanThreads) void plan_kernel(int S, const int32_t* __restrict__ body_scene, Lists L,
                                                            int64_t* __restrict__ seg, int64_t* __restrict__ choff) {
    __shared__ int64_t sh[kPlanThreads];
    __shared__ int64_t tile_total;
    for (int idx = threadIdx.x; idx < kLists * (S + 1); idx += blockDim.x) {
        const int l = idx / (S + 1), s = idx % (S + 1);
        const int64_t n = list_n(L, l);
        int64_t v;
        if (S == 1)
            v = s == 0 ? 0 : n;
        else
            v = n > 0 ? scene_lower_bound(L.l[l], n, body_scene, s) : 0;
        if (s < S) seg[2 * (kLists * (int64_t)s + l)] = v;
        if (s > 0) seg[2 * (kLists * (int64_t)(s - 1) + l) + 1] = v;
    }
    __syncthreads();  // the boundaries written above are read by other threads below
    int64_t carry = 0;
    for (int base = 0; base < S; base += blockDim.x) {
        const int s = base + threadIdx.x;
        int64_t cnt = 0;
        if (s < S)
            for (int l = 0; l < kLists; ++l)
                cnt += n_chunks(seg[2 * (kLists * (int64_t)s + l) + 1] - seg[2 * (kLists * (int64_t)s + l)]);
        const int64_t excl = block_exclusive_scan_i64(cnt, sh);
        if (s < S) choff[s] = carry + excl;
        if (threadIdx.x == blockDim.x - 1) tile_total = excl + cnt;
        __syncthreads();
        carry += tile_total;
        __syncthreads();
    }
    if (threadIdx.x == 0) choff[S] = carry;
}

// the scene whose chunks contain chunk c: the largest s with choff[s] <= c
__device__ __forceinline__ int scene_of_chunk(const int64_t* __restrict__ choff, int S, int64_t c) {
    int lo = 0, hi = S - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (choff[mid] <= c)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Vertex positions at the line-search trial state q - alpha dq of one scene, computed on
// the fly with the arithmetic of abd_update_q + abd_project_vertices (abd.py:305-353), so
// that they equal the projected vertices of the updated state bit for bit.
struct VertsTrial {
    const double *xr, *A, *p, *dq;  // rest positions, state, step
    double alpha;
    __device__ __forceinline__ static d3 place(const double (&Ab)[9], const double (&pb)[3],
                                               const double* __restrict__ r) {
        double x[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double v = pb[c] + r[0] * Ab[c];
            v = v + r[1] * Ab[3 + c];
            v = v + r[2] * Ab[6 + c];
            x[c] = v;
        }
        return mk(x[0], x[1], x[2]);
    }
    __device__ __forceinline__ void load4(int I, int J, const int64_t vi[4], int nI, d3 v[4]) const {
        double AI[9], pI[3], AJ[9], pJ[3];
        trial_state(I, alpha, A, p, dq, AI, pI);
        trial_state(J, alpha, A, p, dq, AJ, pJ);
        for (int k = 0; k < 4; ++k)
            v[k] = k < nI ? place(AI, pI, xr + 3 * vi[k]) : place(AJ, pJ, xr + 3 * vi[k]);
    }
    __device__ __forceinline__ double y(int b, int64_t vid) const {
        double Ab[9], pb[3];
        trial_state(b, alpha, A, p, dq, Ab, pb);
        return place(Ab, pb, xr + 3 * vid).y;
    }
};

// VertsTrial with the trial states of the bodies [b0, b0 + n) computed once per workgroup
// into LDS (st: 12 doubles per body, A then p).
struct VertsTrialLds {
    const double* xr;
    const double* st;
    int b0;
    __device__ __forceinline__ void state(int b, double (&Ab)[9], double (&pb)[3]) const {
        const double* q = st + 12 * (b - b0);
#pragma unroll
        for (int i = 0; i < 9; ++i) Ab[i] = q[i];
#pragma unroll
        for (int i = 0; i < 3; ++i) pb[i] = q[9 + i];
    }
    __device__ __forceinline__ void load4(int I, int J, const int64_t vi[4], int nI, d3 v[4]) const {
        double AI[9], pI[3], AJ[9], pJ[3];
        state(I, AI, pI);
        state(J, AJ, pJ);
        for (int k = 0; k < 4; ++k)
            v[k] = k < nI ? VertsTrial::place(AI, pI, xr + 3 * vi[k]) : VertsTrial::place(AJ, pJ, xr + 3 * vi[k]);
    }
    __device__ __forceinline__ double y(int b, int64_t vid) const {
        double Ab[9], pb[3];
        state(b, Ab, pb);
        return VertsTrial::place(Ab, pb, xr + 3 * vid).y;
    }
};
// bodies of a scene whose trial states a chunk workgroup keeps in LDS
constexpr int kTrialLdsBodies = 128;

// ipcsk/energy.py:10-48 per candidate
template <class Src>
__device__ __forceinline__ double energy_term(int kind, const AbdScene& sc, const SceneView& V, const AbdParams& P,
                                              const Src& X, const int32_t* t, const int32_t* vid) {
    if (P.psd == 3 && kind != kListVG)  // full IPC (beyond the reference)
        return cand_energy_full(kind == kListPT ? 0 : 1, V, X, t, P.d2hat, P.kappa);
    if (kind == kListEE) return cand_energy_ee(V, X, t, P.d2hat, P.kappa, vid);
    if (kind == kListPT) return cand_energy_pt(V, X, t, P.d2hat, P.kappa, vid);
    return cand_energy_vg(sc.voff, X, t, P.ground, P.d2hat, P.kappa);
}

// The line-search trials of chunk_kernel<2>: step lengths alpha [trial][S] (a negative
// length skips that scene's trial: its sums are 0) and the state.
struct TrialArgs {
    const double *alpha, *A, *p, *dq;
};

// One chunk per workgroup.  OP 0 sums energies at positions x0; OP 2 sums energies at the
// trial state blockIdx.y of T (partials part + part_stride blockIdx.y).
template <int OP>
__global__ __launch_bounds__(kRedThreads) void chunk_kernel(AbdScene sc, AbdParams P, const double* __restrict__ x0,
                                                            const double* __restrict__ x1, TrialArgs T, Lists L,
                                                            int S, const int64_t* __restrict__ seg,
                                                            const int64_t* __restrict__ choff,
                                                            double* __restrict__ part, int64_t part_stride) {
    __shared__ double sh[kRedThreads];
    part += part_stride * blockIdx.y;
    const int64_t c = blockIdx.x;
    if (c >= choff[S]) return;
    const int s = scene_of_chunk(choff, S, c);
    // skipped trials (uniform over the workgroup): the neutral element of the sum
    if (OP == 2 && T.alpha[(int64_t)S * blockIdx.y + s] < 0.0) {
        if (threadIdx.x == 0) part[c] = 0.0;
        return;
    }
    int64_t k = c - choff[s];
    int l = 0;
    int64_t start = 0, end = 0;
    for (; l < kLists; ++l) {
        const int64_t a = seg[2 * (kLists * (int64_t)s + l)], b = seg[2 * (kLists * (int64_t)s + l) + 1];
        const int64_t nch = n_chunks(b - a);
        if (k < nch) {
            start = a + k * kChunkItems;
            end = start + kChunkItems < b ? start + kChunkItems : b;
            break;
        }
        k -= nch;
    }
    const SceneView V{sc.voff, sc.foff, sc.eoff, sc.tris, sc.edges, sc.x0};
    const int kind = l < kLists ? L.l[l].kind : 0;
    const int w = width_of(kind);
    double acc = 0.0;
    if (l < kLists) {
        if (OP == 0) {
            const Verts X{x0};
            for (int64_t i = start + threadIdx.x; i < end; i += blockDim.x)
                acc += energy_term(kind, sc, V, P, X, L.l[l].p + w * i, vids_of(L.l[l], i));
        } else if (OP == 2) {
            const double al = T.alpha[(int64_t)S * blockIdx.y + s];
            const int b0 = S > 1 ? sc.scene_boff[s] : 0, b1 = S > 1 ? sc.scene_boff[s + 1] : sc.n_bodies;
            if (b1 - b0 <= kTrialLdsBodies) {  // uniform over the workgroup
                __shared__ double st[12 * kTrialLdsBodies];
                for (int b = b0 + threadIdx.x; b < b1; b += blockDim.x) {
                    double Ab[9], pb[3];
                    trial_state(b, al, T.A, T.p, T.dq, Ab, pb);
                    double* q = st + 12 * (b - b0);
                    for (int i = 0; i < 9; ++i) q[i] = Ab[i];
                    for (int i = 0; i < 3; ++i) q[9 + i] = pb[i];
                }
                __syncthreads();
                const VertsTrialLds X{sc.x0, st, b0};
                for (int64_t i = start + threadIdx.x; i < end; i += blockDim.x)
                    acc += energy_term(kind, sc, V, P, X, L.l[l].p + w * i, vids_of(L.l[l], i));
            } else {
                const VertsTrial X{sc.x0, T.A, T.p, T.dq, al};
                for (int64_t i = start + threadIdx.x; i < end; i += blockDim.x)
                    acc += energy_term(kind, sc, V, P, X, L.l[l].p + w * i, vids_of(L.l[l], i));
            }
        }
    }
    const double r = block_sum_fast(acc, sh);
    if (threadIdx.x == 0) part[c] = r;
}

template <int OP>
__global__ __launch_bounds__(kRedThreads) void scene_final_kernel(const int64_t* __restrict__ choff,
                                                                  const double* __restrict__ part,
                                                                  double* __restrict__ out, int64_t part_stride,
                                                                  int64_t out_stride) {
    __shared__ double sh[kRedThreads];
    part += part_stride * blockIdx.y;
    out += out_stride * blockIdx.y;
    const int s = blockIdx.x;
    double acc = 0.0;
    for (int64_t i = choff[s] + threadIdx.x; i < choff[s + 1]; i += blockDim.x) acc += part[i];
    const double r = block_sum_fast(acc, sh);
    if (threadIdx.x == 0) out[s] = r;
}

// ---------------------------------------------------------------- CCD TOI
// The scene minimum of the TOIs below 1 (ccd.py:9-40, abd.py:284-299) in two passes.
// toi_scan_kernel: one candidate per thread (kToiSplit workgroups of kRedThreads per chunk
// of the plan): the vertex-ground TOI directly, and for a PT / EE candidate its
// coplanarity cubic and bracket test; the few candidates with a bracketed root in [0, 1]
// go to a queue (one atomic per wave).  toi_solve_kernel: the full collision_time of each
// queued candidate (root searches and validity), its TOI below 1 min-ed into its scene's
// slot (non-negative doubles order like their bit patterns, so an integer atomic min is
// exact).  toi_final_kernel: per scene the minimum of its chunks' partials and its slot.
// A minimum does not depend on the order, so the result is the one of evaluating every
// candidate in turn.
constexpr int kToiSplit = kChunkItems / kRedThreads;  // one candidate per thread
constexpr int kToiSolveThreads = 64;
constexpr int kToiSolveGrid = 2048;
constexpr unsigned long long kBitsOne = 0x3FF0000000000000ull;  // 1.0

__global__ void toi_init_kernel(int S, unsigned long long* __restrict__ slots, int32_t* __restrict__ qcount) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < S) slots[s] = kBitsOne;
    if (s == 0) *qcount = 0;
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const double u = __shfl_xor(v, o);
        v = u < v ? u : v;
    }
    return v;
}

__global__ __launch_bounds__(kRedThreads, 4) void toi_scan_kernel(AbdScene sc, AbdParams P, const double* __restrict__ x0,
                                                               const double* __restrict__ x1, Lists L, int S,
                                                               const int64_t* __restrict__ seg,
                                                               const int64_t* __restrict__ choff,
                                                               const uint8_t* __restrict__ active,
                                                               double* __restrict__ part, int32_t* __restrict__ queue,
                                                               int32_t* __restrict__ qcount) {
    __shared__ double sh[kRedThreads / kWave];
    const int64_t c = blockIdx.x;
    if (c >= choff[S]) return;  // uniform over the workgroup
    const int s = scene_of_chunk(choff, S, c);
    double acc = 1.0;
    bool queued = false;
    int32_t entry = 0;
    if (!active || active[s]) {
        int64_t k = c - choff[s];
        int l = 0;
        int64_t start = 0, end = 0;
        for (; l < kLists; ++l) {
            const int64_t a = seg[2 * (kLists * (int64_t)s + l)], b = seg[2 * (kLists * (int64_t)s + l) + 1];
            const int64_t nch = n_chunks(b - a);
            if (k < nch) {
                start = a + k * kChunkItems;
                end = start + kChunkItems < b ? start + kChunkItems : b;
                break;
            }
            k -= nch;
        }
        const int64_t i = start + (int64_t)blockIdx.y * kRedThreads + threadIdx.x;
        if (l < kLists && i < end) {
            const int kind = L.l[l].kind;
            const int32_t* t = L.l[l].p + (int64_t)width_of(kind) * i;
            if (kind == kListVG) {
                const double tv = cand_toi_vg(sc.voff, x0, x1, t, P.ground);
                if (tv < 1.0) acc = tv;  // `if t < 1: atomic_min`
            } else {
                const SceneView V{sc.voff, sc.foff, sc.eoff, sc.tris, sc.edges, sc.x0};
                queued = cand_toi_bracketed(kind == kListPT ? 0 : 1, V, x0, x1, t, vids_of(L.l[l], i));
                entry = (int32_t)((i << 1) | (kind == kListEE ? 1 : 0));
            }
        }
    }
    // wave-aggregated push of the queued candidates
    const unsigned long long mask = __ballot(queued);
    if (mask) {
        const int lane = lane_id();
        const int leader = __ffsll((long long)mask) - 1;
        int32_t base = 0;
        if (lane == leader) base = atomicAdd(qcount, (int32_t)__popcll(mask));
        base = __shfl(base, leader);
        if (queued) queue[base + __popcll(mask & ((1ull << lane) - 1ull))] = entry;
    }
    const double wm = wave_min(acc);
    if (lane_id() == 0) sh[wave_id()] = wm;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = sh[0];
        for (int w = 1; w < kRedThreads / kWave; ++w) r = sh[w] < r ? sh[w] : r;
        part[c * kToiSplit + blockIdx.y] = r;
    }
}

__global__ __launch_bounds__(kToiSolveThreads) void toi_solve_kernel(AbdScene sc, Lists L, int S,
                                                                     const double* __restrict__ x0,
                                                                     const double* __restrict__ x1,
                                                                     const int32_t* __restrict__ queue,
                                                                     const int32_t* __restrict__ qcount,
                                                                     unsigned long long* __restrict__ slots) {
    const int32_t n = *qcount;
    const SceneView V{sc.voff, sc.foff, sc.eoff, sc.tris, sc.edges, sc.x0};
    for (int32_t q = blockIdx.x * kToiSolveThreads + threadIdx.x; q < n; q += gridDim.x * kToiSolveThreads) {
        const int32_t e = queue[q];
        const int ee = e & 1;
        const ListRef& Lr = L.l[ee ? 0 : 1];  // the lists' order: EE, PT, VG
        const int32_t* t = Lr.p + 5 * (int64_t)(e >> 1);
        const double tt = cand_toi(ee, V, x0, x1, t, vids_of(Lr, e >> 1));
        if (tt < 1.0) {
            const int s = S > 1 ? sc.body_scene[t[0]] : 0;
            atomicMin(slots + s, (unsigned long long)__double_as_longlong(tt));
        }
    }
}

// alpha (nullable): the line search's first step length of each scene from its TOI
// (first_step_length; -1 for a scene not in `active`)
__global__ __launch_bounds__(kRedThreads) void toi_final_kernel(const int64_t* __restrict__ choff,
                                                                const double* __restrict__ part,
                                                                const unsigned long long* __restrict__ slots,
                                                                double* __restrict__ out,
                                                                const uint8_t* __restrict__ active,
                                                                double* __restrict__ alpha, SceneDots D) {
    __shared__ double sh[kRedThreads];
    const int s = blockIdx.x;
    double acc = __longlong_as_double((long long)slots[s]);
    for (int64_t i = choff[s] * kToiSplit + threadIdx.x; i < choff[s + 1] * kToiSplit; i += blockDim.x) {
        const double v = part[i];
        acc = v < acc ? v : acc;
    }
    const double r = block_min(acc, sh);
    if (threadIdx.x == 0) {
        out[s] = r;
        if (alpha) alpha[s] = !active || active[s] ? first_step_length(r) : -1.0;
    }
    if (D.out1) {  // uniform: seg_dot2_kernel's sums of this scene
        const int64_t i0 = D.sboff ? 12 * (int64_t)D.sboff[s] : 0, i1 = D.sboff ? 12 * (int64_t)D.sboff[s + 1] : D.n;
        double acc1 = 0.0, acc2 = 0.0;
        for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
            acc1 += D.a[i] * D.b[i];
            acc2 += D.c[i] * D.d[i];
        }
        block_sum2_fast(acc1, acc2, sh);
        if (threadIdx.x == 0) {
            D.out1[s] = acc1;
            D.out2[s] = acc2;
        }
    }
}

// out[s] = sum over the rows [12 sboff[s], 12 sboff[s + 1]) of a * b (one workgroup per scene;
// sboff == nullptr: one segment [0, n))
__global__ __launch_bounds__(kRedThreads) void seg_dot_kernel(const int32_t* __restrict__ sboff, int64_t n,
                                                              const double* __restrict__ a,
                                                              const double* __restrict__ b, double* __restrict__ out) {
    __shared__ double sh[17];
    const int s = blockIdx.x;
    const int64_t i0 = sboff ? 12 * (int64_t)sboff[s] : 0, i1 = sboff ? 12 * (int64_t)sboff[s + 1] : n;
    double acc = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) acc += a[i] * b[i];
    const double r = block_sum_fast(acc, sh);
    if (threadIdx.x == 0) out[s] = r;
}

// out1[s] = a . b and out2[s] = c . d over the rows of scene s, each bit-identical to its
// seg_dot_kernel
__global__ __launch_bounds__(kRedThreads) void seg_dot2_kernel(const int32_t* __restrict__ sboff, int64_t n,
                                                               const double* __restrict__ a,
                                                               const double* __restrict__ b,
                                                               const double* __restrict__ c,
                                                               const double* __restrict__ d,
                                                               double* __restrict__ out1, double* __restrict__ out2) {
    __shared__ double sh[34];
    const int s = blockIdx.x;
    const int64_t i0 = sboff ? 12 * (int64_t)sboff[s] : 0, i1 = sboff ? 12 * (int64_t)sboff[s + 1] : n;
    double acc1 = 0.0, acc2 = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        acc1 += a[i] * b[i];
        acc2 += c[i] * d[i];
    }
    block_sum2_fast(acc1, acc2, sh);
    if (threadIdx.x == 0) {
        out1[s] = acc1;
        out2[s] = acc2;
    }
}

// orthogonal_energy.py:25-51 summed over the bodies of each scene (one workgroup per scene)
__global__ __launch_bounds__(kRedThreads) void seg_inertia_kernel(const int32_t* __restrict__ sboff, int nb,
                                                                  AbdParams P, const double* __restrict__ mass,
                                                                  const double* __restrict__ I0v,
                                                                  const double* __restrict__ Ak,
                                                                  const double* __restrict__ pk,
                                                                  const double* __restrict__ A0,
                                                                  const double* __restrict__ p0,
                                                                  const double* __restrict__ Adot,
                                                                  const double* __restrict__ pdot,
                                                                  double* __restrict__ E) {
    __shared__ double sh[17];
    const int s = blockIdx.x;
    const int b0 = sboff ? sboff[s] : 0, b1 = sboff ? sboff[s + 1] : nb;
    double acc = 0.0;
    for (int b = b0 + threadIdx.x; b < b1; b += blockDim.x)
        acc += inertia_energy_body(P, mass, I0v, b, Ak, pk, A0, p0, Adot, pdot);
    const double r = block_sum_fast(acc, sh);
    if (threadIdx.x == 0) E[s] = r;
}

// Upper bound on the chunks of the lists over S scenes (each scene's lists may end in a
// partial chunk).
int64_t chunk_bound(const Lists& L, int S) {
    int64_t bound = 3 * (int64_t)S;
    for (int l = 0; l < kLists; ++l) bound += (L.l[l].n + kChunkItems - 1) / kChunkItems;
    return bound;
}

// Caller-provided scratch of the TOI (the allocation-free entry point's): the plan, the
// scene slots and queue count, room for part_cap chunk partials (kToiSplit per chunk) and
// queue_cap queued candidates.  Null: the arena's.
struct ReduceScratch {
    int64_t* plan = nullptr;
    unsigned long long* slots = nullptr;
    int32_t* qcount = nullptr;
    double* part = nullptr;
    int64_t part_cap = 0;
    int32_t* queue = nullptr;
    int64_t queue_cap = 0;
};

// The per-scene segments and chunk offsets of the lists: plan_in when given (made by
// reduce_plan for these lists), else computed into plan_buf (nullable: the arena).
int get_plan(hipStream_t s, const AbdScene& sc, int S, const Lists& L, const int64_t* plan_in,
             const int64_t** seg, const int64_t** choff, int64_t* plan_buf = nullptr) {
    if (plan_in) {
        *seg = plan_in;
        *choff = plan_in + 2 * kLists * (size_t)S;
        return ABD_OK;
    }
    int64_t* plan = plan_buf;
    int rc = plan ? ABD_OK : arena_get(kSlotPlan, sizeof(int64_t) * reduce_plan_elems(S), (void**)&plan);
    if (rc) return rc;
    hipLaunchKernelGGL(plan_kernel, dim3(1), dim3(kPlanThreads), 0, s, S, sc.body_scene, L, plan,
                       plan + 2 * kLists * (size_t)S);
    ABD_CHECK_LAUNCH();
    *seg = plan;
    *choff = plan + 2 * kLists * (size_t)S;
    return ABD_OK;
}

// Plan + chunks + per-scene sums (OP 0) over the lists; S = 1 treats the lists as one scene.
// Algorithmic bytes: 128 per PT, 132 per EE and 32 per VG candidate.
template <int OP>
int seg_reduce(hipStream_t s, const AbdScene& sc, const AbdParams& P, int S, const double* x0, const double* x1,
               const Lists& L, const int64_t* plan_in, double* out, int prof_id) {
    const int64_t bound = chunk_bound(L, S);
    if (bound > INT32_MAX) return ABD_ERR_ARG;
    const int64_t *seg, *choff;
    double* part = nullptr;
    int rc = get_plan(s, sc, S, L, plan_in, &seg, &choff);
    if (!rc) rc = arena_get(kSlotReduce, sizeof(double) * (size_t)bound, (void**)&part);
    if (rc) return rc;
    int ph = prof_start_ext(prof_id, L.dn ? 1.0 : 128.0 * L.l[1].n + 132.0 * L.l[0].n + 32.0 * L.l[2].n);
    prof_launch(ph, chunk_kernel<OP>, dim3((unsigned)bound), dim3(kRedThreads), 0, s, sc, P, x0, x1,
                TrialArgs{nullptr, nullptr, nullptr, nullptr}, L, S, seg, choff, part, (int64_t)0);
    ABD_CHECK_LAUNCH();
    if ((rc = prof_count_bytes(ph, s, L.dn, 128.0, 132.0, 32.0, 0.0))) return rc;
    hipLaunchKernelGGL(scene_final_kernel<OP>, dim3(S), dim3(kRedThreads), 0, s, choff, part, out, (int64_t)0,
                       (int64_t)0);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

// Algorithmic bytes of a launch over the plan's scenes: c_ee, c_pt, c_vg per EE, PT and VG
// candidate of every (trial k, scene s) it evaluates -- K trials with step lengths alpha
// [K][S] (nullable; a negative step length skips the scene's trial) of the scenes with
// active[s] set (nullable: all) -- from the plan's per-scene segments.
__global__ __launch_bounds__(kRedThreads) void seg_bytes_kernel(int S, int K, const double* __restrict__ alpha,
                                                                const uint8_t* __restrict__ active,
                                                                const int64_t* __restrict__ seg, double c_ee,
                                                                double c_pt, double c_vg, double* __restrict__ out) {
    __shared__ double sh[kRedThreads];
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < (int64_t)S * K; i += blockDim.x) {
        const int64_t sc = i % S;
        if ((alpha && !(alpha[i] >= 0.0)) || (active && !active[sc])) continue;
        const int64_t* q = seg + 2 * (int64_t)kLists * sc;
        acc += c_ee * (double)(q[1] - q[0]) + c_pt * (double)(q[3] - q[2]) + c_vg * (double)(q[5] - q[4]);
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) *out = acc;
}

// Attaches to `handle` the bytes of seg_bytes_kernel (nothing when the handle is -1).
int attach_seg_bytes(int handle, hipStream_t s, int S, int K, const double* alpha, const uint8_t* active,
                     const int64_t* seg, double c_ee, double c_pt, double c_vg) {
    double* out;
    int rc = prof_byte_slot(handle, &out);
    if (rc || !out) return rc;
    hipLaunchKernelGGL(seg_bytes_kernel, dim3(1), dim3(kRedThreads), 0, s, S, K, alpha, active, seg, c_ee, c_pt,
                       c_vg, out);
    ABD_CHECK_LAUNCH();
    prof_stop(handle, s, nullptr, out);
    return ABD_OK;
}

// Per-scene TOI (toi_scan_kernel, toi_solve_kernel, toi_final_kernel) with scratch from
// scr (nullable: the arena; then the scan and the solve are timed when profiling).
int toi_reduce(hipStream_t s, const AbdScene& sc, const AbdParams& P, int S, const double* x0, const double* x1,
               const Lists& L, const int64_t* plan_in, double* out, const uint8_t* active,
               const ReduceScratch* scr, double* first_alpha = nullptr, const SceneDots* dots = nullptr) {
    const int64_t bound = chunk_bound(L, S);
    if (bound * kToiSplit > INT32_MAX) return ABD_ERR_ARG;
    const int64_t nq = L.l[0].n + L.l[1].n;  // every PT and EE candidate may queue
    if (nq > (int64_t)INT32_MAX / 2) return ABD_ERR_ARG;
    if (scr && (bound * kToiSplit > scr->part_cap || nq > scr->queue_cap)) return ABD_ERR_CAPACITY;
    ReduceScratch w;
    if (scr) {
        w = *scr;
    } else {
        int rc;
        if ((rc = arena_get(kSlotReduce, sizeof(double) * (size_t)(bound * kToiSplit), (void**)&w.part))) return rc;
        if ((rc = arena_get(kSlotToiQueue, sizeof(int32_t) * (size_t)(nq + 1), (void**)&w.queue))) return rc;
        if ((rc = arena_get(kSlotToiSlots, sizeof(unsigned long long) * (size_t)(S + 1), (void**)&w.slots)))
            return rc;
        w.qcount = reinterpret_cast<int32_t*>(w.slots + S);
    }
    const int64_t *seg, *choff;
    int rc = get_plan(s, sc, S, L, plan_in, &seg, &choff, w.plan);
    if (rc) return rc;
    hipLaunchKernelGGL(toi_init_kernel, dim3((unsigned)ceil_div(S, kRedThreads)), dim3(kRedThreads), 0, s, S,
                       w.slots, w.qcount);
    ABD_CHECK_LAUNCH();
    // algorithmic bytes of the scan: tuple, indices and four vertices at both ends per PT / EE
    // candidate (224 / 228 B), tuple, offset and a vertex at both ends per VG candidate (56 B),
    // over the scenes it evaluates (counted on the device)
    const int ph = scr ? -1 : prof_start_ext(kProfToi, 1.0);
    prof_launch(ph, toi_scan_kernel, dim3((unsigned)bound, kToiSplit), dim3(kRedThreads), 0, s, sc, P, x0, x1, L, S,
                seg, choff, active, w.part, w.queue, w.qcount);
    ABD_CHECK_LAUNCH();
    if ((rc = attach_seg_bytes(ph, s, S, 1, nullptr, active, seg, 228.0, 224.0, 56.0))) return rc;
    const int64_t gmax = tuning().toi_grid > 0 ? tuning().toi_grid : kToiSolveGrid;
    const unsigned grid = (unsigned)std::min<int64_t>(gmax, std::max<int64_t>(1, ceil_div(nq, 4 * kToiSolveThreads)));
    const int ph2 = scr ? -1 : prof_start_ext(kProfToiSolve, 0.0);
    prof_launch(ph2, toi_solve_kernel, dim3(grid), dim3(kToiSolveThreads), 0, s, sc, L, S, x0, x1,
                (const int32_t*)w.queue, (const int32_t*)w.qcount, w.slots);
    ABD_CHECK_LAUNCH();
    hipLaunchKernelGGL(toi_final_kernel, dim3(S), dim3(kRedThreads), 0, s, choff, (const double*)w.part,
                       (const unsigned long long*)w.slots, out, active, first_alpha, dots ? *dots : SceneDots{});
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

// The candidate lists of energies and TOIs, in the one order every reduction uses (the
// sum order of the energy; a minimum does not depend on it).  The switches of
// const_params.py:42-44 drop a whole list.
Lists contact_lists(const AbdParams& P, const int32_t* pt, int64_t npt, const int32_t* ee, int64_t nee,
                    const int32_t* vg, int64_t nvg, const int64_t* dn = nullptr, VertIds vids = {}) {
    if (!P.pt) npt = 0;
    if (!P.ee) nee = 0;
    if (!P.vg) nvg = 0;
    return Lists{{{ee, nee, kListEE, vids.ee}, {pt, npt, kListPT, vids.pt}, {vg, nvg, kListVG, nullptr}}, dn};
}

size_t reduce_plan_elems(int S) { return 2 * kLists * (size_t)S + (size_t)S + 1; }

int reduce_plan(hipStream_t s, const AbdScene* sc, const AbdParams* P, int S, const int32_t* pt, int64_t npt,
                const int32_t* ee, int64_t nee, const int32_t* vg, int64_t nvg, int64_t* plan, const int64_t* dn) {
    const Lists L = contact_lists(*P, pt, npt, ee, nee, vg, nvg, dn);
    hipLaunchKernelGGL(plan_kernel, dim3(1), dim3(kPlanThreads), 0, s, S, sc->body_scene, L, plan,
                       plan + 2 * kLists * (size_t)S);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

// ---------------------------------------------------------- line-search trials
// K trial states q - alpha dq evaluated together (alpha: [K][S] on the device; S = 1
// treats the scene as one): per scene the contact energy of every trial over the
// candidate lists, each candidate's vertices placed at the trial state on the fly
// (VertsTrial), and the inertia energy of every trial, with one plan of the lists for all.
// E[2S k + s] is the contact energy and E[2S k + S + s] the inertia energy of scene s
// at trial k, each bit-identical to the single-state reductions at that state.

// scene_final_kernel<0> and the inertia energy of one (scene, trial) in one workgroup: the
// contact energy from the trial's chunk partials (E[2 S k + s]) and the inertia energy at the
// trial state (E[2 S k + S + s]), each with the arithmetic of its own kernel.
__global__ __launch_bounds__(kRedThreads) void trial_energies_kernel(const int64_t* __restrict__ choff,
                                                                     const double* __restrict__ part,
                                                                     int64_t part_stride, int S,
                                                                     const double* __restrict__ alpha,
                                                                     const int32_t* __restrict__ sboff, int nb,
                                                                     AbdParams P, const double* __restrict__ mass,
                                                                     const double* __restrict__ I0v,
                                                                     const double* __restrict__ A,
                                                                     const double* __restrict__ p,
                                                                     const double* __restrict__ dq,
                                                                     const double* __restrict__ A0,
                                                                     const double* __restrict__ p0,
                                                                     const double* __restrict__ Adot,
                                                                     const double* __restrict__ pdot,
                                                                     double* __restrict__ E) {
    __shared__ double sh[kRedThreads];
    const int s = blockIdx.x, k = blockIdx.y;
    part += part_stride * k;
    double acc = 0.0;
    for (int64_t i = choff[s] + threadIdx.x; i < choff[s + 1]; i += blockDim.x) acc += part[i];
    const double rc = block_sum_fast(acc, sh);
    if (threadIdx.x == 0) E[(int64_t)2 * S * k + s] = rc;
    const double al = alpha[(int64_t)S * k + s];
    if (al < 0.0) {  // skipped trial (uniform over the workgroup)
        if (threadIdx.x == 0) E[(int64_t)2 * S * k + S + s] = 0.0;
        return;
    }
    const int b0 = S > 1 ? sboff[s] : 0, b1 = S > 1 ? sboff[s + 1] : nb;
    acc = 0.0;
    for (int b = b0 + threadIdx.x; b < b1; b += blockDim.x) {
        double Ab[9], pb[3];
        trial_state(b, al, A, p, dq, Ab, pb);
        acc += inertia_energy_q(P, mass, I0v, b, Ab, pb, A0, p0, Adot, pdot);
    }
    const double ri = block_sum_fast(acc, sh);
    if (threadIdx.x == 0) E[(int64_t)2 * S * k + S + s] = ri;
}

int line_trials(hipStream_t s, const AbdScene* sc, const AbdParams* P, int S, int K, const double* alpha,
                const double* A, const double* p, const double* dq, const double* A0, const double* p0,
                const double* Adot, const double* pdot, const int32_t* pt, int64_t npt, const int32_t* ee, int64_t nee,
                const int32_t* vg, int64_t nvg, const int64_t* plan_in, double* E, const int64_t* dn, VertIds vids) {
    int rc;
    const Lists L = contact_lists(*P, pt, npt, ee, nee, vg, nvg, dn, vids);
    npt = L.l[1].n;
    nee = L.l[0].n;
    nvg = L.l[2].n;
    const int64_t bound = chunk_bound(L, S);
    if (bound > INT32_MAX) return ABD_ERR_ARG;
    const int64_t *seg, *choff;
    double* part;
    if ((rc = get_plan(s, *sc, S, L, plan_in, &seg, &choff))) return rc;
    if ((rc = arena_get(kSlotReduce, sizeof(double) * (size_t)bound * K, (void**)&part))) return rc;
    // bytes: those of the (trial, scene) pairs evaluated, counted on the device
    int ph = prof_start_ext(kProfEnergy, 1.0);
    prof_launch(ph, chunk_kernel<2>, dim3((unsigned)bound, K), dim3(kRedThreads), 0, s, *sc, *P,
                (const double*)nullptr, (const double*)nullptr, TrialArgs{alpha, A, p, dq}, L, S, seg, choff,
                part, bound);
    ABD_CHECK_LAUNCH();
    if ((rc = attach_seg_bytes(ph, s, S, K, alpha, nullptr, seg, 132.0, 128.0, 32.0))) return rc;
    hipLaunchKernelGGL(trial_energies_kernel, dim3(S, K), dim3(kRedThreads), 0, s, choff, (const double*)part, bound,
                       S, alpha, sc->scene_boff, sc->n_bodies, *P, sc->mass, sc->I0, A, p, dq, A0, p0, Adot, pdot, E);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

namespace {

bool scenes_ok(const AbdScene* sc) { return sc && sc->n_scenes > 0 && sc->scene_boff && sc->body_scene; }

int energy_impl(abd_stream_t stream, const AbdScene* sc, const AbdParams* P, int S, const double* xk,
                const int32_t* pt, int64_t npt, const int32_t* ee, int64_t nee, const int32_t* vg, int64_t nvg,
                double* E) {
    if (!P || !xk || !E || npt < 0 || nee < 0 || nvg < 0 || P->psd < 0 || P->psd > 3) return ABD_ERR_ARG;
    if ((npt && !pt) || (nee && !ee) || (nvg && !vg)) return ABD_ERR_ARG;
    return contact_energy(to_stream(stream), sc, P, S, xk, pt, npt, ee, nee, vg, nvg, nullptr, E);
}

}  // namespace

int contact_energy(hipStream_t s, const AbdScene* sc, const AbdParams* P, int S, const double* xk, const int32_t* pt,
                   int64_t npt, const int32_t* ee, int64_t nee, const int32_t* vg, int64_t nvg, const int64_t* dn,
                   double* E) {
    const Lists L = contact_lists(*P, pt, npt, ee, nee, vg, nvg, dn);
    return seg_reduce<0>(s, *sc, *P, S, xk, nullptr, L, nullptr, E, kProfEnergy);
}

int scene_toi(hipStream_t s, const AbdScene* sc, const AbdParams* P, int S, const double* x_t0, const double* x_t1,
              const int32_t* pt, int64_t npt, const int32_t* ee, int64_t nee, const int32_t* vg, int64_t nvg,
              const int64_t* plan, double* toi, const uint8_t* active, const int64_t* dn, double* first_alpha,
              VertIds vids, const SceneDots* dots) {
    if (!P || !x_t0 || !x_t1 || !toi || npt < 0 || nee < 0 || nvg < 0) return ABD_ERR_ARG;
    if ((npt && !pt) || (nee && !ee) || (nvg && !vg)) return ABD_ERR_ARG;
    const Lists L = contact_lists(*P, pt, npt, ee, nee, vg, nvg, dn, vids);  // abd.py:288-299
    return toi_reduce(s, *sc, *P, S, x_t0, x_t1, L, plan, toi, active, nullptr, first_alpha, dots);
}

int dots(hipStream_t s, const AbdScene* sc, int S, const double* a, const double* b, const double* c,
         const double* d, double* out1, double* out2) {
    hipLaunchKernelGGL(seg_dot2_kernel, dim3(S), dim3(kRedThreads), 0, s, S > 1 ? sc->scene_boff : nullptr,
                       12 * (int64_t)sc->n_bodies, a, b, c, d, out1, out2);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

// ------------------------------------------------- allocation-free TOI scratch
// The fixed scratch of abd_toi_noalloc: the chunk plan (kNoAllocPlanElems int64), the scene
// slot and queue count, kNoAllocPartElems chunk partials and a queue of
// ABD_TOI_NOALLOC_MAX_CANDIDATES int32 (in doubles: 2 entries each).  In host code `scratch`
// names a host-side stand-in, not device memory: its device address is looked up once per
// device (noalloc_scratch) and only that address reaches kernels.
constexpr int64_t kNoAllocPlanElems = 64;
constexpr int64_t kNoAllocSlotElems = 64;
// chunks of every list at most: 3 + (all candidates in one list) / 2048, kToiSplit partials each
constexpr int64_t kNoAllocPartElems = kToiSplit * (8 + ABD_TOI_NOALLOC_MAX_CANDIDATES / kChunkItems);
constexpr int64_t kNoAllocQueueElems = ABD_TOI_NOALLOC_MAX_CANDIDATES / 2;
static __device__ __align__(256) double scratch[kNoAllocPlanElems + kNoAllocSlotElems + kNoAllocPartElems +
                                                kNoAllocQueueElems];

namespace {
constexpr int kNoAllocDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kNoAllocDevices] = {};

// The device address of `scratch` on the current device (looked up on its first use there).
int noalloc_scratch(ReduceScratch* out) {
    int dev = -1;
    ABD_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocDevices) return ABD_ERR_NODEVICE;
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        ABD_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_noalloc_ptr[dev] = p;
    }
    double* base = static_cast<double*>(g_noalloc_ptr[dev]);
    out->plan = reinterpret_cast<int64_t*>(base);
    out->slots = reinterpret_cast<unsigned long long*>(base + kNoAllocPlanElems);
    out->qcount = reinterpret_cast<int32_t*>(base + kNoAllocPlanElems + 1);
    out->part = base + kNoAllocPlanElems + kNoAllocSlotElems;
    out->part_cap = kNoAllocPartElems;
    out->queue = reinterpret_cast<int32_t*>(base + kNoAllocPlanElems + kNoAllocSlotElems + kNoAllocPartElems);
    out->queue_cap = ABD_TOI_NOALLOC_MAX_CANDIDATES;
    return ABD_OK;
}
static_assert(2 * kLists + 2 <= kNoAllocPlanElems, "one scene's plan fits the front of the scratch");
}  // namespace

}  // namespace abd

using namespace abd;

extern "C" {

int abd_contact_energy(abd_stream_t stream, const AbdScene* sc, const AbdParams* P, const double* xk,
                       const int32_t* pt, int64_t npt, const int32_t* ee, int64_t nee, const int32_t* vg, int64_t nvg,
                       double* E) {
    if (!sc) return ABD_ERR_ARG;
    return energy_impl(stream, sc, P, 1, xk, pt, npt, ee, nee, vg, nvg, E);
}

int abd_scene_contact_energy(abd_stream_t stream, const AbdScene* sc, const AbdParams* P, const double* xk,
                             const int32_t* pt, int64_t npt, const int32_t* ee, int64_t nee, const int32_t* vg,
                             int64_t nvg, double* E) {
    if (!scenes_ok(sc)) return ABD_ERR_ARG;
    return energy_impl(stream, sc, P, sc->n_scenes, xk, pt, npt, ee, nee, vg, nvg, E);
}

int abd_toi(abd_stream_t stream, const AbdScene* sc, const AbdParams* P, const double* x_t0, const double* x_t1,
            const int32_t* pt, int64_t npt, const int32_t* ee, int64_t nee, const int32_t* vg, int64_t nvg,
            double* toi) {
    if (!sc) return ABD_ERR_ARG;
    return scene_toi(to_stream(stream), sc, P, 1, x_t0, x_t1, pt, npt, ee, nee, vg, nvg, nullptr, toi, nullptr,
                     nullptr);
}

int abd_toi_noalloc(abd_stream_t stream, const AbdScene* sc, const AbdParams* P, const double* x_t0,
                    const double* x_t1, const int32_t* pt, int64_t npt, const int32_t* ee, int64_t nee,
                    const int32_t* vg, int64_t nvg, double* toi) {
    if (!sc || !P || !x_t0 || !x_t1 || !toi || npt < 0 || nee < 0 || nvg < 0) return ABD_ERR_ARG;
    if ((npt && !pt) || (nee && !ee) || (nvg && !vg)) return ABD_ERR_ARG;
    const Lists L = contact_lists(*P, pt, npt, ee, nee, vg, nvg);
    // before any device call
    if (L.l[0].n + L.l[1].n + L.l[2].n > ABD_TOI_NOALLOC_MAX_CANDIDATES) return ABD_ERR_CAPACITY;
    ReduceScratch scr;
    int rc = noalloc_scratch(&scr);
    if (rc) return rc;
    return toi_reduce(to_stream(stream), *sc, *P, 1, x_t0, x_t1, L, nullptr, toi, nullptr, &scr);
}

int abd_scene_toi(abd_stream_t stream, const AbdScene* sc, const AbdParams* P, const double* x_t0,
                  const double* x_t1, const int32_t* pt, int64_t npt, const int32_t* ee, int64_t nee,
                  const int32_t* vg, int64_t nvg, double* toi) {
    if (!scenes_ok(sc)) return ABD_ERR_ARG;
    return scene_toi(to_stream(stream), sc, P, sc->n_scenes, x_t0, x_t1, pt, npt, ee, nee, vg, nvg, nullptr,
                     toi, nullptr, nullptr);
}

int abd_dot(abd_stream_t stream, const double* a, const double* b, int64_t n, double* out) {
    if (n < 0 || (!a && n) || (!b && n) || !out) return ABD_ERR_ARG;
    hipLaunchKernelGGL(seg_dot_kernel, dim3(1), dim3(kRedThreads), 0, to_stream(stream), nullptr, n, a, b, out);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

int abd_scene_dot(abd_stream_t stream, const AbdScene* sc, const double* a, const double* b, double* out) {
    if (!scenes_ok(sc) || !a || !b || !out) return ABD_ERR_ARG;
    hipLaunchKernelGGL(seg_dot_kernel, dim3(sc->n_scenes), dim3(kRedThreads), 0, to_stream(stream), sc->scene_boff,
                       (int64_t)0, a, b, out);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

int abd_inertia_energy(abd_stream_t stream, const AbdScene* sc, const AbdParams* P, const double* Ak, const double* pk,
                       const double* A0, const double* p0, const double* Adot, const double* pdot, double* E) {
    if (!sc || !P || !Ak || !pk || !A0 || !p0 || !Adot || !pdot || !E) return ABD_ERR_ARG;
    hipLaunchKernelGGL(seg_inertia_kernel, dim3(1), dim3(kRedThreads), 0, to_stream(stream), nullptr, sc->n_bodies,
                       *P, sc->mass, sc->I0, Ak, pk, A0, p0, Adot, pdot, E);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

int abd_scene_inertia_energy(abd_stream_t stream, const AbdScene* sc, const AbdParams* P, const double* Ak,
                             const double* pk, const double* A0, const double* p0, const double* Adot,
                             const double* pdot, double* E) {
    if (!scenes_ok(sc) || !P || !Ak || !pk || !A0 || !p0 || !Adot || !pdot || !E) return ABD_ERR_ARG;
    hipLaunchKernelGGL(seg_inertia_kernel, dim3(sc->n_scenes), dim3(kRedThreads), 0, to_stream(stream),
                       sc->scene_boff, sc->n_bodies, *P, sc->mass, sc->I0, Ak, pk, A0, p0, Adot, pdot, E);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

}  // extern "C"
