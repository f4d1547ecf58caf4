// abd_cull.hip — candidate culling (culling.py / culling2.py) and the sparse
// structure of the Newton matrix (orthogonal_energy.py:138-165).
//
// Culling replaces the reference's per-body Warp BVHs with an exact two-level
// scheme that needs no tree build or refit:
//   1. body pairs: every (i < j) pair tested against the FP32 body boxes;
//   2. per body pair and side, the primitives whose box overlaps the hull of the
//      OTHER side's primitive boxes are kept (a primitive box that misses that
//      hull cannot overlap any box of the other side, so nothing is lost);
//   3. the filtered lists are tested exhaustively in chunks of `chunk` tests (a tunable
//      launch shape: the emitted list is the same for every chunk size).
// Every list is produced by a count -> exclusive scan -> write sequence with
// wave-ballot ranks, so outputs are emitted in sorted order (ii, order, a, b)
// and are identical from run to run.
#include <hip/hip_runtime.h>

#include <math.h>

#include "abd_block.h"
#include "abd_device.h"
#include "abd_internal.h"
#include "abd_profile.h"

namespace abd {

__device__ __forceinline__ bool overlap(const float* __restrict__ alo, const float* __restrict__ ahi,
                                        const float* __restrict__ blo, const float* __restrict__ bhi) {
    return !(alo[0] > bhi[0] || ahi[0] < blo[0] || alo[1] > bhi[1] || ahi[1] < blo[1] || alo[2] > bhi[2] ||
             ahi[2] < blo[2]);
}

// ---------------------------------------------------------------- bodies
constexpr int kItems = 8;      // entries per thread in the vertex-ground compaction
constexpr int kPairItems = 1;  // pairs per thread in the body-pair compaction: enough workgroups to fill the device

// k -> (i, j), i < j.  Order 0: row-major (i, j); order 1: sorted by (j, i).
__device__ __forceinline__ void pair_of(int64_t k, int nb, int order, int& i, int& j) {
    if (order == 0) {
        double n2 = 2.0 * nb - 1.0;
        int64_t r = (int64_t)floor((n2 - sqrt(n2 * n2 - 8.0 * (double)k)) * 0.5);
        if (r < 0) r = 0;
        auto off = [nb](int64_t q) { return q * nb - q * (q + 1) / 2; };
        while (r > 0 && off(r) > k) --r;
        while (off(r + 1) <= k) ++r;
        i = (int)r;
        j = (int)(k - off(r) + r + 1);
    } else {
        int64_t c = (int64_t)floor((1.0 + sqrt(1.0 + 8.0 * (double)k)) * 0.5);
        if (c < 1) c = 1;
        while (c > 1 && c * (c - 1) / 2 > k) --c;
        while ((c + 1) * c / 2 <= k) ++c;
        j = (int)c;
        i = (int)(k - c * (c - 1) / 2);
    }
}

__device__ __forceinline__ bool body_pred(int i, int j, const int32_t* __restrict__ scene,
                                          const float* __restrict__ lo, const float* __restrict__ hi) {
    if (scene && scene[i] != scene[j]) return false;
    return overlap(lo + 3 * i, hi + 3 * i, lo + 3 * j, hi + 3 * j);
}

// Scene segmentation: pairs enumerated scene by scene (bodies of a scene are
// contiguous), so the concatenation stays sorted in both orders.
struct SceneSeg {
    int n;                 // scenes (<= 1: no segmentation)
    const int32_t* boff;   // [n+1] body offsets
    const int64_t* poff;   // [n+1] pair offsets
};

__device__ __forceinline__ void seg_pair_of(int64_t k, int nb, int order, const SceneSeg& sg, int& i, int& j) {
    if (sg.n <= 1) {
        pair_of(k, nb, order, i, j);
        return;
    }
    int lo = 0, hi = sg.n;  // last s with poff[s] <= k
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (sg.poff[mid] <= k)
            lo = mid;
        else
            hi = mid;
    }
    const int b0 = sg.boff[lo];
    pair_of(k - sg.poff[lo], sg.boff[lo + 1] - b0, order, i, j);
    i += b0;
    j += b0;
}

template <bool WRITE>
__global__ __launch_bounds__(kBlock) void body_pairs_kernel(int nb, int64_t npairs, int order, SceneSeg sg,
                                                             const int32_t* __restrict__ scene,
                                                             const float* __restrict__ lo,
                                                             const float* __restrict__ hi, int64_t* __restrict__ bcnt,
                                                             const int64_t* __restrict__ boff, int32_t* out,
                                                             int64_t cap) {
    __shared__ int shw[kBlock / kWave];
    int64_t base = (int64_t)blockIdx.x * kBlock * kPairItems;
    int64_t run = WRITE ? boff[blockIdx.x] : 0;
    for (int it = 0; it < kPairItems; ++it) {
        int64_t k = base + (int64_t)it * kBlock + threadIdx.x;
        bool pred = false;
        int i = 0, j = 0;
        if (k < npairs) {
            seg_pair_of(k, nb, order, sg, i, j);
            pred = body_pred(i, j, scene, lo, hi);
        }
        int tot;
        int r = block_rank(pred, shw, &tot);
        if (WRITE && pred) {
            int64_t o = run + r;
            if (o < cap) {
                out[2 * o] = i;
                out[2 * o + 1] = j;
            }
        }
        run += tot;
    }
    if (!WRITE && threadIdx.x == 0) bcnt[blockIdx.x] = run;
}

// ------------------------------------------------------------- primitives

// Hull of the FP32 boxes of each body's primitives: one wave per body (kBlock / 64 bodies
// per workgroup), the six extrema reduced by lane shuffles (a minimum or maximum does not
// depend on the order).  blockIdx.y selects one of two primitive sets (set 1 only when
// nsets == 1).
struct HullSet {
    const int32_t* offs;
    const float *lo, *hi;
    float *hlo, *hhi;
};
__global__ __launch_bounds__(kBlock) void hull_kernel(int nb, HullSet H1, HullSet H2) {
    const HullSet& H = blockIdx.y ? H2 : H1;
    const int b = blockIdx.x * (kBlock / kWave) + wave_id();
    if (b >= nb) return;  // the whole wave (no barrier below)
    const int p0 = H.offs[b], p1 = H.offs[b + 1];
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int p = p0 + lane_id(); p < p1; p += kWave) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            mn[k] = fminf(mn[k], H.lo[3 * (int64_t)p + k]);
            mx[k] = fmaxf(mx[k], H.hi[3 * (int64_t)p + k]);
        }
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            mn[k] = fminf(mn[k], __shfl_xor(mn[k], off, kWave));
            mx[k] = fmaxf(mx[k], __shfl_xor(mx[k], off, kWave));
        }
    if (lane_id() < 3) {
        const int k = lane_id();
        H.hlo[3 * b + k] = k == 0 ? mn[0] : (k == 1 ? mn[1] : mn[2]);
        H.hhi[3 * b + k] = k == 0 ? mx[0] : (k == 1 ? mx[1] : mx[2]);
    }
}

// unit -> (ii, I (set-1 body), J (set-2 body))
__device__ __forceinline__ void unit_bodies(int pt, int64_t u, const int32_t* __restrict__ ij, int64_t& ii, int& I,
                                            int& J) {
    if (pt) {
        ii = u >> 1;
        int a = ij[2 * ii], b = ij[2 * ii + 1];
        I = (u & 1) ? b : a;  // intersection_prims: odd threads take (ij1, ij0), culling.py:88-91
        J = a + b - I;
    } else {
        ii = u;
        I = ij[2 * ii];
        J = ij[2 * ii + 1];
    }
}

// One workgroup per (unit, side): side 0 keeps set-1 prims of I overlapping hull2[J],
// side 1 keeps set-2 prims of J overlapping hull1[I]; local ids in ascending order.
__global__ __launch_bounds__(kBlock) void filter_kernel(int pt, const int32_t* __restrict__ ij, PrimSets S,
                                                         int64_t maxs, int32_t* __restrict__ flist,
                                                         int64_t* __restrict__ fcnt) {
    __shared__ int shw[kBlock / kWave];
    const int64_t u = blockIdx.x >> 1;
    const int side = blockIdx.x & 1;
    int64_t ii;
    int I, J;
    unit_bodies(pt, u, ij, ii, I, J);
    const int32_t* offs = side ? S.off2 : S.off1;
    const float* lo = side ? S.lo2 : S.lo1;
    const float* hi = side ? S.hi2 : S.hi1;
    const int body = side ? J : I;
    const int other = side ? I : J;
    const float* hlo = side ? S.hlo1 + 3 * other : S.hlo2 + 3 * other;
    const float* hhi = side ? S.hhi1 + 3 * other : S.hhi2 + 3 * other;
    const int p0 = offs[body], p1 = offs[body + 1];
    int32_t* dst = flist + blockIdx.x * maxs;
    int64_t run = 0;
    for (int base = p0; base < p1; base += kBlock) {
        int p = base + threadIdx.x;
        bool pred = p < p1 && overlap(lo + 3 * (int64_t)p, hi + 3 * (int64_t)p, hlo, hhi);
        int tot;
        int r = block_rank(pred, shw, &tot);
        if (pred) dst[run + r] = p - p0;
        run += tot;
    }
    if (threadIdx.x == 0) fcnt[blockIdx.x] = run;
}

__global__ void unit_chunks_kernel(int64_t nunits, int chunk, const int64_t* __restrict__ fcnt,
                                   int64_t* __restrict__ nchunk) {
    int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nunits) return;
    int64_t nt = fcnt[2 * u] * fcnt[2 * u + 1];
    nchunk[u] = (nt + chunk - 1) / chunk;
}

// last u with choff[u] <= c (choff non-decreasing, choff[0] = 0)
__device__ __forceinline__ int64_t find_unit(const int64_t* __restrict__ choff, int64_t nunits, int64_t c) {
    int64_t lo = 0, hi = nunits;  // invariant: choff[lo] <= c, answer in [lo, hi)
    while (hi - lo > 1) {
        int64_t mid = (lo + hi) >> 1;
        if (choff[mid] <= c)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

template <bool WRITE>
__global__ __launch_bounds__(kBlock) void prim_pairs_kernel(int pt, int chunk, const int32_t* __restrict__ ij,
                                                             PrimSets S, int64_t maxs, const int32_t* __restrict__ flist,
                                                             const int64_t* __restrict__ fcnt,
                                                             const int64_t* __restrict__ choff, int64_t nunits,
                                                             const int64_t* __restrict__ total_chunks,
                                                             int64_t* __restrict__ ccnt,
                                                             const int64_t* __restrict__ ccoff,
                                                             int32_t* __restrict__ out, int64_t cap) {
    __shared__ int shw[kBlock / kWave];
    const int64_t T = *total_chunks;
    for (int64_t c = blockIdx.x; c < T; c += gridDim.x) {
        int64_t u = find_unit(choff, nunits, c);
        int64_t ii;
        int I, J;
        unit_bodies(pt, u, ij, ii, I, J);
        const uint32_t nA = (uint32_t)fcnt[2 * u], nB = (uint32_t)fcnt[2 * u + 1];
        const uint32_t nt = nA * nB;
        const int32_t* LA = flist + (2 * u) * maxs;
        const int32_t* LB = flist + (2 * u + 1) * maxs;
        const int oA = S.off1[I], oB = S.off2[J];
        const uint32_t f0 = (uint32_t)(c - choff[u]) * (uint32_t)chunk;
        const uint32_t f1 = f0 + (uint32_t)chunk < nt ? f0 + (uint32_t)chunk : nt;
        int64_t run = WRITE ? ccoff[c] : 0;
        for (uint32_t base = f0; base < f1; base += kBlock) {
            uint32_t f = base + threadIdx.x;
            bool pred = false;
            int32_t a = 0, b = 0;
            if (f < f1) {
                uint32_t i = f / nB, j = f - i * nB;
                a = LA[i];
                b = LB[j];
                int64_t ga = oA + a, gb = oB + b;
                pred = overlap(S.lo1 + 3 * ga, S.hi1 + 3 * ga, S.lo2 + 3 * gb, S.hi2 + 3 * gb);
            }
            int tot;
            int r = block_rank(pred, shw, &tot);
            if (WRITE && pred) {
                int64_t o = run + r;
                if (o < cap) {
                    int32_t* t5 = out + 5 * o;
                    t5[0] = I;
                    t5[1] = J;
                    t5[2] = (int32_t)ii;
                    t5[3] = a;
                    t5[4] = b;
                }
            }
            run += tot;
        }
        if (!WRITE && threadIdx.x == 0) ccnt[c] = run;
    }
}

// ------------------------------------------------------------------- VG
template <bool WRITE>
__global__ __launch_bounds__(kBlock) void vg_kernel(int64_t nv, const int32_t* __restrict__ vert_body,
                                                     const int32_t* __restrict__ voff, const float* __restrict__ blo,
                                                     const float* __restrict__ plo, const float* __restrict__ phi,
                                                     double ythresh, float fground, int64_t* __restrict__ bcnt,
                                                     const int64_t* __restrict__ boff, int32_t* __restrict__ out,
                                                     int64_t cap) {
    __shared__ int shw[kBlock / kWave];
    const float qlo[3] = {-1e3f, -1e3f, -1e3f};
    const float qhi[3] = {1e3f, fground, 1e3f};
    int64_t base = (int64_t)blockIdx.x * kBlock * kItems;
    int64_t run = WRITE ? boff[blockIdx.x] : 0;
    for (int it = 0; it < kItems; ++it) {
        int64_t v = base + (int64_t)it * kBlock + threadIdx.x;
        bool pred = false;
        int b = 0;
        if (v < nv) {
            b = vert_body[v];
            // culling.py:67: lowers[i][1] < 0.5 * dhat + ground (FP64)
            pred = (double)blo[3 * b + 1] < ythresh && overlap(plo + 3 * v, phi + 3 * v, qlo, qhi);
        }
        int tot;
        int r = block_rank(pred, shw, &tot);
        if (WRITE && pred) {
            int64_t o = run + r;
            if (o < cap) {
                out[2 * o] = b;
                out[2 * o + 1] = (int32_t)(v - voff[b]);
            }
        }
        run += tot;
    }
    if (!WRITE && threadIdx.x == 0) bcnt[blockIdx.x] = run;
}

// ------------------------------------------------------------ structure
__global__ void triplets_kernel(int nb, int64_t nij, const int32_t* __restrict__ ij, int32_t* __restrict__ rows,
                                int32_t* __restrict__ cols) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t nslot = nb + 2 * nij;
    if (t >= 16 * nslot) return;
    int64_t i = t / 16;
    int e = (int)(t % 16), ii = e % 4, jj = e / 4;
    int64_t I, J;
    if (i < nb) {
        I = J = i;
    } else if (i < nb + nij) {
        I = ij[2 * (i - nb)];
        J = ij[2 * (i - nb) + 1];
    } else {
        I = ij[2 * (i - nb - nij) + 1];
        J = ij[2 * (i - nb - nij)];
    }
    rows[16 * i + ii + jj * 4] = (int32_t)(ii + 4 * I);
    cols[16 * i + ii + jj * 4] = (int32_t)(jj + 4 * J);
}

// first k in [lo, hi) with key(k) >= v; keys non-decreasing, key(k) = a[2k + comp]
__device__ __forceinline__ int64_t lower_bound2(const int32_t* __restrict__ a, int comp, int64_t lo, int64_t hi,
                                                int v) {
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (a[2 * mid + comp] < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// dnij (nullable): the pair count on the device (nij is then unused)
__global__ void rowlen_kernel(int nb, int64_t nij, const int64_t* __restrict__ dnij, const int32_t* __restrict__ ij,
                              const int32_t* __restrict__ ji, int64_t* __restrict__ len) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    if (dnij) nij = *dnij;
    int64_t nu = lower_bound2(ij, 0, 0, nij, b + 1) - lower_bound2(ij, 0, 0, nij, b);
    int64_t nl = lower_bound2(ji, 1, 0, nij, b + 1) - lower_bound2(ji, 1, 0, nij, b);
    len[b] = nl + 1 + nu;
}

__global__ void rowfill_kernel(int nb, int64_t nij, const int64_t* __restrict__ dnij, const int32_t* __restrict__ ij,
                               const int32_t* __restrict__ ji, const int64_t* __restrict__ off,
                               int32_t* __restrict__ row_ptr, int32_t* __restrict__ cols,
                               int32_t* __restrict__ slots) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    if (dnij) nij = *dnij;
    const int64_t total = (int64_t)nb + 2 * nij;
    int64_t o = off[b];
    row_ptr[b] = (int32_t)o;
    if (b == nb - 1) row_ptr[nb] = (int32_t)total;
    int64_t l0 = lower_bound2(ji, 1, 0, nij, b), l1 = lower_bound2(ji, 1, 0, nij, b + 1);
    for (int64_t k = l0; k < l1; ++k) {
        int i = ji[2 * k];
        int64_t u0 = lower_bound2(ij, 0, 0, nij, i), u1 = lower_bound2(ij, 0, 0, nij, i + 1);
        int64_t ii = lower_bound2(ij, 1, u0, u1, b);
        cols[o] = i;
        slots[o] = (int32_t)(nb + nij + ii);
        ++o;
    }
    cols[o] = b;
    slots[o] = b;
    ++o;
    int64_t u0 = lower_bound2(ij, 0, 0, nij, b), u1 = lower_bound2(ij, 0, 0, nij, b + 1);
    for (int64_t k = u0; k < u1; ++k) {
        cols[o] = ij[2 * k + 1];
        slots[o] = (int32_t)(nb + k);
        ++o;
    }
}

// ------------------------------------------------------- staged culling
// The primitive-pair culling in three stages that leave their counts on the device,
// so that a caller can run several cullings side by side and read all their counts in
// one host synchronization.  `set` selects one of two independent workspaces.
namespace {
const int kCullSlots[2][8] = {
    {kSlotHull1, kSlotHull2, kSlotFilter, kSlotFilterCount, kSlotScanA, kSlotScanB, kSlotChunkCount, kSlotChunk},
    {kSlotHull1B, kSlotHull2B, kSlotFilterB, kSlotFilterCountB, kSlotScanAB, kSlotScanBB, kSlotChunkCountB,
     kSlotChunkB}};
}  // namespace

int prim_cull_begin(hipStream_t s, const AbdScene* sc, int mode, const int32_t* ij, int64_t nij, const float* lo1,
                    const float* hi1, const float* lo2, const float* hi2, int set, PrimCull& C,
                    int64_t* total_copy) {
    const int* slot = kCullSlots[set];
    C.set = set;
    C.pt = mode == ABD_CULL_PT;
    C.ij = ij;
    const int nb = sc->n_bodies;
    PrimSets& S = C.S;
    S.off1 = C.pt ? sc->voff : sc->eoff;
    S.off2 = C.pt ? sc->foff : sc->eoff;
    S.lo1 = lo1;
    S.hi1 = hi1;
    S.lo2 = lo2;
    S.hi2 = hi2;
    const int64_t max1 = C.pt ? sc->max_verts : sc->max_edges;
    const int64_t max2 = C.pt ? sc->max_tris : sc->max_edges;
    C.maxs = max1 > max2 ? max1 : max2;
    if (C.maxs > 65535) return ABD_ERR_ARG;  // n1 * n2 tests per body pair must fit in 32 bits
    C.nunits = C.pt ? 2 * nij : nij;
    if (2 * C.nunits > INT32_MAX) return ABD_ERR_ARG;
    int rc;
    float *h1, *h2;
    if ((rc = arena_get(slot[0], sizeof(float) * 6 * (size_t)nb, (void**)&h1))) return rc;
    if ((rc = arena_get(slot[1], sizeof(float) * 6 * (size_t)nb, (void**)&h2))) return rc;
    S.hlo1 = h1;
    S.hhi1 = h1 + 3 * nb;
    S.hlo2 = h2;
    S.hhi2 = h2 + 3 * nb;
    // both hulls in one launch; the EE culling compares one set of boxes with itself
    const bool same = S.off1 == S.off2 && lo1 == lo2 && hi1 == hi2;
    if (same) {
        S.hlo2 = S.hlo1;
        S.hhi2 = S.hhi1;
    }
    if (nb > 0) {
        const HullSet H1{S.off1, lo1, hi1, h1, h1 + 3 * nb}, H2{S.off2, lo2, hi2, h2, h2 + 3 * nb};
        hipLaunchKernelGGL(hull_kernel, dim3((unsigned)ceil_div(nb, kBlock / kWave), same ? 1 : 2), dim3(kBlock), 0,
                           s, nb, H1, H2);
        ABD_CHECK_LAUNCH();
    }
    if ((rc = arena_get(slot[2], sizeof(int32_t) * (size_t)(2 * C.nunits * C.maxs), (void**)&C.flist))) return rc;
    if ((rc = arena_get(slot[3], sizeof(int64_t) * (size_t)(2 * C.nunits), (void**)&C.fcnt))) return rc;
    if ((rc = arena_get(slot[4], sizeof(int64_t) * (size_t)(C.nunits + 1), (void**)&C.nchunk))) return rc;
    if ((rc = arena_get(slot[5], sizeof(int64_t) * (size_t)(C.nunits + 1), (void**)&C.choff))) return rc;
    int ph = prof_start(kProfFilter, s, 0.0);
    hipLaunchKernelGGL(filter_kernel, dim3((unsigned)(2 * C.nunits)), dim3(kBlock), 0, s, C.pt, ij, S, C.maxs,
                       C.flist, C.fcnt);
    ABD_CHECK_LAUNCH();
    prof_stop(ph, s);
    C.tune = tuning();
    hipLaunchKernelGGL(unit_chunks_kernel, dim3(grid_for(C.nunits)), dim3(kBlock), 0, s, C.nunits, C.tune.cull_chunk,
                       C.fcnt, C.nchunk);
    ABD_CHECK_LAUNCH();
    if ((rc = scan_exclusive_i64(s, C.nchunk, C.choff, C.nunits, C.choff + C.nunits, total_copy))) return rc;
    C.total_chunks = C.choff + C.nunits;
    C.total_pairs = nullptr;
    return ABD_OK;
}

int prim_cull_count(hipStream_t s, PrimCull& C, int64_t T, int64_t* total_copy) {
    const int* slot = kCullSlots[C.set];
    C.T = T;
    if (T == 0) return ABD_OK;
    int rc;
    if ((rc = arena_get(slot[6], sizeof(int64_t) * (size_t)(T + 1), (void**)&C.ccnt))) return rc;
    if ((rc = arena_get(slot[7], sizeof(int64_t) * (size_t)(T + 1), (void**)&C.ccoff))) return rc;
    C.grid = (unsigned)(T < C.tune.cull_grid ? T : C.tune.cull_grid);
    int ph = prof_start(kProfPrimCount, s, 0.0);
    hipLaunchKernelGGL(prim_pairs_kernel<false>, dim3(C.grid), dim3(kBlock), 0, s, C.pt, C.tune.cull_chunk, C.ij, C.S,
                       C.maxs, (const int32_t*)C.flist, (const int64_t*)C.fcnt, (const int64_t*)C.choff, C.nunits,
                       (const int64_t*)C.total_chunks, C.ccnt, (const int64_t*)nullptr, (int32_t*)nullptr, (int64_t)0);
    ABD_CHECK_LAUNCH();
    prof_stop(ph, s);
    if ((rc = scan_exclusive_i64(s, C.ccnt, C.ccoff, T, C.ccoff + T, total_copy))) return rc;
    C.total_pairs = C.ccoff + T;
    return ABD_OK;
}

int prim_cull_fill(hipStream_t s, PrimCull& C, int32_t* out, int64_t cap, int64_t n) {
    if (C.T == 0 || n == 0) return ABD_OK;
    int ph = prof_start(kProfPrimFill, s, 20.0 * n);
    hipLaunchKernelGGL(prim_pairs_kernel<true>, dim3(C.grid), dim3(kBlock), 0, s, C.pt, C.tune.cull_chunk, C.ij, C.S,
                       C.maxs, (const int32_t*)C.flist, (const int64_t*)C.fcnt, (const int64_t*)C.choff, C.nunits,
                       (const int64_t*)C.total_chunks, C.ccnt, (const int64_t*)C.ccoff, out, cap);
    ABD_CHECK_LAUNCH();
    prof_stop(ph, s);
    return ABD_OK;
}

int vg_cull_begin(hipStream_t s, const AbdScene* sc, const float* blo, const float* plo, const float* phi, double dhat,
                  double ground, VgCull& C,
                  int64_t* total_copy) {
    C.sc = sc;
    C.blo = blo;
    C.plo = plo;
    C.phi = phi;
    C.yth = 0.5 * dhat + ground;
    C.fg = (float)ground;
    const int64_t nv = sc->n_verts;
    C.nblk = ceil_div(nv, (int64_t)kBlock * kItems);
    C.total = nullptr;
    if (nv == 0) return ABD_OK;
    int rc;
    if ((rc = arena_get(kSlotVgA, sizeof(int64_t) * (C.nblk + 1), (void**)&C.bcnt))) return rc;
    if ((rc = arena_get(kSlotVgB, sizeof(int64_t) * (C.nblk + 1), (void**)&C.boff))) return rc;
    int ph = prof_start(kProfVg, s, 28.0 * nv);
    hipLaunchKernelGGL(vg_kernel<false>, dim3((unsigned)C.nblk), dim3(kBlock), 0, s, nv, sc->vert_body, sc->voff, blo,
                       plo, phi, C.yth, C.fg, C.bcnt, (const int64_t*)nullptr, (int32_t*)nullptr, (int64_t)0);
    ABD_CHECK_LAUNCH();
    prof_stop(ph, s);
    if ((rc = scan_exclusive_i64(s, C.bcnt, C.boff, C.nblk, C.boff + C.nblk, total_copy))) return rc;
    C.total = C.boff + C.nblk;
    return ABD_OK;
}

int vg_cull_fill(hipStream_t s, VgCull& C, int32_t* out, int64_t cap, int64_t n) {
    if (n == 0 || !C.total) return ABD_OK;
    const AbdScene* sc = C.sc;
    hipLaunchKernelGGL(vg_kernel<true>, dim3((unsigned)C.nblk), dim3(kBlock), 0, s, (int64_t)sc->n_verts,
                       sc->vert_body, sc->voff, C.blo, C.plo, C.phi, C.yth, C.fg, C.bcnt, (const int64_t*)C.boff, out,
                       cap);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

// abd_cull_bodies; ji == nullptr skips the (j, i) order.
int cull_bodies(hipStream_t s, const AbdScene* sc, const float* lo, const float* hi, int32_t* ij, int32_t* ji,
                int64_t cap, int64_t* count) {
    const int nb = sc->n_bodies;
    SceneSeg sg{0, nullptr, nullptr};
    int64_t npairs = (int64_t)nb * (nb - 1) / 2;
    if (sc->n_scenes > 1 && sc->scene_boff && sc->scene_poff) {
        sg = SceneSeg{sc->n_scenes, sc->scene_boff, sc->scene_poff};
        npairs = sc->n_scene_pairs;
    }
    *count = 0;
    if (npairs <= 0) return ABD_OK;
    const int64_t nblk = ceil_div(npairs, (int64_t)kBlock * kPairItems);
    if (nblk > INT32_MAX) return ABD_ERR_ARG;
    int64_t *bcnt, *boff;
    int rc = arena_get(kSlotScanA, sizeof(int64_t) * (nblk + 1), (void**)&bcnt);
    if (rc) return rc;
    rc = arena_get(kSlotScanB, sizeof(int64_t) * (nblk + 1), (void**)&boff);
    if (rc) return rc;
    int ph = prof_start(kProfBodyPairs, s, 24.0 * nb);
    hipLaunchKernelGGL(body_pairs_kernel<false>, dim3((unsigned)nblk), dim3(kBlock), 0, s, nb, npairs, 0, sg,
                       sc->body_scene, lo, hi, bcnt, (const int64_t*)nullptr, (int32_t*)nullptr, (int64_t)0);
    ABD_CHECK_LAUNCH();
    prof_stop(ph, s);
    if ((rc = scan_exclusive_i64(s, bcnt, boff, nblk, boff + nblk))) return rc;
    int64_t n;
    if ((rc = read_i64(s, boff + nblk, &n))) return rc;
    *count = n;
    if (n > cap) return ABD_ERR_OVERFLOW;
    if (n == 0) return ABD_OK;
    hipLaunchKernelGGL(body_pairs_kernel<true>, dim3((unsigned)nblk), dim3(kBlock), 0, s, nb, npairs, 0, sg,
                       sc->body_scene, lo, hi, bcnt, (const int64_t*)boff, ij, cap);
    ABD_CHECK_LAUNCH();
    if (!ji) return ABD_OK;
    // (j, i) order: same predicate, second enumeration
    hipLaunchKernelGGL(body_pairs_kernel<false>, dim3((unsigned)nblk), dim3(kBlock), 0, s, nb, npairs, 1, sg,
                       sc->body_scene, lo, hi, bcnt, (const int64_t*)nullptr, (int32_t*)nullptr, (int64_t)0);
    ABD_CHECK_LAUNCH();
    if ((rc = scan_exclusive_i64(s, bcnt, boff, nblk, boff + nblk))) return rc;
    hipLaunchKernelGGL(body_pairs_kernel<true>, dim3((unsigned)nblk), dim3(kBlock), 0, s, nb, npairs, 1, sg,
                       sc->body_scene, lo, hi, bcnt, (const int64_t*)boff, ji, cap);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}


}  // namespace abd

using namespace abd;

extern "C" {

int abd_cull_bodies(abd_stream_t stream, const AbdScene* sc, const float* lo, const float* hi, int32_t* ij,
                    int32_t* ji, int64_t cap, int64_t* count) {
    if (!sc || !lo || !hi || !count || cap < 0 || (cap > 0 && (!ij || !ji))) return ABD_ERR_ARG;
    return cull_bodies(to_stream(stream), sc, lo, hi, ij, ji, cap, count);
}

int abd_cull_prims(abd_stream_t stream, const AbdScene* sc, int32_t mode, const int32_t* ij, int64_t nij,
                   const float* lo1, const float* hi1, const float* lo2, const float* hi2, int32_t* out, int64_t cap,
                   int64_t* count) {
    if (!sc || !count || nij < 0 || cap < 0 || (nij > 0 && !ij) || !lo1 || !hi1 || !lo2 || !hi2 ||
        (cap > 0 && !out))
        return ABD_ERR_ARG;
    if (mode != ABD_CULL_PT && mode != ABD_CULL_EE) return ABD_ERR_ARG;
    *count = 0;
    if (nij == 0) return ABD_OK;
    hipStream_t s = to_stream(stream);
    PrimCull C;
    int rc = prim_cull_begin(s, sc, mode, ij, nij, lo1, hi1, lo2, hi2, 0, C);
    if (rc) return rc;
    int64_t T;
    if ((rc = read_i64(s, C.total_chunks, &T))) return rc;
    if ((rc = prim_cull_count(s, C, T))) return rc;
    if (T == 0) return ABD_OK;
    int64_t n;
    if ((rc = read_i64(s, C.total_pairs, &n))) return rc;
    *count = n;
    if (n > cap) return ABD_ERR_OVERFLOW;
    if (n == 0) return ABD_OK;
    return prim_cull_fill(s, C, out, cap, n);
}

int abd_cull_vg(abd_stream_t stream, const AbdScene* sc, const float* blo, const float* plo, const float* phi,
                double dhat, double ground, int32_t* out, int64_t cap, int64_t* count) {
    if (!sc || !blo || !plo || !phi || !count || cap < 0 || (cap > 0 && !out)) return ABD_ERR_ARG;
    hipStream_t s = to_stream(stream);
    *count = 0;
    if (sc->n_verts == 0) return ABD_OK;
    VgCull C;
    int rc = vg_cull_begin(s, sc, blo, plo, phi, dhat, ground, C);
    if (rc) return rc;
    int64_t n;
    if ((rc = read_i64(s, C.total, &n))) return rc;
    *count = n;
    if (n > cap) return ABD_ERR_OVERFLOW;
    return vg_cull_fill(s, C, out, cap, n);
}

int abd_set_triplets(abd_stream_t stream, int32_t nb, const int32_t* ij, int64_t nij, int32_t* rows,
                     int32_t* cols) {
    if (nb < 0 || nij < 0 || (nij > 0 && !ij) || !rows || !cols) return ABD_ERR_ARG;
    int64_t n = 16 * ((int64_t)nb + 2 * nij);
    if (n == 0) return ABD_OK;
    hipLaunchKernelGGL(triplets_kernel, dim3(grid_for(n, kBlock, INT32_MAX)), dim3(kBlock), 0, to_stream(stream), nb,
                       nij, ij, rows, cols);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

int abd_bsr_structure(abd_stream_t stream, int32_t nb, const int32_t* ij, const int32_t* ji, int64_t nij,
                      int32_t* row_ptr, int32_t* cols, int32_t* slots) {
    if (nb <= 0 || nij < 0 || (nij > 0 && (!ij || !ji)) || !row_ptr || !cols || !slots) return ABD_ERR_ARG;
    return bsr_structure(to_stream(stream), nb, ij, ji, nij, nullptr, row_ptr, cols, slots);
}

}  // extern "C"

namespace abd {

int bsr_structure(hipStream_t s, int nb, const int32_t* ij, const int32_t* ji, int64_t nij, const int64_t* dnij,
                  int32_t* row_ptr, int32_t* cols, int32_t* slots) {
    int64_t *len, *off;
    int rc;
    if ((rc = arena_get(kSlotScanA, sizeof(int64_t) * (nb + 1), (void**)&len))) return rc;
    if ((rc = arena_get(kSlotScanB, sizeof(int64_t) * (nb + 1), (void**)&off))) return rc;
    hipLaunchKernelGGL(rowlen_kernel, dim3(grid_for(nb)), dim3(kBlock), 0, s, nb, nij, dnij, ij, ji, len);
    ABD_CHECK_LAUNCH();
    if ((rc = scan_exclusive_i64(s, len, off, nb, off + nb))) return rc;
    hipLaunchKernelGGL(rowfill_kernel, dim3(grid_for(nb)), dim3(kBlock), 0, s, nb, nij, dnij, ij, ji,
                       (const int64_t*)off, row_ptr, cols, slots);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

// ------------------------------------------------- collision set from a superset
// The pairs and candidates of a candidate set made from larger boxes (the swept set of
// the last step: boxes over x - alpha dq, alpha in [0, 1], dilated as the collision set's)
// that pass the culling predicates on the current boxes.  When the current boxes lie
// inside those the superset was made from, this is the culling's own output: the same
// predicates on the same boxes, and the superset's sorted order restricted to the kept
// entries (pair indices renumbered through `remap`).
namespace {

// The five lists of a keep pass: 0 ij, 1 ji, 2 PT, 3 EE, 4 VG.
constexpr int kKeepLists = 5;
struct KeepArgs {
    KeepInput in;
    KeepBoxes bx;
    KeepOutput out;
    const int32_t *voff, *foff, *eoff;
    int32_t* remap;
    double yth;
    float fg;
    int64_t n[kKeepLists];
    int64_t blk0[kKeepLists + 1];  // first workgroup (block of kBlock entries) of each list
    int64_t* bcnt[kKeepLists];
    int64_t* boff[kKeepLists];
    uint8_t* flag;                 // each entry's predicate (count pass), read by the write pass
    int64_t flag0[kKeepLists];     // first flag of each list
    // primitive boxes from the vertices when x is set (KeepAtX)
    const double *x, *x1;
    const int32_t *tris, *edges;
    float pdil;
    double edil, tdil;
};

// point_boxes_kernel's box of vertex v at x, or with x1 set swept_point_boxes_kernel's
// (the union with the box at x1)
__device__ __forceinline__ void point_box_at(const double* __restrict__ x, const double* __restrict__ x1,
                                             int64_t v, float dil, float* lo, float* hi) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float c = (float)x[3 * v + k];
        if (x1) {
            const float c1 = (float)x1[3 * v + k];
            lo[k] = fminf(c - dil, c1 - dil);
            hi[k] = fmaxf(c + dil, c1 + dil);
        } else {
            lo[k] = c - dil;
            hi[k] = c + dil;
        }
    }
}

// prim_boxes_kernel's box of the primitive with scene vertices vi[0..NV) (its topology's
// order) at x, or with x1 set swept_prim_boxes_kernel's
template <int NV>
__device__ __forceinline__ void prim_box_of(const double* __restrict__ x, const double* __restrict__ x1,
                                            const int64_t (&vi)[NV], double dil, float* lo, float* hi) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double mn = x[3 * vi[0] + k], mx = mn;
#pragma unroll
        for (int j = 1; j < NV; ++j) {
            const double c = x[3 * vi[j] + k];
            mn = c < mn ? c : mn;
            mx = c > mx ? c : mx;
        }
        float l = (float)(mn - dil), h = (float)(mx + dil);
        if (x1) {
            double mn1 = x1[3 * vi[0] + k], mx1 = mn1;
#pragma unroll
            for (int j = 1; j < NV; ++j) {
                const double c = x1[3 * vi[j] + k];
                mn1 = c < mn1 ? c : mn1;
                mx1 = c > mx1 ? c : mx1;
            }
            l = fminf(l, (float)(mn1 - dil));
            h = fmaxf(h, (float)(mx1 + dil));
        }
        lo[k] = l;
        hi[k] = h;
    }
}
// the same for primitive i (NV vertices from topo, numbered from vb)
template <int NV>
__device__ __forceinline__ void prim_box_at(const double* __restrict__ x, const double* __restrict__ x1,
                                            const int32_t* __restrict__ topo, int64_t i, int vb, double dil,
                                            float* lo, float* hi) {
    int64_t vi[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) vi[j] = vb + topo[NV * i + j];
    prim_box_of<NV>(x, x1, vi, dil, lo, hi);
}

__device__ __forceinline__ bool keep_pred(const KeepArgs& A, int k, int64_t i) {
    const KeepBoxes& B = A.bx;
    if (k <= 1) {
        const int32_t* pr = (k == 0 ? A.in.ij : A.in.ji) + 2 * i;
        const int a = pr[0], b = pr[1];
        return overlap(B.blo + 3 * a, B.bhi + 3 * a, B.blo + 3 * b, B.bhi + 3 * b);
    }
    if (k <= 3) {  // (I, J, ii, a, b): both bodies' boxes and the two primitives' boxes
        const int32_t* t5 = (k == 2 ? A.in.pt : A.in.ee) + 5 * i;
        const int I = t5[0], J = t5[1];
        const int32_t* vid = k == 2 ? A.in.vids.pt : A.in.vids.ee;
        if (A.x && vid) {  // the candidate's vertex ids stored beside the list
            vid += 4 * i;
            const int64_t v0 = vid[0], v1 = vid[1], v2 = vid[2], v3 = vid[3];
            if (!overlap(B.blo + 3 * I, B.bhi + 3 * I, B.blo + 3 * J, B.bhi + 3 * J)) return false;
            float l1[3], h1[3], l2[3], h2[3];
            if (k == 2) {
                point_box_at(A.x, A.x1, v0, A.pdil, l1, h1);
                const int64_t tv[3] = {v1, v2, v3};
                prim_box_of<3>(A.x, A.x1, tv, A.tdil, l2, h2);
            } else {
                const int64_t ea[2] = {v0, v1}, eb[2] = {v2, v3};
                prim_box_of<2>(A.x, A.x1, ea, A.edil, l1, h1);
                prim_box_of<2>(A.x, A.x1, eb, A.edil, l2, h2);
            }
            return overlap(l1, h1, l2, h2);
        }
        if (!overlap(B.blo + 3 * I, B.bhi + 3 * I, B.blo + 3 * J, B.bhi + 3 * J)) return false;
        const int64_t a = (k == 2 ? A.voff : A.eoff)[I] + (int64_t)t5[3];
        const int64_t b = (k == 2 ? A.foff : A.eoff)[J] + (int64_t)t5[4];
        if (A.x) {
            // a: a point of I (PT) or an edge of I; b: a triangle or an edge of J
            float l1[3], h1[3], l2[3], h2[3];
            if (k == 2) {
                point_box_at(A.x, A.x1, a, A.pdil, l1, h1);
                prim_box_at<3>(A.x, A.x1, A.tris, b, A.voff[J], A.tdil, l2, h2);
            } else {
                prim_box_at<2>(A.x, A.x1, A.edges, a, A.voff[I], A.edil, l1, h1);
                prim_box_at<2>(A.x, A.x1, A.edges, b, A.voff[J], A.edil, l2, h2);
            }
            return overlap(l1, h1, l2, h2);
        }
        const float *lo1 = k == 2 ? B.plo : B.elo, *hi1 = k == 2 ? B.phi : B.ehi;
        const float *lo2 = k == 2 ? B.tlo : B.elo, *hi2 = k == 2 ? B.thi : B.ehi;
        return overlap(lo1 + 3 * a, hi1 + 3 * a, lo2 + 3 * b, hi2 + 3 * b);
    }
    // VG (b, v): the predicate of vg_kernel
    const float qlo[3] = {-1e3f, -1e3f, -1e3f};
    const float qhi[3] = {1e3f, A.fg, 1e3f};
    const int b = A.in.vg[2 * i];
    const int64_t v = A.voff[b] + (int64_t)A.in.vg[2 * i + 1];
    if (!((double)B.blo[3 * b + 1] < A.yth)) return false;
    if (A.x) {
        float l[3], h[3];
        point_box_at(A.x, A.x1, v, A.pdil, l, h);
        return overlap(l, h, qlo, qhi);
    }
    return overlap(B.plo + 3 * v, B.phi + 3 * v, qlo, qhi);
}

__device__ __forceinline__ void keep_write(const KeepArgs& A, int k, int64_t i, int64_t o) {
    if (k <= 1) {
        const int32_t* pr = (k == 0 ? A.in.ij : A.in.ji) + 2 * i;
        int32_t* q = (k == 0 ? A.out.ij : A.out.ji) + 2 * o;
        q[0] = pr[0];
        q[1] = pr[1];
    } else if (k <= 3) {
        const int32_t* t5 = (k == 2 ? A.in.pt : A.in.ee) + 5 * i;
        int32_t* q = (k == 2 ? A.out.pt : A.out.ee) + 5 * o;
        q[0] = t5[0];
        q[1] = t5[1];
        // the pair's index in the kept ij: its block's offset + its rank in the block (the
        // tuple is kept only if its pair's body boxes overlap, so the pair is kept)
        q[2] = (int32_t)(A.boff[0][t5[2] / kBlock] + A.remap[t5[2]]);
        q[3] = t5[3];
        q[4] = t5[4];
        const int32_t* vin = k == 2 ? A.in.vids.pt : A.in.vids.ee;
        int32_t* vout = k == 2 ? A.out.ptv : A.out.eev;
        if (vin && vout) {
#pragma unroll
            for (int j = 0; j < 4; ++j) vout[4 * o + j] = vin[4 * i + j];
        }
    } else {
        A.out.vg[2 * o] = A.in.vg[2 * i];
        A.out.vg[2 * o + 1] = A.in.vg[2 * i + 1];
    }
}

// The count: one entry per thread (its predicate's gathers need many waves in flight to
// hide their latency), the kept entries of each workgroup's kBlock entries to bcnt; each
// entry's predicate is stored for the write pass, and each kept pair's rank in its
// workgroup's kept pairs (ij) to remap, so that the write pass places pairs and tuples in one
// launch.
// entries of input list k: A.n[k], or with the input sizes on the device (A.in.dn, the
// keep output layout) that count (A.n[k] then bounds it)
__device__ __forceinline__ int64_t keep_n(const KeepArgs& A, int k) {
    return A.in.dn ? (k == 1 && !A.in.ji ? 0 : A.in.dn[k]) : A.n[k];
}

__global__ __launch_bounds__(kBlock) void keep_count_kernel(KeepArgs A) {
    __shared__ int shw[kBlock / kWave];
    const int64_t g = blockIdx.x;
    int k = 0;
    while (k + 1 < kKeepLists && g >= A.blk0[k + 1]) ++k;
    const int64_t blk = g - A.blk0[k];
    const int64_t i = blk * kBlock + threadIdx.x;
    bool pred = false;
    const bool in_list = i < keep_n(A, k);
    if (in_list) {
        pred = keep_pred(A, k, i);
        A.flag[A.flag0[k] + i] = pred ? 1 : 0;
    }
    if (k == 0) {  // uniform over the workgroup
        int tot;
        const int r = block_rank(pred, shw, &tot);
        if (in_list) A.remap[i] = pred ? r : -1;
        if (threadIdx.x == 0) A.bcnt[k][blk] = tot;
        return;
    }
    const int c = __popcll(__ballot(pred ? 1 : 0));
    if (lane_id() == 0) shw[wave_id()] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < kBlock / kWave; ++w) t += shw[w];
        A.bcnt[k][blk] = t;
    }
}

// The write pass over lists [k0, k1), one entry per thread: the kept entries at their
// ordered ranks (the block's scanned offset + the rank in the block).
__global__ __launch_bounds__(kBlock) void keep_write_kernel(KeepArgs A, int k0, int k1) {
    __shared__ int shw[kBlock / kWave];
    const int64_t g = (int64_t)blockIdx.x + A.blk0[k0];
    int k = k0;
    while (k + 1 < k1 && g >= A.blk0[k + 1]) ++k;
    const int64_t blk = g - A.blk0[k];
    const int64_t i = blk * kBlock + threadIdx.x;
    const int64_t nk = keep_n(A, k);
    const bool pred = i < nk && A.flag[A.flag0[k] + i] != 0;
    int tot;
    const int64_t r = A.boff[k][blk] + block_rank(pred, shw, &tot);
    if (pred) keep_write(A, k, i, r);
}

}  // namespace

namespace {
__global__ __launch_bounds__(kBlock) void cand_vid_kernel(SceneView V, int kind, const int32_t* __restrict__ tuples,
                                                          int64_t n, int32_t* __restrict__ vid) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        int64_t v[4];
        cand_idx(kind, V, tuples + 5 * i, nullptr, v);
#pragma unroll
        for (int j = 0; j < 4; ++j) vid[4 * i + j] = (int32_t)v[j];
    }
}
}  // namespace

int cand_vertex_ids(hipStream_t s, const AbdScene* sc, int kind, const int32_t* tuples, int64_t n, int32_t* vid) {
    if (n <= 0) return ABD_OK;
    const SceneView V{sc->voff, sc->foff, sc->eoff, sc->tris, sc->edges, sc->x0};
    hipLaunchKernelGGL(cand_vid_kernel, dim3(grid_for(n, kBlock, 1 << 16)), dim3(kBlock), 0, s, V, kind, tuples, n, vid);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

int keep_from_superset(hipStream_t s, const AbdScene* sc, const KeepInput& in, const KeepBoxes& bx, double dhat,
                       double ground, const KeepOutput& out, int64_t* dcount, int64_t* hcount, bool sync,
                       const KeepAtX* at) {
    KeepArgs A;
    A.in = in;
    A.bx = bx;
    A.out = out;
    A.voff = sc->voff;
    A.foff = sc->foff;
    A.eoff = sc->eoff;
    A.x = at ? at->x : nullptr;
    A.x1 = at ? at->x1 : nullptr;
    A.tris = sc->tris;
    A.edges = sc->edges;
    A.pdil = at ? (float)at->pdil : 0.0f;
    A.edil = at ? at->edil : 0.0;
    A.tdil = at ? at->tdil : 0.0;
    A.yth = 0.5 * dhat + ground;
    A.fg = (float)ground;
    const int64_t n[kKeepLists] = {in.nij, in.ji ? in.nij : 0, in.npt, in.nee, in.nvg};
    int64_t nblk[kKeepLists], total = 0, nflag = 0;
    A.blk0[0] = 0;
    for (int k = 0; k < kKeepLists; ++k) {
        A.flag0[k] = nflag;
        nflag += n[k];
        A.n[k] = n[k];
        nblk[k] = ceil_div(n[k], (int64_t)kBlock);
        A.blk0[k + 1] = A.blk0[k] + nblk[k];
        total += 2 * (nblk[k] + 1);
    }
    if (A.blk0[kKeepLists] > INT32_MAX) return ABD_ERR_ARG;
    int64_t* buf;
    int rc;
    if ((rc = arena_get(kSlotKeep, sizeof(int64_t) * (size_t)total, (void**)&buf))) return rc;
    if ((rc = arena_get(kSlotKeepMap, sizeof(int32_t) * ((size_t)in.nij + 1), (void**)&A.remap))) return rc;
    for (int k = 0; k < kKeepLists; ++k) {
        A.bcnt[k] = buf;
        A.boff[k] = buf + nblk[k] + 1;
        buf += 2 * (nblk[k] + 1);
    }
    // counts of all five lists in one launch (with the kept pairs' ranks in their
    // workgroups), their scans (totals to dcount), then every list written in one launch
    const int64_t nall = A.blk0[kKeepLists];
    if ((rc = arena_get(kSlotKeepFlags, (size_t)nflag + 1, (void**)&A.flag))) return rc;
    const int ph = prof_start(kProfKeep, s, 0.0);
    if (nall > 0) {
        hipLaunchKernelGGL(keep_count_kernel, dim3((unsigned)nall), dim3(kBlock), 0, s, A);
        ABD_CHECK_LAUNCH();
    }
    ScanSeg sg[kKeepLists];
    int nseg = 0;
    for (int k = 0; k < kKeepLists; ++k)
        if (n[k] > 0) sg[nseg++] = ScanSeg{A.bcnt[k], A.boff[k], nblk[k], A.boff[k] + nblk[k], dcount + k};
    if ((rc = scan_exclusive_i64_multi(s, sg, nseg))) return rc;
    if (nall > 0) {
        hipLaunchKernelGGL(keep_write_kernel, dim3((unsigned)nall), dim3(kBlock), 0, s, A, 0, kKeepLists);
        ABD_CHECK_LAUNCH();
    }
    prof_stop(ph, s);
    // an empty list's total is 0 (its scan segment is skipped)
    for (int k = 0; k < kKeepLists; ++k)
        if (n[k] == 0 && (rc = zero_fill(s, dcount + k, sizeof(int64_t)))) return rc;
    if (hcount) {
        ABD_CHECK(hipMemcpyAsync(hcount, dcount, kKeepLists * sizeof(int64_t), hipMemcpyDeviceToHost, s));
        if (sync) ABD_CHECK(hipStreamSynchronize(s));
    }
    return ABD_OK;
}

}  // namespace abd

using namespace abd;

extern "C" int abd_cull_keep(abd_stream_t stream, const AbdScene* sc, const int32_t* ij, const int32_t* ji,
                             int64_t nij, const int32_t* pt, int64_t npt, const int32_t* ee, int64_t nee,
                             const int32_t* vg, int64_t nvg, const float* blo, const float* bhi, const float* plo,
                             const float* phi, const float* elo, const float* ehi, const float* tlo,
                             const float* thi, double dhat, double ground, int32_t* ij_out, int32_t* ji_out,
                             int32_t* pt_out, int32_t* ee_out, int32_t* vg_out, int64_t* counts) {
    if (!sc || !counts || nij < 0 || npt < 0 || nee < 0 || nvg < 0) return ABD_ERR_ARG;
    if ((nij && (!ij || !ij_out || (ji && !ji_out))) || (npt && (!pt || !pt_out)) || (nee && (!ee || !ee_out)) ||
        (nvg && (!vg || !vg_out)))
        return ABD_ERR_ARG;
    if (!blo || !bhi || !plo || !phi || !elo || !ehi || !tlo || !thi) return ABD_ERR_ARG;
    hipStream_t s = to_stream(stream);
    int64_t *dcount, *hcount;
    int rc = arena_get(kSlotKeepCount, 8 * sizeof(int64_t), (void**)&dcount);
    if (!rc) rc = arena_host(8 * sizeof(int64_t), (void**)&hcount);
    if (rc) return rc;
    const KeepInput in{ij, ji, pt, ee, vg, nij, npt, nee, nvg};
    const KeepBoxes bx{blo, bhi, plo, phi, elo, ehi, tlo, thi};
    const KeepOutput out{ij_out, ji_out, pt_out, ee_out, vg_out};
    if ((rc = keep_from_superset(s, sc, in, bx, dhat, ground, out, dcount, hcount))) return rc;
    for (int k = 0; k < 5; ++k) counts[k] = hcount[k];
    return ABD_OK;
}

