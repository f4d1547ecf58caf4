// abd_block.h — workgroup-level building blocks (wave64 ballots, fixed-order
// reductions and scans through LDS).  Every helper ends with a barrier so the
// LDS it used can be reused immediately by the caller.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace abd {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }
__device__ __forceinline__ int wave_id() { return (int)(threadIdx.x / kWave); }
__device__ __forceinline__ int n_waves() { return (int)((blockDim.x + kWave - 1) / kWave); }

// Sum over the workgroup, fixed pairing (deterministic); result returned to every thread.
// sh must hold blockDim.x doubles.
__device__ __forceinline__ double block_sum(double v, double* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = (int)blockDim.x / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] = sh[t] + sh[t + s];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}
// Minimum over the workgroup (blockDim.x a multiple of 64; sh: n_waves() entries): a
// shuffle tree inside each wave, then the waves' minima (two barriers).  NaN never wins a
// comparison, so a NaN is kept only where every other value is NaN.
__device__ __forceinline__ double block_min(double v, double* sh) {
    double r = v;
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const double y = __shfl_xor(r, off, kWave);
        r = y < r ? y : r;
    }
    if (lane_id() == 0) sh[wave_id()] = r;
    __syncthreads();
    r = sh[0];
    for (int k = 1; k < n_waves(); ++k) r = sh[k] < r ? sh[k] : r;
    __syncthreads();
    return r;
}
__device__ __forceinline__ float block_minf(float v, float* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = (int)blockDim.x / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] = fminf(sh[t], sh[t + s]);
        __syncthreads();
    }
    float r = sh[0];
    __syncthreads();
    return r;
}
__device__ __forceinline__ float block_maxf(float v, float* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = (int)blockDim.x / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] = fmaxf(sh[t], sh[t + s]);
        __syncthreads();
    }
    float r = sh[0];
    __syncthreads();
    return r;
}

// Sum over the workgroup with two barriers: a fixed shuffle tree inside each
// wave (lane 0 ends with the wave's sum), wave sums through LDS, then the first
// wave reduces those in the same fixed order.  Every thread receives the value
// computed by one lane, so all waves branch alike on it.  blockDim.x must be a
// multiple of 64 and at most 1024; sh must hold 17 doubles.
__device__ __forceinline__ double wave_sum_down(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
    return v;
}
__device__ __forceinline__ double block_sum_fast(double v, double* sh) {
    const int lane = lane_id(), w = wave_id(), nw = n_waves();
    v = wave_sum_down(v);
    if (lane == 0) sh[w] = v;
    __syncthreads();
    if (w == 0) {
        double t = lane < nw ? sh[lane] : 0.0;
#pragma unroll
        for (int off = 8; off > 0; off >>= 1) t += __shfl_down(t, off, kWave);
        if (lane == 0) sh[16] = t;
    }
    __syncthreads();
    return sh[16];
}

// Two independent sums with the reduction pattern of block_sum_fast (each result
// bit-identical to a block_sum_fast of that value alone) and the same two
// barriers.  sh must hold 34 doubles.
__device__ __forceinline__ void block_sum2_fast(double& a, double& b, double* sh) {
    const int lane = lane_id(), w = wave_id(), nw = n_waves();
    a = wave_sum_down(a);
    b = wave_sum_down(b);
    if (lane == 0) {
        sh[w] = a;
        sh[17 + w] = b;
    }
    __syncthreads();
    if (w == 0) {
        double ta = lane < nw ? sh[lane] : 0.0;
        double tb = lane < nw ? sh[17 + lane] : 0.0;
#pragma unroll
        for (int off = 8; off > 0; off >>= 1) {
            ta += __shfl_down(ta, off, kWave);
            tb += __shfl_down(tb, off, kWave);
        }
        if (lane == 0) {
            sh[16] = ta;
            sh[33] = tb;
        }
    }
    __syncthreads();
    a = sh[16];
    b = sh[33];
}

// Exclusive prefix sum of one int64 per thread (sh: at least n_waves() entries): a
// shuffle scan inside each wave, then the totals of the waves before it (two barriers).
__device__ __forceinline__ int64_t block_exclusive_scan_i64(int64_t v, int64_t* sh) {
    const int lane = lane_id(), w = wave_id();
    int64_t incl = v;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const int64_t y = __shfl_up(incl, off, kWave);
        if (lane >= off) incl += y;
    }
    if (lane == kWave - 1) sh[w] = incl;
    __syncthreads();
    int64_t before = 0;
    for (int k = 0; k < w; ++k) before += sh[k];
    __syncthreads();  // sh is free again for the caller
    return before + incl - v;
}

// Ordered compaction rank inside the workgroup: returns the number of threads
// with a lower thread index whose predicate holds; *total = count over the
// workgroup.  sh_waves: n_waves() ints.
__device__ __forceinline__ int block_rank(bool pred, int* sh_waves, int* total) {
    unsigned long long mask = __ballot(pred ? 1 : 0);
    const int lane = lane_id();
    const int w = wave_id();
    unsigned long long below = lane == 0 ? 0ull : (mask & ((~0ull) >> (kWave - lane)));
    int r = __popcll(below);
    if (lane == 0) sh_waves[w] = __popcll(mask);
    __syncthreads();
    int off = 0, tot = 0;
    const int nw = n_waves();
    for (int k = 0; k < nw; ++k) {
        int c = sh_waves[k];
        if (k < w) off += c;
        tot += c;
    }
    __syncthreads();
    *total = tot;
    return off + r;
}

// Single workgroup (kScanThreads threads) exclusive scan of load(0), ..., load(n - 1) into out;
// each pass covers 1024 threads x 8 items.  total / total2 (either may be null) receive the sum.
constexpr int kScanThreads = 1024;
constexpr int kScanItems = 8;

template <class Load>
__device__ __forceinline__ void scan_block_exclusive_by(Load load, int64_t* __restrict__ out, int64_t n,
                                                        int64_t* __restrict__ total, int64_t* __restrict__ total2) {
    __shared__ int64_t sh[kScanThreads];
    int64_t carry = 0;
    const int t = threadIdx.x;
    for (int64_t base = 0; base < n; base += (int64_t)kScanThreads * kScanItems) {
        int64_t v[kScanItems];
        int64_t s = 0;
        int64_t i0 = base + (int64_t)t * kScanItems;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            int64_t i = i0 + k;
            v[k] = i < n ? load(i) : 0;
            s += v[k];
        }
        int64_t excl = block_exclusive_scan_i64(s, sh);
        int64_t run = carry + excl;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            int64_t i = i0 + k;
            if (i < n) out[i] = run;
            run += v[k];
        }
        // last thread holds the tile total
        if (t == kScanThreads - 1) sh[0] = excl + s;
        __syncthreads();
        carry += sh[0];
        __syncthreads();
    }
    if (t == 0 && total) *total = carry;
    if (t == 0 && total2) *total2 = carry;
}

}  // namespace abd
