// abd_common.hip — library housekeeping (version, errors, scratch arena) and the
// generic deterministic scan / reduction kernels the stages share.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <mutex>

#include "abd_block.h"
#include "abd_internal.h"

namespace abd {

namespace {
constexpr int kMaxDevices = 64;
Arena g_arena[kMaxDevices];
std::mutex g_arena_mutex;
thread_local Arena* t_arena = nullptr;  // a simulator's own arena during its calls
thread_local bool t_capturing = false;    // this thread is recording a HIP graph
std::atomic<uint64_t> g_alloc_gen{1};

int current_device(int* dev) {
    hipError_t e = hipGetDevice(dev);
    if (e != hipSuccess) return -static_cast<int>(e);
    if (*dev < 0 || *dev >= kMaxDevices) return ABD_ERR_NODEVICE;
    return ABD_OK;
}
}  // namespace

int hip_failure(hipError_t e, const char* file, int line, const char* expr) {
    static const bool report = std::getenv("ABD_DEBUG_ERRORS") != nullptr;
    if (report) std::fprintf(stderr, "abd: %s failed at %s:%d: %s\n", expr, file, line, hipGetErrorString(e));
    return -static_cast<int>(e);
}

ArenaScope::ArenaScope(Arena* a) : prev_(t_arena) { t_arena = a; }
ArenaScope::~ArenaScope() { t_arena = prev_; }

uint64_t alloc_generation() { return g_alloc_gen.load(std::memory_order_relaxed); }
void note_reallocation() { g_alloc_gen.fetch_add(1, std::memory_order_relaxed); }
void set_capturing(bool on) { t_capturing = on; }
bool capturing() { return t_capturing; }

int arena_get(int slot, size_t bytes, void** ptr) {
    int dev;
    int rc = current_device(&dev);
    if (rc) return rc;
    std::unique_lock<std::mutex> lock(g_arena_mutex, std::defer_lock);
    if (!t_arena) lock.lock();
    Arena& A = t_arena ? *t_arena : g_arena[dev];
    if (bytes == 0) bytes = 256;
    if (A.size[slot] < bytes) {
        if (t_capturing) return ABD_ERR_CAPACITY;  // no allocation inside a recorded graph
        note_reallocation();
        size_t want = bytes + bytes / 2;
        want = (want + 255) & ~size_t(255);
        if (A.ptr[slot]) {
            ABD_CHECK(hipDeviceSynchronize());
            ABD_CHECK(hipFree(A.ptr[slot]));
            A.ptr[slot] = nullptr;
            A.size[slot] = 0;
        }
        const hipError_t e = hipMalloc(&A.ptr[slot], want);
        if (e != hipSuccess) {
            A.ptr[slot] = nullptr;
            if (std::getenv("ABD_DEBUG_ERRORS"))
                std::fprintf(stderr, "abd: scratch slot %d: %zu bytes not allocated\n", slot, want);
            return hip_failure(e, __FILE__, __LINE__, "hipMalloc");
        }
        A.size[slot] = want;
    }
    *ptr = A.ptr[slot];
    return ABD_OK;
}

int arena_host(size_t bytes, void** ptr) {
    int dev;
    int rc = current_device(&dev);
    if (rc) return rc;
    std::unique_lock<std::mutex> lock(g_arena_mutex, std::defer_lock);
    if (!t_arena) lock.lock();
    Arena& A = t_arena ? *t_arena : g_arena[dev];
    if (A.host_size < bytes) {
        if (t_capturing) return ABD_ERR_CAPACITY;
        note_reallocation();
        if (A.host) {
            ABD_CHECK(hipDeviceSynchronize());
            ABD_CHECK(hipHostFree(A.host));
            A.host = nullptr;
        }
        size_t want = bytes < 4096 ? 4096 : bytes;
        ABD_CHECK(hipHostMalloc(&A.host, want, hipHostMallocDefault));
        A.host_size = want;
    }
    *ptr = A.host;
    return ABD_OK;
}

int arena_release_current() {
    int dev;
    int rc = current_device(&dev);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(g_arena_mutex);
    return arena_free(g_arena[dev]);
}

int arena_free(Arena& A) {
    ABD_CHECK(hipDeviceSynchronize());
    for (int i = 0; i < kArenaSlots; ++i) {
        if (A.ptr[i]) ABD_CHECK(hipFree(A.ptr[i]));
        A.ptr[i] = nullptr;
        A.size[i] = 0;
    }
    if (A.host) ABD_CHECK(hipHostFree(A.host));
    A.host = nullptr;
    A.host_size = 0;
    return ABD_OK;
}

// ------------------------------------------------------------------ fill
__global__ __launch_bounds__(kBlock) void zero_words_kernel(uint64_t* __restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) p[i] = 0;
}

int zero_fill(hipStream_t s, void* p, size_t bytes) {
    if (bytes % 8 || ((uintptr_t)p & 7)) return ABD_ERR_ARG;
    const int64_t n = (int64_t)(bytes / 8);
    if (n == 0) return ABD_OK;
    hipLaunchKernelGGL(zero_words_kernel, dim3(grid_for(n, kBlock, 1 << 16)), dim3(kBlock), 0, s, (uint64_t*)p, n);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

// ------------------------------------------------------------------ scan
__device__ __forceinline__ void scan_block_exclusive(const int64_t* __restrict__ in, int64_t* __restrict__ out,
                                                     int64_t n, int64_t* __restrict__ total,
                                                     int64_t* __restrict__ total2) {
    scan_block_exclusive_by([in](int64_t i) { return in[i]; }, out, n, total, total2);
}

__global__ __launch_bounds__(kScanThreads) void scan_i64_kernel(const int64_t* __restrict__ in,
                                                                 int64_t* __restrict__ out, int64_t n,
                                                                 int64_t* __restrict__ total,
                                                                 int64_t* __restrict__ total2) {
    scan_block_exclusive(in, out, n, total, total2);
}

struct ScanSegs {
    ScanSeg s[kMaxScanSegs];
};
// one workgroup per segment
__global__ __launch_bounds__(kScanThreads) void scan_i64_multi_kernel(ScanSegs S) {
    const ScanSeg& g = S.s[blockIdx.x];
    scan_block_exclusive(g.in, g.out, g.n, g.total, g.total2);
}

int scan_exclusive_i64_multi(hipStream_t s, const ScanSeg* segs, int nseg) {
    if (nseg < 0 || nseg > kMaxScanSegs) return ABD_ERR_ARG;
    if (nseg == 0) return ABD_OK;
    ScanSegs S;
    for (int k = 0; k < nseg; ++k) S.s[k] = segs[k];
    hipLaunchKernelGGL(scan_i64_multi_kernel, dim3(nseg), dim3(kScanThreads), 0, s, S);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

int scan_exclusive_i64(hipStream_t s, const int64_t* in, int64_t* out, int64_t n, int64_t* total, int64_t* total2) {
    hipLaunchKernelGGL(scan_i64_kernel, dim3(1), dim3(kScanThreads), 0, s, in, out, n, total, total2);
    ABD_CHECK_LAUNCH();
    return ABD_OK;
}

namespace {
AbdTuning g_tuning{4096, 16384, 0, 0, 0};
}
AbdTuning tuning() { return g_tuning; }

int read_i64(hipStream_t s, const int64_t* dev, int64_t* host) {
    int64_t* h;
    int rc = arena_host(sizeof(int64_t), (void**)&h);
    if (rc) return rc;
    ABD_CHECK(hipMemcpyAsync(h, dev, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    ABD_CHECK(hipStreamSynchronize(s));
    *host = *h;
    return ABD_OK;
}

}  // namespace abd

using namespace abd;

extern "C" {

int abd_version(void) { return 10000; }

const char* abd_error_string(int code) {
    switch (code) {
        case ABD_OK: return "ok";
        case ABD_ERR_OVERFLOW: return "output list too small (required count returned)";
        case ABD_ERR_ARG: return "invalid argument";
        case ABD_ERR_UNSORTED: return "candidate list not grouped by body pair / body";
        case ABD_ERR_NODEVICE: return "no HIP device";
        case ABD_ERR_CAPACITY: return "input exceeds the fixed scratch of an allocation-free entry point";
        default: break;
    }
    if (code < 0) return hipGetErrorString(static_cast<hipError_t>(-code));
    return "unknown status";
}

int abd_release(void) { return arena_release_current(); }

int abd_get_tuning(AbdTuning* t) {
    if (!t) return ABD_ERR_ARG;
    *t = g_tuning;
    return ABD_OK;
}

int abd_set_tuning(const AbdTuning* t) {
    if (!t || t->cull_chunk < 256 || t->cull_chunk > 65536 || (t->cull_chunk & (t->cull_chunk - 1)) ||
        t->cull_grid < 64 || t->cull_grid > (1 << 20) || t->contact_bound_mib < 0 ||
        (t->pair_grid != 0 && (t->pair_grid < 64 || t->pair_grid > 65536)) ||
        (t->toi_grid != 0 && (t->toi_grid < 64 || t->toi_grid > 65536)))
        return ABD_ERR_ARG;
    g_tuning = *t;
    note_reallocation();  // graphs recorded under the old shapes are not replayed
    return ABD_OK;
}

}  // extern "C"
