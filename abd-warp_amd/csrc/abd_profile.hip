// abd_profile.hip — see abd_profile.h.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "abd_internal.h"
#include "abd_profile.h"

namespace abd {

namespace {
struct Rec {
    int kernel;
    hipEvent_t a, b;
    hipEvent_t c;    // after the copies of scale / byte values (nullptr: none)
    double bytes;
    int scale_slot;  // index into the pinned scale buffer, -1 when unscaled
    int bytes_slot;  // index into the pinned byte buffer, -1 when unscaled
    bool stopped;    // prof_stop recorded the end event
};
constexpr int kScaleSlots = 1 << 16;
int32_t* g_scale = nullptr;  // pinned host copies of per-launch scale values
double* g_dbytes = nullptr;  // pinned host copies of per-launch device byte counts
int g_scale_used = 0;
int g_bytes_used = 0;
std::mutex g_mu;
bool g_on = false;
uint32_t g_mask = 0xffffffffu;  // kernel ids whose launches are timed
std::vector<Rec> g_recs;
// events are created once and recycled: creating two events per launch would cost
// more host time than many of the kernels being timed
std::vector<hipEvent_t> g_pool;

hipEvent_t take_event() {
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
double g_ms[kProfCount] = {};
double g_bytes[kProfCount] = {};
int64_t g_n[kProfCount] = {};

size_t g_done = 0;  // records before this index are resolved (their events recycled)

void account(const Rec& r, float ms) {
    double mult = r.scale_slot >= 0 ? (double)g_scale[r.scale_slot] : 1.0;
    if (r.bytes_slot >= 0) mult *= g_dbytes[r.bytes_slot];
    g_ms[r.kernel] += ms;
    g_bytes[r.kernel] += r.bytes * mult;
    g_n[r.kernel] += 1;
    g_pool.push_back(r.a);
    g_pool.push_back(r.b);
    if (r.c) g_pool.push_back(r.c);
}

// Resolves the records whose end event has completed, in order, without waiting.
void resolve_ready_locked() {
    while (g_done < g_recs.size()) {
        Rec& r = g_recs[g_done];
        if (!r.stopped || hipEventQuery(r.c ? r.c : r.b) != hipSuccess) break;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) break;
        account(r, ms);
        ++g_done;
    }
}

int resolve_locked() {
    for (; g_done < g_recs.size(); ++g_done) {
        Rec& r = g_recs[g_done];
        if (!r.stopped) continue;
        ABD_CHECK(hipEventSynchronize(r.c ? r.c : r.b));
        float ms = 0.f;
        ABD_CHECK(hipEventElapsedTime(&ms, r.a, r.b));
        account(r, ms);
    }
    g_recs.clear();
    g_done = 0;
    g_scale_used = 0;
    g_bytes_used = 0;
    return ABD_OK;
}
}  // namespace

__global__ void prof_bytes_kernel(const int64_t* __restrict__ dn, double c_pt, double c_ee, double c_vg,
                                  double c_nij, int single, double base, double* __restrict__ out) {
    out[0] = single ? base + c_pt * (double)dn[0]
                    : base + c_pt * (double)dn[2] + c_ee * (double)dn[3] + c_vg * (double)dn[4] +
                          c_nij * (double)dn[0];
}

namespace {
constexpr int kByteSlots = 1 << 14;  // device byte counts in flight (reused in stream order)
int g_byte_next = 0;
}  // namespace

int prof_byte_slot(int handle, double** out) {
    *out = nullptr;
    if (handle < 0) return ABD_OK;
    double* slots;
    int rc = arena_get(kSlotProfBytes, sizeof(double) * kByteSlots, (void**)&slots);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(g_mu);
    *out = slots + (g_byte_next++ % kByteSlots);
    return ABD_OK;
}

int prof_count_bytes(int handle, hipStream_t s, const int64_t* dn, double c_pt, double c_ee, double c_vg,
                     double c_nij, bool single, double base) {
    if (handle < 0 || !dn) return ABD_OK;
    double* out;
    int rc = prof_byte_slot(handle, &out);
    if (rc) return rc;
    hipLaunchKernelGGL(prof_bytes_kernel, dim3(1), dim3(1), 0, s, dn, c_pt, c_ee, c_vg, c_nij, single ? 1 : 0, base,
                       out);
    ABD_CHECK_LAUNCH();
    prof_stop(handle, s, nullptr, out);
    return ABD_OK;
}

bool prof_enabled() { return g_on; }

int prof_start(int kernel, hipStream_t s, double bytes) {
    if (!g_on || !((g_mask >> kernel) & 1u)) return -1;
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_pool.size() < 2 && g_recs.size() - g_done > 1024) resolve_ready_locked();
    Rec r{kernel, take_event(), take_event(), nullptr, bytes, -1, -1, false};
    if (!r.a || !r.b) {
        if (r.a) g_pool.push_back(r.a);
        if (r.b) g_pool.push_back(r.b);
        return -1;
    }
    if (hipEventRecord(r.a, s) != hipSuccess) {
        g_pool.push_back(r.a);
        g_pool.push_back(r.b);
        return -1;
    }
    g_recs.push_back(r);
    return (int)g_recs.size() - 1;
}

int prof_start_ext(int kernel, double bytes) {
    if (!g_on || !((g_mask >> kernel) & 1u)) return -1;
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_pool.size() < 2 && g_recs.size() - g_done > 1024) resolve_ready_locked();
    Rec r{kernel, take_event(), take_event(), nullptr, bytes, -1, -1, false};
    if (!r.a || !r.b) {
        if (r.a) g_pool.push_back(r.a);
        if (r.b) g_pool.push_back(r.b);
        return -1;
    }
    g_recs.push_back(r);
    return (int)g_recs.size() - 1;
}

void prof_ext_events(int handle, hipEvent_t* a, hipEvent_t* b) {
    std::lock_guard<std::mutex> lock(g_mu);
    *a = g_recs[handle].a;
    *b = g_recs[handle].b;
}

void prof_ext_done(int handle) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_recs[handle].stopped = true;
}

void prof_stop(int handle, hipStream_t s, const int32_t* dev_scale, const double* dev_bytes) {
    if (handle < 0) return;
    std::lock_guard<std::mutex> lock(g_mu);
    if (handle >= (int)g_recs.size()) return;
    Rec& r = g_recs[handle];
    // the timed region ends here, before any copy (a second call only adds values)
    if (!r.stopped && hipEventRecord(r.b, s) == hipSuccess) r.stopped = true;
    bool copied = false;
    if (dev_scale && g_scale && g_scale_used < kScaleSlots) {
        // the value as of this launch, copied on the launch stream
        const int slot = g_scale_used++;
        if (hipMemcpyAsync(&g_scale[slot], dev_scale, sizeof(int32_t), hipMemcpyDeviceToHost, s) == hipSuccess) {
            r.scale_slot = slot;
            copied = true;
        }
    }
    if (dev_bytes && g_dbytes && g_bytes_used < kScaleSlots) {
        const int slot = g_bytes_used++;
        if (hipMemcpyAsync(&g_dbytes[slot], dev_bytes, sizeof(double), hipMemcpyDeviceToHost, s) == hipSuccess) {
            r.bytes_slot = slot;
            copied = true;
        }
    }
    if (copied && r.stopped && !r.c) {
        r.c = take_event();
        if (r.c && hipEventRecord(r.c, s) != hipSuccess) {
            g_pool.push_back(r.c);
            r.c = nullptr;
        }
    }
}

}  // namespace abd

using namespace abd;

extern "C" {

int abd_profile_begin(void) {
    std::lock_guard<std::mutex> lock(g_mu);
    int rc = resolve_locked();
    for (int k = 0; k < kProfCount; ++k) {
        g_ms[k] = 0.0;
        g_bytes[k] = 0.0;
        g_n[k] = 0;
    }
    if (!g_scale && hipHostMalloc((void**)&g_scale, sizeof(int32_t) * kScaleSlots, hipHostMallocDefault) != hipSuccess)
        g_scale = nullptr;
    if (!g_dbytes && hipHostMalloc((void**)&g_dbytes, sizeof(double) * kScaleSlots, hipHostMallocDefault) != hipSuccess)
        g_dbytes = nullptr;
    // create the events a profiled region recycles before it starts
    while (g_pool.size() < 4096) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) break;
        g_pool.push_back(e);
    }
    g_on = true;
    return rc;
}

int abd_profile_select(uint32_t mask) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_mask = mask;
    return ABD_OK;
}

int abd_profile_end(void) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_on = false;
    return resolve_locked();
}

int abd_profile_query(int32_t kernel, double* ms, int64_t* launches, double* bytes) {
    if (kernel < 0 || kernel >= kProfCount) return ABD_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_mu);
    int rc = resolve_locked();
    if (rc) return rc;
    if (ms) *ms = g_ms[kernel];
    if (launches) *launches = g_n[kernel];
    if (bytes) *bytes = g_bytes[kernel];
    return ABD_OK;
}

}  // extern "C"
