// C-ABI plumbing: errors, device memory, streams, timers.
#include "common.h"
#include "../../include/atomorph_hip.h"

#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace amh {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int check_hip(hipError_t e, const char *what) {
    if (e == hipSuccess) return AMH_OK;
    set_error("%s: %s", what, hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? AMH_E_NOMEM : AMH_E_HIP;
}

// ---------------------------------------------------- kernel profiling
namespace prof {
struct Pending {
    std::string name;
    hipEvent_t a, b;
};
static std::mutex mu;
static std::atomic<bool> on{false};
static std::vector<Pending> pending;
static std::vector<std::pair<hipEvent_t, hipEvent_t>> pool;
static std::map<std::string, std::pair<uint64_t, double>> totals;  // launches, ms
static thread_local std::vector<Pending> open;
}  // namespace prof

bool profiling() { return prof::on.load(std::memory_order_relaxed); }

void prof_begin(const char *kernel, hipStream_t s) {
    prof::Pending p;
    p.name = kernel;
    {
        std::lock_guard<std::mutex> lk(prof::mu);
        if (!prof::pool.empty()) {
            p.a = prof::pool.back().first;
            p.b = prof::pool.back().second;
            prof::pool.pop_back();
        } else if (hipEventCreate(&p.a) != hipSuccess || hipEventCreate(&p.b) != hipSuccess) {
            return;
        }
    }
    if (hipEventRecord(p.a, s) != hipSuccess) return;
    prof::open.push_back(p);
}

void prof_end(const char *kernel, hipStream_t s) {
    for (size_t i = prof::open.size(); i-- > 0;) {
        if (prof::open[i].name != kernel) continue;
        prof::Pending p = prof::open[i];
        prof::open.erase(prof::open.begin() + long(i));
        if (hipEventRecord(p.b, s) != hipSuccess) return;
        std::lock_guard<std::mutex> lk(prof::mu);
        prof::pending.push_back(p);
        return;
    }
}

bool prof_take(hipEvent_t &start, hipEvent_t &stop) {
    std::lock_guard<std::mutex> lk(prof::mu);
    if (!prof::pool.empty()) {
        start = prof::pool.back().first;
        stop = prof::pool.back().second;
        prof::pool.pop_back();
        return true;
    }
    return hipEventCreate(&start) == hipSuccess && hipEventCreate(&stop) == hipSuccess;
}

void prof_put(const char *kernel, hipEvent_t start, hipEvent_t stop) {
    prof::Pending p;
    p.name = kernel;
    p.a = start;
    p.b = stop;
    std::lock_guard<std::mutex> lk(prof::mu);
    prof::pending.push_back(p);
}

}  // namespace amh

using amh::check_hip;
using amh::set_error;

extern "C" int amh_profile_enable(int on) {
    amh::prof::on.store(on != 0);
    return AMH_OK;
}

// Device time of the launches recorded under `kernel` since the last reset
// (synchronous: waits for their events).
extern "C" int amh_profile_read(const char *kernel, uint64_t *launches, double *ms) {
    std::lock_guard<std::mutex> lk(amh::prof::mu);
    for (auto &p : amh::prof::pending) {
        float t = 0.0f;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&t, p.a, p.b) == hipSuccess) {
            auto &e = amh::prof::totals[p.name];
            e.first += 1;
            e.second += t;
        }
        amh::prof::pool.emplace_back(p.a, p.b);
    }
    amh::prof::pending.clear();
    auto it = amh::prof::totals.find(kernel ? kernel : "");
    if (launches) *launches = it == amh::prof::totals.end() ? 0 : it->second.first;
    if (ms) *ms = it == amh::prof::totals.end() ? 0.0 : it->second.second;
    return AMH_OK;
}

extern "C" int amh_profile_reset(void) {
    std::lock_guard<std::mutex> lk(amh::prof::mu);
    for (auto &p : amh::prof::pending) amh::prof::pool.emplace_back(p.a, p.b);
    amh::prof::pending.clear();
    amh::prof::totals.clear();
    return AMH_OK;
}

extern "C" {

const char *amh_version(void) { return "atomorph-mi355x 0.1 (gfx950)"; }
const char *amh_last_error(void) { return amh::g_err; }

int amh_device_info(int32_t *compute_units, uint64_t *hbm_bytes, char *name, size_t name_len) {
    hipDeviceProp_t p;
    int r = check_hip(hipGetDeviceProperties(&p, 0), "hipGetDeviceProperties");
    if (r) return r;
    if (compute_units) *compute_units = p.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = uint64_t(p.totalGlobalMem);
    if (name && name_len) {
        snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
    }
    return AMH_OK;
}

int amh_malloc(void **ptr, size_t bytes) {
    if (!ptr) { set_error("amh_malloc: ptr is NULL"); return AMH_E_ARG; }
    *ptr = nullptr;
    if (bytes == 0) return AMH_OK;
    return check_hip(hipMalloc(ptr, bytes), "hipMalloc");
}
int amh_free(void *ptr) { return ptr ? check_hip(hipFree(ptr), "hipFree") : AMH_OK; }

int amh_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream) {
    if (!bytes) return AMH_OK;
    return check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, static_cast<hipStream_t>(stream)),
                     "hipMemcpyAsync H2D");
}
int amh_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream) {
    if (!bytes) return AMH_OK;
    return check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, static_cast<hipStream_t>(stream)),
                     "hipMemcpyAsync D2H");
}
int amh_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream) {
    if (!bytes) return AMH_OK;
    return check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, static_cast<hipStream_t>(stream)),
                     "hipMemcpyAsync D2D");
}
int amh_memset(void *dst, int value, size_t bytes, void *stream) {
    if (!bytes) return AMH_OK;
    return check_hip(hipMemsetAsync(dst, value, bytes, static_cast<hipStream_t>(stream)), "hipMemsetAsync");
}
int amh_stream_create(void **stream) {
    if (!stream) { set_error("amh_stream_create: NULL"); return AMH_E_ARG; }
    hipStream_t s = nullptr;
    int r = check_hip(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    *stream = s;
    return r;
}
int amh_stream_destroy(void *stream) {
    return stream ? check_hip(hipStreamDestroy(static_cast<hipStream_t>(stream)), "hipStreamDestroy") : AMH_OK;
}
int amh_stream_sync(void *stream) {
    return check_hip(hipStreamSynchronize(static_cast<hipStream_t>(stream)), "hipStreamSynchronize");
}

struct amh_timer_s { hipEvent_t a, b; };

int amh_timer_begin(void **timer, void *stream) {
    if (!timer) { set_error("amh_timer_begin: NULL"); return AMH_E_ARG; }
    amh_timer_s *t = static_cast<amh_timer_s *>(*timer);
    if (!t) {
        t = new amh_timer_s;
        int r = check_hip(hipEventCreate(&t->a), "hipEventCreate");
        if (!r) r = check_hip(hipEventCreate(&t->b), "hipEventCreate");
        if (r) { delete t; return r; }
        *timer = t;
    }
    return check_hip(hipEventRecord(t->a, static_cast<hipStream_t>(stream)), "hipEventRecord");
}
int amh_timer_end(void *timer, void *stream, float *ms) {
    amh_timer_s *t = static_cast<amh_timer_s *>(timer);
    if (!t || !ms) { set_error("amh_timer_end: NULL"); return AMH_E_ARG; }
    int r = check_hip(hipEventRecord(t->b, static_cast<hipStream_t>(stream)), "hipEventRecord");
    if (!r) r = check_hip(hipEventSynchronize(t->b), "hipEventSynchronize");
    if (!r) r = check_hip(hipEventElapsedTime(ms, t->a, t->b), "hipEventElapsedTime");
    return r;
}
int amh_timer_free(void *timer) {
    amh_timer_s *t = static_cast<amh_timer_s *>(timer);
    if (!t) return AMH_OK;
    (void)hipEventDestroy(t->a);
    (void)hipEventDestroy(t->b);
    delete t;
    return AMH_OK;
}

}  // extern "C"
