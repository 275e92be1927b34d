#include "profiler.h"

#include <atomic>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fhehip.h"
#include "capi_internal.h"

namespace fhehip {

namespace {
std::atomic<bool> g_on{false};
std::mutex g_mu;
struct Pending {
    std::string name;
    hipEvent_t start, stop;
};
std::vector<Pending> g_pending;
std::vector<hipEvent_t> g_free;
struct Stat {
    unsigned long calls = 0;
    double ms = 0;
};
std::map<std::string, Stat> g_stats;

void drain_locked() {
    for (auto& p : g_pending) {
        float ms = 0;
        if (hipEventSynchronize(p.stop) == hipSuccess && hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
            Stat& s = g_stats[p.name];
            s.calls++;
            s.ms += ms;
        }
        g_free.push_back(p.start);
        g_free.push_back(p.stop);
    }
    g_pending.clear();
}
}  // namespace

thread_local int t_capturing = 0;

CaptureGuard::CaptureGuard() { t_capturing++; }
CaptureGuard::~CaptureGuard() { t_capturing--; }

bool profiler_on() { return !t_capturing && g_on.load(std::memory_order_relaxed); }

hipEvent_t profiler_event() {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_free.empty()) {
        hipEvent_t e = g_free.back();
        g_free.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void profiler_record(const char* name, int logn, hipEvent_t start, hipEvent_t stop) {
    if (!start || !stop) return;
    std::string key = std::string(name) + "/N=" + std::to_string(1u << logn);
    std::lock_guard<std::mutex> lk(g_mu);
    g_pending.push_back({key, start, stop});
    if (g_pending.size() > 4096) drain_locked();
}

}  // namespace fhehip

using namespace fhehip;

extern "C" {

int fhehip_profiler_enable(int on) {
    FHEHIP_API_BEGIN
    std::lock_guard<std::mutex> lk(g_mu);
    drain_locked();
    g_stats.clear();
    g_on.store(on != 0);
    FHEHIP_API_END
}

int fhehip_profiler_report(char* buf, size_t cap, size_t* needed) {
    FHEHIP_API_BEGIN
    std::lock_guard<std::mutex> lk(g_mu);
    drain_locked();
    std::string out;
    char line[512];
    for (auto& kv : g_stats) {
        std::snprintf(line, sizeof line, "%s %lu %.6f\n", kv.first.c_str(), kv.second.calls, kv.second.ms);
        out += line;
    }
    if (needed) *needed = out.size() + 1;
    if (buf && cap) {
        size_t n = out.size() < cap - 1 ? out.size() : cap - 1;
        out.copy(buf, n);
        buf[n] = 0;
    }
    FHEHIP_API_END
}

}  // extern "C"
