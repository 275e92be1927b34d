// Per-degree launch shapes of the fused row kernels (see tuning.h).
#include "tuning.h"

#include <atomic>

namespace fhehip {

namespace {
constexpr int kDegrees = 16;  // log2 degrees 0 .. 15
constexpr int kKnobs = 7;
std::atomic<int> g_knob[kDegrees][kKnobs];
std::atomic<bool> g_set[kDegrees];
}  // namespace

FusedTuning fused_tuning(int logn) {
    FusedTuning t;
    if (logn < 0 || logn >= kDegrees || !g_set[logn].load(std::memory_order_acquire)) return t;
    t.xcd = g_knob[logn][0].load(std::memory_order_relaxed);
    t.ks_mac_inst = g_knob[logn][1].load(std::memory_order_relaxed);
    t.re_tensor = g_knob[logn][2].load(std::memory_order_relaxed);
    t.re_lift = g_knob[logn][3].load(std::memory_order_relaxed);
    t.re_finish = g_knob[logn][4].load(std::memory_order_relaxed);
    t.halves = g_knob[logn][5].load(std::memory_order_relaxed);
    t.scaler_mfma = g_knob[logn][6].load(std::memory_order_relaxed);
    return t;
}

void set_fused_tuning(int logn, const FusedTuning& t) {
    if (logn < 3 || logn >= kDegrees) throw Error(-1, "tuning: degree 8 .. 32768");
    auto re_ok = [](int re) { return re == 4 || re == 5; };
    if ((t.xcd != 0 && t.xcd != 1) || (t.ks_mac_inst != 2 && t.ks_mac_inst != 4 && t.ks_mac_inst != 8) ||
        !re_ok(t.re_tensor) || !re_ok(t.re_lift) || !re_ok(t.re_finish) || t.halves < 0 || t.halves > 7 || (t.scaler_mfma != 0 && t.scaler_mfma != 1))
        throw Error(-1, "tuning: xcd 0 or 1, ks_mac instances 2, 4 or 8, log2 values per thread 4 or 5, halves 0 .. 7, "
                        "scaler_mfma 0 or 1");
    g_knob[logn][0].store(t.xcd, std::memory_order_relaxed);
    g_knob[logn][1].store(t.ks_mac_inst, std::memory_order_relaxed);
    g_knob[logn][2].store(t.re_tensor, std::memory_order_relaxed);
    g_knob[logn][3].store(t.re_lift, std::memory_order_relaxed);
    g_knob[logn][4].store(t.re_finish, std::memory_order_relaxed);
    g_knob[logn][5].store(t.halves, std::memory_order_relaxed);
    g_knob[logn][6].store(t.scaler_mfma, std::memory_order_relaxed);
    g_set[logn].store(true, std::memory_order_release);
}

}  // namespace fhehip
