// Optional per-kernel timing with HIP events recorded on the launch stream.
// Disabled by default; when enabled every launcher brackets its kernel with an
// event pair, and fhehip_profiler_report() sums the elapsed times per kernel.
#pragma once

#include <hip/hip_runtime.h>

namespace fhehip {

bool profiler_on();

// While a CaptureGuard lives on a thread, launches on that thread record no
// profiling events (they are being captured into a HIP graph).
struct CaptureGuard {
    CaptureGuard();
    ~CaptureGuard();
};
void profiler_record(const char* name, int logn, hipEvent_t start, hipEvent_t stop);
hipEvent_t profiler_event();

struct ProfScope {
    const char* name;
    int logn;
    hipStream_t stream;
    hipEvent_t start = nullptr;
    ProfScope(const char* n, int lg, hipStream_t s) : name(n), logn(lg), stream(s) {
        if (profiler_on()) {
            start = profiler_event();
            (void)hipEventRecord(start, stream);
        }
    }
    ~ProfScope() {
        if (start) {
            hipEvent_t stop = profiler_event();
            (void)hipEventRecord(stop, stream);
            profiler_record(name, logn, start, stop);
        }
    }
};

}  // namespace fhehip
