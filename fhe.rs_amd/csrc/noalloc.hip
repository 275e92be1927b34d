// fhehip_multiply for callers that cannot allocate device memory: every
// scratch buffer of the multiplication is carved from one fixed device array
// per GPU instead of the per-stream pools.
#include <mutex>

#include "../../include/fhehip.h"
#include "bfv.h"
#include "profiler.h"

// The fixed scratch region. Each GPU the library runs on holds its own copy
// from the code object's load until the process ends. In host code this name
// is a stand-in, not a device address: host code only uses the address that
// hipGetSymbolAddress returns for the current device. (Inside namespace
// fhehip the unqualified name would be the scratch-pool function, so the
// lookup below names it ::scratch.)
static __device__ __align__(256) uint8_t scratch[FHEHIP_NOALLOC_SCRATCH_BYTES];

namespace fhehip {

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};
}  // namespace

// The current device's copy of `scratch`, looked up once per device (with
// that device current) and kept.
static void* device_scratch() {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error(-6, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch)));
        if (!p) throw Error(-6, "the fixed scratch region has no device address");
        g_scratch[dev] = p;
    }
    return g_scratch[dev];
}

void multiply_noalloc(Multiplicator& m, const u64* lhs, const u64* rhs, u64* out, u32 inst, hipStream_t s) {
    if (inst > 65535) throw Error(-1, "at most 65535 instances per call");
    const size_t need = m.scratch_bytes(inst);
    if (need > FHEHIP_NOALLOC_SCRATCH_BYTES)
        throw Error(-5, "this batch needs " + std::to_string(need) + " bytes of scratch; the fixed region holds " +
                            std::to_string((size_t)FHEHIP_NOALLOC_SCRATCH_BYTES) + " (" +
                            std::to_string(FHEHIP_NOALLOC_SCRATCH_BYTES / m.scratch_bytes(1)) +
                            " instances of this multiplicator)");
    if (!inst) return;
    FixedArena arena(device_scratch(), FHEHIP_NOALLOC_SCRATCH_BYTES);
    ArenaScope route(arena);
    // No per-kernel profiler events either: they are created on demand.
    CaptureGuard no_profiling;
    m.multiply(lhs, rhs, out, inst, s);
}

}  // namespace fhehip
