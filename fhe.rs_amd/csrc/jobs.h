// Batched evaluation of jobs with differing parameter sets.
#pragma once

#include "bfv.h"

namespace fhehip {

constexpr u32 kJobMaxRotations = 4;

// One evaluation: multiply + relinearize, then column rotations in order.
struct Job {
    Multiplicator* mult;
    EvaluationKey* ek;
    const u64* lhs;
    const u64* rhs;
    u64* out;
    u32 num_rotations;
    u32 steps[kJobMaxRotations];
};

// Groups jobs by Multiplicator (one parameter set and relinearization key per
// group), runs each group as batched launches on a library-owned stream and
// orders the caller's stream after all of them. Calls are serialized on the
// host; two consecutive submits may be in flight on the device at once.
class JobRunner {
  public:
    static JobRunner& instance();
    void submit(const Job* jobs, u32 n, hipStream_t s);

  private:
    JobRunner() {}
    hipStream_t stream(u32 k);

    // Per-submit resources, alternated so that a submit can be prepared
    // while the previous one still runs.
    struct Slot {
        std::vector<hipEvent_t> events;
        u32 used = 0;
        hipEvent_t done = nullptr;  // recorded on the caller stream at the end
        u64* pinned = nullptr;      // host staging of the pointer table
        size_t pinned_words = 0;
        DevBuf table;               // device pointer table
        hipEvent_t next_event();
    };

    std::mutex m_;
    std::vector<hipStream_t> streams_;
    Slot slots_[2];
    u32 turn_ = 0;
    StreamScratch in_a_, in_b_, cur_, tmp_, sub_, sub2_;
};

}  // namespace fhehip
