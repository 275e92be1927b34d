// Request server: continuous batching of encrypted evaluations on one GPU.
#pragma once

#include <atomic>
#include <condition_variable>
#include <deque>
#include <thread>

#include "bfv.h"

namespace fhehip {

struct ServerConfig {
    u64 hbm_budget = 0;  // device bytes admitted requests may occupy at once (0: no limit)
    u32 max_batch = 64;  // requests per batch
    u32 lanes = 4;       // batches in flight at once, one stream each
};

// One request kind: ct x ct multiply + relinearize, then the sum over
// 2^log2_slots consecutive slots by rotations and additions.
struct ServerParams {
    Multiplicator* mult;
    EvaluationKey* ek;
    std::vector<u32> exponents;
    size_t ct = 0;         // words per ciphertext
    u64 footprint = 0;     // device bytes one admitted request occupies
};

struct ServerStats {
    u64 submitted = 0, completed = 0, failed = 0, batches = 0;
    u64 peak_bytes = 0;
    u32 max_batch_seen = 0;
};

// Requests arrive through submit() from any thread; a server thread admits
// them in arrival order into batches of one parameter set while their
// footprint fits the budget, launches each batch on a free lane as soon as
// one is free (so requests that arrive while others run are batched and
// overlap on the device) and records completion times.
class Server {
  public:
    explicit Server(const ServerConfig& cfg);
    ~Server();
    u32 add_params(Multiplicator* m, EvaluationKey* ek, u32 log2_slots);
    const ServerParams& params(u32 id) const { return params_.at(id); }
    u64 submit(u32 params, const u64* lhs, const u64* rhs, u64* out);
    // state: 0 queued, 1 running, 2 done, -1 failed; latency in microseconds
    // from submit() to the observed completion (done requests only).
    void query(u64 ticket, int& state, double& latency_us);
    // Waits until every request submitted so far has finished; false on timeout.
    bool drain(double timeout_s);
    ServerStats stats();
    std::string error();

  private:
    struct Request {
        u32 params;
        const u64* lhs;
        const u64* rhs;
        u64* out;
        double t_submit = 0, t_done = 0;
        int state = 0;
    };
    struct Lane {
        hipStream_t st = nullptr;
        hipEvent_t done = nullptr;
        bool busy = false;
        std::vector<u64> reqs;
        u64 bytes = 0;
        u64* pinned = nullptr;
        size_t pinned_words = 0;
        DevBuf table;
    };
    void loop();
    bool launch(Lane& lane);  // under m_
    void finish(Lane& lane, bool ok);

    ServerConfig cfg_;
    int device_ = 0;
    std::mutex m_;
    std::condition_variable cv_;       // new arrivals, completions
    std::deque<ServerParams> params_;  // stable references while serving
    std::vector<Request> reqs_;
    std::deque<u64> queue_;            // arrival order
    std::vector<std::unique_ptr<Lane>> lanes_;
    u64 in_flight_ = 0;
    ServerStats stats_;
    std::string error_;
    bool stop_ = false;
    std::thread th_;
    StreamScratch in_a_, in_b_, cur_, res_;
};

}  // namespace fhehip
