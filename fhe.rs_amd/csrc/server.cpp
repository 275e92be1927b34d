// Request server: arrival queue, admission by device footprint, batches of
// one parameter set launched on free lanes, completion polling.
#include "server.h"

#include <algorithm>
#include <chrono>
#include <cstring>

namespace fhehip {

namespace {
double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}
}  // namespace

Server::Server(const ServerConfig& cfg) : cfg_(cfg) {
    if (cfg_.max_batch == 0 || cfg_.max_batch > 65535 || cfg_.lanes == 0 || cfg_.lanes > 16)
        throw Error(-1, "server needs 1..65535 requests per batch and 1..16 lanes");
    HIPCHK(hipGetDevice(&device_));
    for (u32 k = 0; k < cfg_.lanes; k++) {
        lanes_.emplace_back(new Lane);
        Lane& l = *lanes_.back();
        HIPCHK(hipStreamCreateWithFlags(&l.st, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&l.done, hipEventDisableTiming));
    }
    th_ = std::thread([this] { loop(); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> g(m_);
        stop_ = true;
    }
    cv_.notify_all();
    th_.join();
    for (auto& lp : lanes_) {
        Lane& l = *lp;
        // teardown: failures here have no caller to report to
        (void)hipStreamSynchronize(l.st);
        (void)hipEventDestroy(l.done);
        (void)hipStreamDestroy(l.st);
        if (l.pinned) (void)hipHostFree(l.pinned);
    }
}

u32 Server::add_params(Multiplicator* m, EvaluationKey* ek, u32 log2_slots) {
    if (!m || !ek) throw Error(-1, "null argument");
    if (!m->rk) throw Error(-1, "the multiplicator has no relinearization key");
    if (ek->keys.empty()) throw Error(-1, "empty evaluation key");
    const DeviceContext& b = *m->base;
    const auto& ec = *ek->keys[0]->ksk->ct_ctx;
    if (ec.N != b.N || ec.moduli != b.moduli)
        throw Error(-4, "evaluation key and multiplicator use different ciphertext contexts");
    if (((size_t)1 << log2_slots) > b.N / 2) throw Error(-1, "log2_slots too large");
    ServerParams p;
    p.mult = m;
    p.ek = ek;
    for (u32 k = 0; k < log2_slots; k++) {
        const u32 e = ek->column_exponent(1u << k);
        if (!ek->find(e)) throw Error(-1, "the evaluation key lacks a rotation of the slot sum");
        p.exponents.push_back(e);
    }
    const size_t N = b.N, L = b.L, M = m->mul->L, Lk = m->rk->key_ctx->L;
    size_t Lg = 0;
    for (auto& g : ek->keys) Lg = std::max<size_t>(Lg, g->ksk->key_ctx->L);
    const size_t common = std::min(m->ext_lhs->common, m->ext_rhs->common);
    p.ct = 2 * L * N;
    const size_t words = 4 * p.ct                                          // lane inputs, product, result
                         + (4 * L + 4 * (M - common) + 3 * M + 3 * L) * N   // multiplication scratch
                         + (L * Lk + 2 * (Lk + L)) * N                      // relinearization key switch
                         + L * N + (L * Lg + 2 * (Lg + L)) * N + p.ct;      // rotations and sums
    p.footprint = words * 8;
    std::lock_guard<std::mutex> g(m_);
    params_.push_back(p);
    return (u32)params_.size() - 1;
}

u64 Server::submit(u32 params, const u64* lhs, const u64* rhs, u64* out) {
    if (!lhs || !rhs || !out) throw Error(-1, "null argument");
    std::lock_guard<std::mutex> g(m_);
    if (params >= params_.size()) throw Error(-1, "unknown parameter set");
    if (cfg_.hbm_budget && params_[params].footprint > cfg_.hbm_budget)
        throw Error(-5, "one request needs more device memory than the server budget");
    if (stop_) throw Error(-1, "server is stopping");
    Request r;
    r.params = params;
    r.lhs = lhs;
    r.rhs = rhs;
    r.out = out;
    r.t_submit = now_s();
    reqs_.push_back(r);
    queue_.push_back(reqs_.size() - 1);
    stats_.submitted++;
    cv_.notify_all();
    return reqs_.size() - 1;
}

void Server::query(u64 ticket, int& state, double& latency_us) {
    std::lock_guard<std::mutex> g(m_);
    if (ticket >= reqs_.size()) throw Error(-1, "unknown ticket");
    const Request& r = reqs_[ticket];
    state = r.state;
    latency_us = r.state == 2 ? (r.t_done - r.t_submit) * 1e6 : 0.0;
}

bool Server::drain(double timeout_s) {
    std::unique_lock<std::mutex> lk(m_);
    const auto until = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeout_s);
    return cv_.wait_until(lk, until, [&] { return stats_.completed + stats_.failed == reqs_.size(); });
}

ServerStats Server::stats() {
    std::lock_guard<std::mutex> g(m_);
    return stats_;
}

std::string Server::error() {
    std::lock_guard<std::mutex> g(m_);
    return error_;
}

void Server::finish(Lane& lane, bool ok) {
    const double t = now_s();
    for (u64 id : lane.reqs) {
        reqs_[id].state = ok ? 2 : -1;
        reqs_[id].t_done = t;
        if (ok)
            stats_.completed++;
        else
            stats_.failed++;
    }
    in_flight_ -= lane.bytes;
    lane.bytes = 0;
    lane.reqs.clear();
    lane.busy = false;
    cv_.notify_all();
}

// Forms a batch for `lane` from the queue (called with m_ held; releases it
// while issuing device work). Returns false when nothing could be admitted.
bool Server::launch(Lane& lane) {
    if (queue_.empty()) return false;
    const u32 p = reqs_[queue_.front()].params;
    const ServerParams& P = params_[p];
    std::vector<u64> batch;
    u64 bytes = 0;
    for (auto it = queue_.begin(); it != queue_.end() && batch.size() < cfg_.max_batch;) {
        if (reqs_[*it].params != p) {
            ++it;
            continue;
        }
        if (cfg_.hbm_budget && in_flight_ + bytes + P.footprint > cfg_.hbm_budget) break;
        batch.push_back(*it);
        bytes += P.footprint;
        it = queue_.erase(it);
    }
    if (batch.empty()) return false;
    in_flight_ += bytes;
    stats_.peak_bytes = std::max(stats_.peak_bytes, in_flight_);
    stats_.batches++;
    stats_.max_batch_seen = std::max<u32>(stats_.max_batch_seen, (u32)batch.size());
    for (u64 id : batch) reqs_[id].state = 1;
    lane.busy = true;
    lane.reqs = batch;
    lane.bytes = bytes;
    // pointer table: inputs into the lane's contiguous buffers, results out
    const size_t B = batch.size(), ct = P.ct;
    std::vector<u64> tab;
    tab.reserve(6 * B);
    std::vector<const u64*> lhs(B), rhs(B);
    std::vector<u64*> out(B);
    for (size_t i = 0; i < B; i++) {
        lhs[i] = reqs_[batch[i]].lhs;
        rhs[i] = reqs_[batch[i]].rhs;
        out[i] = reqs_[batch[i]].out;
    }
    m_.unlock();
    try {
        hipStream_t st = lane.st;
        u64* a = in_a_.get(st, B * ct * 8).as<u64>();
        u64* b = in_b_.get(st, B * ct * 8).as<u64>();
        u64* cur = cur_.get(st, B * ct * 8).as<u64>();
        u64* res = res_.get(st, B * ct * 8).as<u64>();
        for (size_t i = 0; i < B; i++) tab.insert(tab.end(), {(u64)lhs[i], (u64)(a + i * ct)});
        for (size_t i = 0; i < B; i++) tab.insert(tab.end(), {(u64)rhs[i], (u64)(b + i * ct)});
        for (size_t i = 0; i < B; i++) tab.insert(tab.end(), {(u64)(res + i * ct), (u64)out[i]});
        if (lane.pinned_words < tab.size()) {
            if (lane.pinned) HIPCHK(hipHostFree(lane.pinned));
            lane.pinned = nullptr;
            HIPCHK(hipHostMalloc((void**)&lane.pinned, tab.size() * 8, hipHostMallocDefault));
            lane.pinned_words = tab.size();
        }
        if (lane.table.bytes < tab.size() * 8) lane.table.reset(tab.size() * 8);
        std::memcpy(lane.pinned, tab.data(), tab.size() * 8);
        HIPCHK(hipMemcpyAsync(lane.table.ptr, lane.pinned, tab.size() * 8, hipMemcpyHostToDevice, st));
        const u64* const* t = lane.table.as<const u64*>();
        launch_copy_slots(t, (u32)B, ct, st);
        launch_copy_slots(t + 2 * B, (u32)B, ct, st);
        P.mult->multiply(a, b, cur, (u32)B, st);
        P.ek->rotate_and_sum(P.exponents, cur, res, (u32)B, st);
        launch_copy_slots(t + 4 * B, (u32)B, ct, st);
        HIPCHK(hipEventRecord(lane.done, st));
    } catch (const std::exception& e) {
        m_.lock();
        error_ = e.what();
        (void)hipStreamSynchronize(lane.st);
        finish(lane, false);
        return true;
    }
    m_.lock();
    return true;
}

void Server::loop() {
    if (hipSetDevice(device_) != hipSuccess) {
        std::lock_guard<std::mutex> g(m_);
        error_ = "server thread could not select the device";
        return;
    }
    std::unique_lock<std::mutex> lk(m_);
    for (;;) {
        bool progress = false;
        bool busy = false;
        for (auto& lp : lanes_) {
            Lane& l = *lp;
            if (!l.busy) continue;
            const hipError_t q = hipEventQuery(l.done);
            if (q == hipSuccess) {
                finish(l, true);
                progress = true;
            } else if (q != hipErrorNotReady) {
                error_ = hipGetErrorString(q);
                finish(l, false);
                progress = true;
            } else {
                busy = true;
            }
        }
        if (stop_) {
            // refuse what is still queued; wait for running batches
            for (u64 id : queue_) {
                reqs_[id].state = -1;
                stats_.failed++;
            }
            queue_.clear();
            if (!busy) break;
        }
        for (auto& lp : lanes_) {
            Lane& l = *lp;
            if (l.busy || queue_.empty()) continue;
            if (launch(l)) {
                progress = true;
                busy = true;
            }
        }
        if (!progress) {
            if (busy)
                cv_.wait_for(lk, std::chrono::microseconds(20));
            else
                cv_.wait_for(lk, std::chrono::milliseconds(5), [&] { return stop_ || !queue_.empty(); });
        }
    }
    cv_.notify_all();
}

}  // namespace fhehip
