// Batched evaluation of jobs with differing parameter sets: grouping by
// Multiplicator, per-group staging through pointer tables, per-group streams.
#include "jobs.h"

#include <algorithm>
#include <array>
#include <cstring>
#include <map>
#include <unordered_map>

namespace fhehip {

namespace {
constexpr u32 kStreams = 8;

// Device work of one group, planned on the host before anything is issued.
struct Op {
    enum Kind { COPY, MUL, ROT } kind;
    size_t table_off = 0;  // COPY: first pair in the table
    u32 count = 0;         // COPY: pairs; MUL/ROT: instances
    size_t elems = 0;      // COPY: words per slot
    Multiplicator* mult = nullptr;
    EvaluationKey* ek = nullptr;
    u32 exponent = 0;
    const u64* a = nullptr;
    const u64* b = nullptr;
    u64* out = nullptr;
};

struct Group {
    Multiplicator* mult;
    std::vector<u32> members;
    hipStream_t st = nullptr;
    size_t ct = 0;  // words per ciphertext
    std::vector<Op> ops;
};

bool contiguous(const std::vector<const u64*>& p, size_t stride) {
    for (size_t k = 1; k < p.size(); k++)
        if (p[k] != p[0] + k * stride) return false;
    return true;
}
}  // namespace

JobRunner& JobRunner::instance() {
    // Never destroyed: its streams and events live until the process ends.
    static JobRunner* r = new JobRunner;
    return *r;
}

hipStream_t JobRunner::stream(u32 k) {
    while (streams_.size() <= k) {
        hipStream_t st;
        HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        streams_.push_back(st);
    }
    return streams_[k];
}

hipEvent_t JobRunner::Slot::next_event() {
    if (used == events.size()) {
        hipEvent_t e;
        HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        events.push_back(e);
    }
    return events[used++];
}

void JobRunner::submit(const Job* jobs, u32 n, hipStream_t s) {
    std::lock_guard<std::mutex> lock(m_);
    if (!n) return;
    // ---- validation (nothing is issued before every job is checked)
    for (u32 k = 0; k < n; k++) {
        const Job& j = jobs[k];
        if (!j.mult || !j.lhs || !j.rhs || !j.out) throw Error(-1, "null job argument");
        if (!j.mult->rk) throw Error(-1, "job multiplicator has no relinearization key");
        if (j.num_rotations > kJobMaxRotations) throw Error(-1, "too many rotations in a job");
        if (!j.num_rotations) continue;
        if (!j.ek || j.ek->keys.empty()) throw Error(-1, "job with rotations but no evaluation key");
        const auto& ec = *j.ek->keys[0]->ksk->ct_ctx;
        if (ec.N != j.mult->base->N || ec.moduli != j.mult->base->moduli)
            throw Error(-4, "evaluation key and multiplicator use different ciphertext contexts");
        for (u32 r = 0; r < j.num_rotations; r++) {
            if (j.steps[r] == 0 || j.steps[r] >= ec.N / 2) throw Error(-1, "rotation step out of range");
            if (!j.ek->find(j.ek->column_exponent(j.steps[r])))
                throw Error(-1, "the evaluation key has no Galois key for a requested step");
        }
    }
    // ---- groups, in order of first appearance
    std::vector<Group> groups;
    std::unordered_map<Multiplicator*, size_t> gid;
    for (u32 k = 0; k < n; k++) {
        auto it = gid.find(jobs[k].mult);
        if (it == gid.end()) {
            it = gid.emplace(jobs[k].mult, groups.size()).first;
            groups.push_back(Group{jobs[k].mult, {}, nullptr, 0, {}});
        }
        groups[it->second].members.push_back(k);
    }
    // ---- per-stream scratch, sized for the largest group on each stream
    std::map<hipStream_t, size_t> need;
    for (size_t g = 0; g < groups.size(); g++) {
        Group& G = groups[g];
        G.st = stream((u32)(g % kStreams));
        G.ct = 2 * (size_t)G.mult->base->L * G.mult->base->N;
        size_t& b = need[G.st];
        b = std::max(b, G.members.size() * G.ct * 8);
    }
    std::map<hipStream_t, std::array<u64*, 6>> buf;
    for (auto& kv : need) {
        hipStream_t st = kv.first;
        buf[st] = {in_a_.get(st, kv.second).as<u64>(), in_b_.get(st, kv.second).as<u64>(),
                   cur_.get(st, kv.second).as<u64>(), tmp_.get(st, kv.second).as<u64>(),
                   sub_.get(st, kv.second).as<u64>(), sub2_.get(st, kv.second).as<u64>()};
    }
    // ---- plan: pointer table and per-group operations
    std::vector<u64> table;
    auto copy = [&](Group& G, const std::vector<std::pair<const u64*, u64*>>& pairs) {
        if (pairs.empty()) return;
        Op op;
        op.kind = Op::COPY;
        op.table_off = table.size() / 2;
        op.count = (u32)pairs.size();
        op.elems = G.ct;
        for (auto& pr : pairs) {
            table.push_back((u64)pr.first);
            table.push_back((u64)pr.second);
        }
        G.ops.push_back(op);
    };
    for (Group& G : groups) {
        const auto& B = buf[G.st];
        const size_t m = G.members.size(), ct = G.ct;
        std::vector<const u64*> lhs(m), rhs(m);
        for (size_t k = 0; k < m; k++) {
            lhs[k] = jobs[G.members[k]].lhs;
            rhs[k] = jobs[G.members[k]].rhs;
        }
        const u64* a = lhs[0];
        const u64* b = rhs[0];
        if (!contiguous(lhs, ct)) {
            std::vector<std::pair<const u64*, u64*>> pr;
            for (size_t k = 0; k < m; k++) pr.push_back({lhs[k], B[0] + k * ct});
            copy(G, pr);
            a = B[0];
        }
        if (!contiguous(rhs, ct)) {
            std::vector<std::pair<const u64*, u64*>> pr;
            for (size_t k = 0; k < m; k++) pr.push_back({rhs[k], B[1] + k * ct});
            copy(G, pr);
            b = B[1];
        }
        u64* cur = B[2];
        u64* tmp = B[3];
        Op mul;
        mul.kind = Op::MUL;
        mul.mult = G.mult;
        mul.a = a;
        mul.b = b;
        mul.out = cur;
        mul.count = (u32)m;
        G.ops.push_back(mul);
        for (u32 r = 0; r < kJobMaxRotations; r++) {
            // members still rotating in round r, by (evaluation key, step)
            std::map<std::pair<EvaluationKey*, u32>, std::vector<u32>> by;
            for (size_t k = 0; k < m; k++) {
                const Job& j = jobs[G.members[k]];
                if (j.num_rotations > r) by[{j.ek, j.steps[r]}].push_back((u32)k);
            }
            for (auto& kv : by) {
                EvaluationKey* ek = kv.first.first;
                const u32 exponent = ek->column_exponent(kv.first.second);
                const auto& pos = kv.second;
                Op rot;
                rot.kind = Op::ROT;
                rot.ek = ek;
                rot.exponent = exponent;
                rot.count = (u32)pos.size();
                if (pos.size() == m) {
                    rot.a = cur;
                    rot.out = tmp;
                    G.ops.push_back(rot);
                    std::swap(cur, tmp);
                } else {
                    std::vector<std::pair<const u64*, u64*>> in, back;
                    for (size_t i = 0; i < pos.size(); i++) {
                        in.push_back({cur + pos[i] * ct, B[4] + i * ct});
                        back.push_back({B[5] + i * ct, cur + pos[i] * ct});
                    }
                    copy(G, in);
                    rot.a = B[4];
                    rot.out = B[5];
                    G.ops.push_back(rot);
                    copy(G, back);
                }
            }
        }
        std::vector<std::pair<const u64*, u64*>> outp;
        for (size_t k = 0; k < m; k++) outp.push_back({cur + k * ct, jobs[G.members[k]].out});
        copy(G, outp);
    }
    // ---- upload the table, then issue every group on its stream
    Slot& sl = slots_[turn_];
    turn_ ^= 1;
    if (sl.done) HIPCHK(hipEventSynchronize(sl.done));  // this slot's previous submit has finished
    sl.used = 0;
    if (!sl.done) HIPCHK(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    const size_t words = std::max<size_t>(table.size(), 2);
    if (sl.pinned_words < words) {
        if (sl.pinned) HIPCHK(hipHostFree(sl.pinned));
        sl.pinned = nullptr;
        HIPCHK(hipHostMalloc((void**)&sl.pinned, words * 8, hipHostMallocDefault));
        sl.pinned_words = words;
    }
    if (sl.table.bytes < words * 8) sl.table.reset(words * 8);
    if (!table.empty()) {
        std::memcpy(sl.pinned, table.data(), table.size() * 8);
        HIPCHK(hipMemcpyAsync(sl.table.ptr, sl.pinned, table.size() * 8, hipMemcpyHostToDevice, s));
    }
    hipEvent_t ready = sl.next_event();
    HIPCHK(hipEventRecord(ready, s));
    const u64* const* dev_table = sl.table.as<const u64*>();
    for (Group& G : groups) {
        HIPCHK(hipStreamWaitEvent(G.st, ready, 0));
        for (const Op& op : G.ops) {
            switch (op.kind) {
                case Op::COPY: launch_copy_slots(dev_table + 2 * op.table_off, op.count, op.elems, G.st); break;
                case Op::MUL: op.mult->multiply(op.a, op.b, op.out, op.count, G.st); break;
                case Op::ROT: op.ek->rotate(op.exponent, op.a, op.out, op.count, G.st); break;
            }
        }
        hipEvent_t e = sl.next_event();
        HIPCHK(hipEventRecord(e, G.st));
        HIPCHK(hipStreamWaitEvent(s, e, 0));
    }
    HIPCHK(hipEventRecord(sl.done, s));
}

}  // namespace fhehip
