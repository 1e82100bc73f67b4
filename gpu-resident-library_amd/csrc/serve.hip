// Serving loop: continuous batching of homomorphic operations of mixed ring
// size, level and kind over resident ciphertexts.
//
// Requests queue in arrival order.  Each scheduling step
//   1. retires lanes whose batch has completed (event query, no blocking) and
//      returns their workspace slots;
//   2. for every idle lane, takes the ring of the oldest waiting request and
//      admits that ring's waiting requests (oldest first, up to max_batch) for
//      which a workspace slot exists or fits in the HBM budget;
//   3. builds one plan for the admitted operations (EvalMult+Rescale, Rotate
//      and EvalAdd share the plan's stage launches), stages its job arrays
//      through the lane's pinned buffer into the lane's device arena, launches
//      it on the lane's stream and records a completion event.
// Requests that arrive while batches run are admitted as soon as a lane and
// slots free up; batches on different lanes overlap on the device.  Nothing
// is allocated per request once the slot pool has grown to the working set.
//
// Completion times come from HIP events measured against an event recorded
// when the server was created, so latencies are device-side completion times
// on the server clock.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <thread>

#include "../../include/gpufhe_capi.h"
#include "handles.hpp"

using namespace gf;

namespace {
using Clock = std::chrono::steady_clock;

struct SRing {
    gf_ctx* ctx = nullptr;
    gf_key* rlk = nullptr;
    std::vector<gf_key*> rot;
    std::vector<uint64_t> rot_k;
    size_t ws_words = 0, out_words = 0;   // workspace, then room for a result
    std::vector<uint64_t*> free_slots;
};
struct SReq {
    uint64_t id;
    gf_request r;
};
struct SLane {
    hipStream_t s = nullptr;
    hipEvent_t done = nullptr;
    void* arena = nullptr;
    void* stage = nullptr;
    size_t cap = 0;
    bool busy = false;
    uint32_t ring = 0;
    double launch_ms = 0;
    std::vector<SReq> reqs;
    std::vector<uint64_t*> slots;
    std::unique_ptr<Plan> plan;
};
}  // namespace

struct gf_server {
    std::vector<SRing> rings;
    std::vector<SLane> lanes;
    std::deque<SReq> queue;
    size_t budget = 0, used = 0;
    uint32_t max_batch = 64;
    hipEvent_t t0_ev = nullptr;
    Clock::time_point t0;
    std::vector<gf_completion> done;
    uint64_t next_id = 0;
    uint64_t batches = 0, admitted = 0;
    std::vector<void*> allocations;

    double now_ms() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }

    ~gf_server() {
        for (auto& l : lanes) {
            if (l.s) (void)hipStreamSynchronize(l.s);
        }
        for (auto& l : lanes) {
            if (l.done) (void)hipEventDestroy(l.done);
            if (l.s) (void)hipStreamDestroy(l.s);
            if (l.arena) (void)hipFree(l.arena);
            if (l.stage) (void)hipHostFree(l.stage);
        }
        for (void* p : allocations) (void)hipFree(p);
        if (t0_ev) (void)hipEventDestroy(t0_ev);
    }

    void retire() {
        for (auto& l : lanes) {
            if (!l.busy) continue;
            const hipError_t q = hipEventQuery(l.done);
            if (q == hipErrorNotReady) continue;
            hip_check(q, "batch completion");
            float ms = 0;
            hip_check(hipEventElapsedTime(&ms, t0_ev, l.done), "completion time");
            for (size_t i = 0; i < l.reqs.size(); ++i) {
                gf_completion c{};
                c.id = l.reqs[i].id;
                c.arrival_ms = l.reqs[i].r.arrival_ms;
                c.launch_ms = l.launch_ms;
                c.done_ms = ms;
                c.batch_size = (uint32_t)l.reqs.size();
                c.ring = l.ring;
                c.op = l.reqs[i].r.op;
                c.level = l.reqs[i].r.level;
                done.push_back(c);
            }
            SRing& R = rings[l.ring];
            for (uint64_t* s : l.slots) R.free_slots.push_back(s);
            l.reqs.clear();
            l.slots.clear();
            l.plan.reset();
            l.busy = false;
        }
    }

    uint64_t* take_slot(SRing& R) {
        if (!R.free_slots.empty()) {
            uint64_t* s = R.free_slots.back();
            R.free_slots.pop_back();
            return s;
        }
        const size_t bytes = (R.ws_words + R.out_words) * 8;
        if (used + bytes > budget) return nullptr;
        void* p = nullptr;
        hip_check(hipMalloc(&p, bytes), "workspace slot");
        allocations.push_back(p);
        used += bytes;
        return (uint64_t*)p;
    }

    // the lane's job arena and staging area hold at least `bytes` (the lane is
    // idle: its last batch, and the copy that staged it, have completed)
    void fit_lane(SLane& l, size_t bytes) {
        if (bytes <= l.cap) return;
        size_t cap = l.cap;
        while (cap < bytes) cap *= 2;
        void *arena = nullptr, *stage = nullptr;
        hip_check(hipMalloc(&arena, cap), "lane job arena");
        if (hipHostMalloc(&stage, cap, hipHostMallocDefault) != hipSuccess) {
            (void)hipFree(arena);
            throw std::runtime_error("gpufhe: lane staging: out of pinned host memory");
        }
        (void)hipFree(l.arena);
        (void)hipHostFree(l.stage);
        l.arena = arena;
        l.stage = stage;
        l.cap = cap;
    }

    // Admits waiting requests into every idle lane.  A lane's batch is
    // committed (requests leave the queue, the lane turns busy) only once its
    // plan has been staged and launched; on any failure before that the
    // requests stay queued in order, their slots return to the pool, and the
    // error propagates.
    uint32_t admit() {
        uint32_t launched = 0;
        for (auto& l : lanes) {
            if (l.busy || queue.empty()) continue;
            const uint32_t ring = queue.front().r.ring;
            SRing& R = rings[ring];
            PlanBuilder b(*R.ctx->C);
            const Ctx& C = *R.ctx->C;
            std::vector<size_t> pick;
            std::vector<uint64_t*> slots;
            std::unique_ptr<Plan> plan;
            try {
                size_t idx = 0;
                for (auto it = queue.begin(); it != queue.end() && pick.size() < max_batch; ++it, ++idx) {
                    if (it->r.ring != ring) continue;
                    uint64_t* slot = take_slot(R);
                    if (!slot) break;   // footprint: wait for slots to come back
                    slots.push_back(slot);
                    const gf_request& r = it->r;
                    uint64_t* o0 = r.o0 ? r.o0 : slot + R.ws_words;
                    uint64_t* o1 = r.o1 ? r.o1 : slot + R.ws_words + (size_t)C.L * C.n;
                    switch (r.op) {
                        case GF_OP_EVALMULT_RESCALE:
                            b.evalmult_rescale(r.level, r.a0, r.a1, r.b0, r.b1, o0, o1, *R.rlk->K, slot);
                            break;
                        case GF_OP_ROTATE:
                            b.rotate(r.level, R.rot_k[r.rot], r.a0, r.a1, o0, o1, *R.rot[r.rot]->K, slot);
                            break;
                        default:
                            b.add(r.level, r.a0, r.a1, r.b0, r.b1, o0, o1);
                            break;
                    }
                    pick.push_back(idx);
                }
                if (pick.empty()) continue;
                std::vector<char> blob;
                plan = b.finish_host(blob);
                fit_lane(l, blob.size());
                if (!blob.empty()) {
                    std::memcpy(l.stage, blob.data(), blob.size());
                    hip_check(hipMemcpyAsync(l.arena, l.stage, blob.size(), hipMemcpyHostToDevice, l.s), "stage plan");
                }
                plan->base = l.arena;
                plan->run(l.s);
                hip_check(hipEventRecord(l.done, l.s), "completion event");
                hip_check(hipGetLastError(), "batch launch");
            } catch (...) {
                for (uint64_t* s : slots) R.free_slots.push_back(s);
                throw;
            }
            // committed: move the picked requests (in queue order) to the lane
            l.reqs.clear();
            for (size_t i : pick) l.reqs.push_back(queue[i]);
            for (size_t k = pick.size(); k-- > 0;) queue.erase(queue.begin() + (std::ptrdiff_t)pick[k]);
            l.slots = std::move(slots);
            l.plan = std::move(plan);
            l.ring = ring;
            l.launch_ms = now_ms();
            l.busy = true;
            ++launched;
            ++batches;
            admitted += l.reqs.size();
        }
        return launched;
    }

    bool busy() const {
        for (auto& l : lanes)
            if (l.busy) return true;
        return false;
    }

    void check_request(const gf_request& r) const {
        if (r.ring >= rings.size()) throw std::runtime_error("gpufhe: unknown ring");
        const SRing& R = rings[r.ring];
        const Ctx& C = *R.ctx->C;
        if (r.level == 0 || r.level > C.L) throw std::runtime_error("gpufhe: request level outside the ring's context");
        if (!r.a0 || !r.a1) throw std::runtime_error("gpufhe: request without input ciphertext");
        switch (r.op) {
            case GF_OP_EVALMULT_RESCALE:
                if (!R.rlk) throw std::runtime_error("gpufhe: ring has no relinearisation key");
                if (r.level < 2) throw std::runtime_error("gpufhe: EvalMult+Rescale needs two towers");
                if (!r.b0 || !r.b1) throw std::runtime_error("gpufhe: EvalMult+Rescale needs two ciphertexts");
                break;
            case GF_OP_ROTATE:
                if (r.rot >= R.rot.size()) throw std::runtime_error("gpufhe: unknown rotation key");
                break;
            case GF_OP_ADD:
                if (!r.b0 || !r.b1) throw std::runtime_error("gpufhe: EvalAdd needs two ciphertexts");
                break;
            default:
                throw std::runtime_error("gpufhe: unknown operation");
        }
    }
};

extern "C" {

gf_server* gf_server_create(size_t ws_budget_bytes, uint32_t lanes, uint32_t max_batch) {
    return guard_ptr<gf_server>([&] {
        if (lanes == 0 || lanes > 8) throw std::runtime_error("gpufhe: lanes must be in [1, 8]");
        if (max_batch == 0 || max_batch > 4096) throw std::runtime_error("gpufhe: max_batch must be in [1, 4096]");
        auto S = std::make_unique<gf_server>();
        S->budget = ws_budget_bytes;
        S->max_batch = max_batch;
        S->lanes.resize(lanes);
        for (auto& l : S->lanes) {
            hip_check(hipStreamCreateWithFlags(&l.s, hipStreamNonBlocking), "lane stream");
            hip_check(hipEventCreateWithFlags(&l.done, hipEventDefault), "lane event");
            l.cap = (size_t)32 << 20;   // grows (fit_lane) if a batch's job arrays need more
            hip_check(hipMalloc(&l.arena, l.cap), "lane job arena");
            hip_check(hipHostMalloc(&l.stage, l.cap, hipHostMallocDefault), "lane staging");
        }
        hip_check(hipEventCreateWithFlags(&S->t0_ev, hipEventDefault), "clock event");
        hip_check(hipEventRecord(S->t0_ev, S->lanes[0].s), "clock event");
        hip_check(hipEventSynchronize(S->t0_ev), "clock event");
        S->t0 = Clock::now();
        return S.release();
    });
}

void gf_server_destroy(gf_server* S) { delete S; }

int gf_server_add_ring(gf_server* S, gf_ctx* c, gf_key* rlk, gf_key* const* rot, const uint64_t* rot_k, uint32_t nrot,
                       uint32_t* ring_out) {
    return guard([&] {
        SRing R;
        R.ctx = c;
        R.rlk = rlk;
        for (uint32_t i = 0; i < nrot; ++i) {
            if (!(rot_k[i] & 1) || rot_k[i] >= 2ull * c->C->n)
                throw std::runtime_error("gpufhe: automorphism index must be odd and < 2n");
            R.rot.push_back(rot[i]);
            R.rot_k.push_back(rot_k[i]);
        }
        R.ws_words = c->C->ws_words();
        R.out_words = 2 * (size_t)c->C->L * c->C->n;
        S->rings.push_back(std::move(R));
        *ring_out = (uint32_t)S->rings.size() - 1;
    });
}

size_t gf_server_slot_bytes(const gf_server* S, uint32_t ring) {
    if (ring >= S->rings.size()) return 0;
    return (S->rings[ring].ws_words + S->rings[ring].out_words) * 8;
}

double gf_server_now_ms(const gf_server* S) { return S->now_ms(); }

int gf_server_submit(gf_server* S, const gf_request* r, uint64_t* id_out) {
    return guard([&] {
        S->check_request(*r);
        SReq q{S->next_id++, *r};
        if (q.r.arrival_ms < 0) q.r.arrival_ms = S->now_ms();
        S->queue.push_back(q);
        if (id_out) *id_out = q.id;
    });
}

int gf_server_step(gf_server* S, uint32_t* launched) {
    return guard([&] {
        S->retire();
        const uint32_t n = S->admit();
        if (launched) *launched = n;
    });
}

uint32_t gf_server_poll(gf_server* S, gf_completion* out, uint32_t max) {
    const uint32_t n = (uint32_t)std::min<size_t>(max, S->done.size());
    std::copy(S->done.begin(), S->done.begin() + n, out);
    S->done.erase(S->done.begin(), S->done.begin() + n);
    return n;
}

uint32_t gf_server_pending(const gf_server* S) {
    size_t n = S->queue.size();
    for (auto& l : S->lanes) n += l.reqs.size();
    return (uint32_t)n;
}

int gf_server_drain(gf_server* S) {
    return guard([&] {
        while (!S->queue.empty() || S->busy()) {
            S->retire();
            if (S->admit() > 0) continue;
            if (S->busy()) std::this_thread::sleep_for(std::chrono::microseconds(20));
            else if (!S->queue.empty()) throw std::runtime_error("gpufhe: queued requests do not fit the workspace budget");
        }
    });
}

int gf_server_run_trace(gf_server* S, uint32_t count, const gf_request* reqs, gf_completion* out) {
    return guard([&] {
        for (uint32_t i = 1; i < count; ++i)
            if (reqs[i].arrival_ms < reqs[i - 1].arrival_ms) throw std::runtime_error("gpufhe: trace arrivals must be sorted");
        for (uint32_t i = 0; i < count; ++i) S->check_request(reqs[i]);
        std::vector<int64_t> slot_of(count, -1);
        const uint64_t id0 = S->next_id;
        const double base = S->now_ms();
        uint32_t next = 0, finished = 0;
        std::vector<gf_completion> tmp;
        while (finished < count) {
            const double now = S->now_ms() - base;
            while (next < count && reqs[next].arrival_ms <= now) {
                SReq q{S->next_id++, reqs[next]};
                q.r.arrival_ms = base + reqs[next].arrival_ms;   // scheduled arrival, server clock
                S->queue.push_back(q);
                ++next;
            }
            S->retire();
            const uint32_t launched = S->admit();
            for (const gf_completion& c : S->done) {
                if (c.id >= id0 && c.id < id0 + count) {
                    out[c.id - id0] = c;
                    ++finished;
                }
            }
            S->done.clear();
            if (!S->busy() && !S->queue.empty() && launched == 0)
                throw std::runtime_error("gpufhe: queued requests do not fit the workspace budget");
            if (launched == 0) {
                if (S->busy() || next < count) std::this_thread::sleep_for(std::chrono::microseconds(20));
            }
        }
    });
}

int gf_server_stats(const gf_server* S, uint64_t* batches, uint64_t* admitted, size_t* slot_bytes_used) {
    *batches = S->batches;
    *admitted = S->admitted;
    *slot_bytes_used = S->used;
    return 0;
}

}  // extern "C"
