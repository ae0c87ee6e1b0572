// Request server: Paterson-Stockmeyer evaluations submitted at any time, admitted by device
// working-set size, and advanced together with continuous batching.
//
// A request is a compiled PS program (algorithms.hpp) over its own device registers, carved out of one
// device pool the server allocates once (its byte budget) by a first-fit allocator: admission in arrival
// order (no overtaking) as soon as a block of the request's working-set size is free.
//
// A worker thread issues rounds without waiting for the GPU.  A round advances every admitted request
// by one ct x ct multiplication: the requests whose next multiplication shares a ring, a level and a
// relinearization key form ONE batched mult + relinearize + rescale (run_op_group: the key is streamed
// once for all of them), preceded and followed by each request's cheap instructions (constant products
// and adds, truncations) on the same stream lane.  Requests whose program has been issued to its end get
// their output copy on that lane.  Each group records an event on its lane; a request whose next group
// lands on another lane makes that lane wait for its last event, so every request's instructions stay
// in order while independent groups run concurrently on up to 8 lanes.  At most kMaxInflight rounds are
// queued on the GPU: a new arrival joins the next round issued, which is queued while the previous one
// runs, so the device does not idle between rounds.  Rounds retire in order through their events
// (polled, then waited for when the queue is full); a request retires with the round that copied its
// output, and its pool block is released.  The arithmetic of every request is exactly rck_ps_eval's
// (same instructions, same kernels), so a result does not depend on what it was batched with.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <sstream>
#include <thread>

#include "../../include/rnsckks.h"
#include "algorithms.hpp"
#include "capi_types.hpp"

using namespace rck;

namespace {

using Clock = std::chrono::steady_clock;
constexpr int kMaxInflight = 2;      // rounds queued on the device at once
constexpr size_t kPoolAlign = 256;   // sub-allocation granule of the request pool

struct Ring {
    std::vector<Context*> lane;       // one scratch clone per stream lane: the server's own workspaces
    std::vector<const KeyDev*> keys;  // relinearization keys by level
};

struct Round;

struct Request {
    uint64_t id = 0;
    rck_context* ctx = nullptr;
    PsProgram prog;
    const uint64_t* in = nullptr;
    uint64_t* out = nullptr;
    size_t footprint = 0;
    size_t pool_off = 0;
    bool has_block = false;
    std::vector<uint64_t*> regs;
    size_t pc = 0;               // next instruction to issue
    bool issued_all = false;     // every instruction and the output copy are queued
    int last_lane = -1;          // lane of the request's last queued work
    hipEvent_t last_event = nullptr;  // recorded after that work (null: nothing pending)
    Round* last_round = nullptr;      // round that owns last_event
    Clock::time_point t_submit, t_admit, t_done;
    int status = 1;  // 1 waiting, 2 running, 0 done, <0 failed
    std::string error;
};

struct Round {
    std::vector<hipEvent_t> events;     // one per group launch and per output copy, in issue order
    std::vector<Request*> finishing;    // requests whose output copy is in this round
};

// First-fit allocator over [0, size) with coalescing free blocks (host bookkeeping only).
struct PoolAlloc {
    size_t size = 0;
    std::map<size_t, size_t> free;  // offset -> length
    void init(size_t n) {
        size = n;
        free.clear();
        if (n) free[0] = n;
    }
    bool alloc(size_t n, size_t* off) {
        n = (n + kPoolAlign - 1) / kPoolAlign * kPoolAlign;
        for (auto it = free.begin(); it != free.end(); ++it) {
            if (it->second < n) continue;
            *off = it->first;
            const size_t rest = it->second - n;
            const size_t at = it->first + n;
            free.erase(it);
            if (rest) free[at] = rest;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t n) {
        n = (n + kPoolAlign - 1) / kPoolAlign * kPoolAlign;
        auto next = free.lower_bound(off);
        if (next != free.end() && off + n == next->first) {
            n += next->second;
            next = free.erase(next);
        }
        if (next != free.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += n;
                return;
            }
        }
        free[off] = n;
    }
};

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

double percentile(std::vector<double> v, double p) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = p * (double)(v.size() - 1);
    const size_t lo = (size_t)pos;
    const size_t hi = std::min(v.size() - 1, lo + 1);
    return v[lo] + (v[hi] - v[lo]) * (pos - (double)lo);
}

}  // namespace

struct rck_server {
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> waiting;
    std::vector<Request*> active;
    std::map<uint64_t, Request*> requests;
    std::map<rck_context*, Ring> rings;
    size_t budget = 0, used = 0, peak = 0;
    uint32_t max_active = 0;
    uint64_t next_id = 1;
    bool stop = false;
    uint8_t* pool = nullptr;         // the requests' working sets (budget bytes, allocated once)
    PoolAlloc alloc;
    std::vector<hipStream_t> lanes;  // independent multiplication groups run concurrently, one per lane
    std::vector<hipEvent_t> event_pool;
    std::deque<Round*> inflight;     // issued rounds, oldest first
    std::thread worker;
    // statistics
    uint64_t rounds = 0, mult_launches = 0, mults = 0, completed = 0, failed = 0;
    std::map<int, uint64_t> batch_hist;  // batched multiplications per launch -> launches
    Clock::time_point first_submit, last_done;
    bool any_submit = false;
    std::string fatal;

    void run();
    void admit();
    void issue_round();
    bool retire_ready(bool wait_oldest);
    void fail_all(const std::string& msg);
    hipEvent_t take_event();
};

namespace {

template <typename F>
int sguard(F&& f) {
    try {
        f();
        return RCK_OK;
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return RCK_ERR_INTERNAL;
    } catch (const std::exception& e) {
        set_error(e.what());
        return RCK_ERR_INTERNAL;
    }
}

}  // namespace

hipEvent_t rck_server::take_event() {
    if (event_pool.empty()) {
        hipEvent_t e;
        RCK_CHECK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        return e;
    }
    hipEvent_t e = event_pool.back();
    event_pool.pop_back();
    return e;
}

// Admission in arrival order while the pool has a free block of the next request's working-set size.
void rck_server::admit() {
    std::vector<Request*> admitted;
    {
        std::lock_guard<std::mutex> lk(mu);
        while (!waiting.empty() && (max_active == 0 || active.size() < max_active)) {
            Request* r = waiting.front();
            size_t off = 0;
            if (r->footprint && !alloc.alloc(r->footprint, &off)) break;
            waiting.pop_front();
            r->pool_off = off;
            r->has_block = r->footprint != 0;
            used += r->footprint;
            peak = std::max(peak, used);
            r->status = 2;
            r->t_admit = Clock::now();
            active.push_back(r);
            admitted.push_back(r);
        }
    }
    for (Request* r : admitted) {
        const uint32_t n = rings.at(r->ctx).lane[0]->n;
        uint64_t* base = reinterpret_cast<uint64_t*>(pool + r->pool_off);
        std::vector<uint64_t*> slot_ptr;
        size_t off = 0;
        for (int sl : r->prog.slot_limbs) {
            slot_ptr.push_back(base + off);
            off += (size_t)sl * 2 * n;
        }
        r->regs.assign(r->prog.reg_limbs.size(), nullptr);
        r->regs[0] = const_cast<uint64_t*>(r->in);
        for (size_t k = 1; k < r->regs.size(); k++)
            if (r->prog.slot_of[k] >= 0) r->regs[k] = slot_ptr[r->prog.slot_of[k]];
    }
}

// One round: every admitted request with instructions left advances by one multiplication (grouped by
// ring and level into batched launches spread over the lanes), or to its end.
void rck_server::issue_round() {
    std::vector<Request*> run;
    {
        std::lock_guard<std::mutex> lk(mu);
        for (Request* r : active)
            if (!r->issued_all) run.push_back(r);
    }
    Round* rd = new Round();
    const int nl = (int)lanes.size();
    // make lane li wait for r's last work when that was queued on another lane
    auto order_on = [&](Request* r, int li) {
        if (r->last_event && r->last_lane != li) RCK_CHECK_HIP(hipStreamWaitEvent(lanes[li], r->last_event, 0));
    };
    auto mark = [&](const std::vector<Request*>& rs, int li) {
        hipEvent_t e = take_event();
        RCK_CHECK_HIP(hipEventRecord(e, lanes[li]));
        rd->events.push_back(e);
        for (Request* r : rs) {
            r->last_lane = li;
            r->last_event = e;
            r->last_round = rd;
        }
    };
    auto cheap = [&](Request* r, int li) {
        Ring& g = rings.at(r->ctx);
        while (r->pc < r->prog.code.size() && r->prog.code[r->pc].kind != kPsMultRelinRescale) {
            ps_exec(*g.lane[li], r->prog, r->prog.code[r->pc], r->regs, g.keys, lanes[li]);
            r->pc++;
        }
    };
    // the output copy of a request issued to its end (on the lane of its last work)
    auto finish = [&](Request* r, int li) {
        const uint32_t n = rings.at(r->ctx).lane[0]->n;
        const size_t rb = (size_t)2 * r->prog.reg_limbs[r->prog.result] * n * 8;
        RCK_CHECK_HIP(hipMemcpyAsync(r->out, r->regs[r->prog.result], rb, hipMemcpyDeviceToDevice, lanes[li]));
        r->issued_all = true;
        rd->finishing.push_back(r);
    };
    std::map<std::pair<rck_context*, int>, std::vector<Request*>> groups;
    std::vector<Request*> tail;  // no multiplication left: only cheap instructions
    for (Request* r : run) {
        size_t k = r->pc;
        while (k < r->prog.code.size() && r->prog.code[k].kind != kPsMultRelinRescale) k++;
        if (k < r->prog.code.size())
            groups[{r->ctx, r->prog.code[k].key_level}].push_back(r);
        else
            tail.push_back(r);
    }
    std::vector<std::pair<double, std::pair<rck_context*, int>>> order;
    for (auto& kv : groups)
        order.push_back({(double)kv.second.size() * (kv.first.second + 1) * rings.at(kv.first.first).lane[0]->n,
                         kv.first});
    std::sort(order.begin(), order.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
    std::vector<double> load(nl, 0.0);
    for (auto& o : order) {
        const int li = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[li] += o.first;
        std::vector<Request*>& grp = groups[o.second];
        Ring& g = rings.at(o.second.first);
        const int level = o.second.second;
        std::vector<const uint64_t*> as, bs;
        std::vector<uint64_t*> outs;
        for (Request* r : grp) {
            order_on(r, li);
            cheap(r, li);
            const PsInstr& ins = r->prog.code[r->pc];
            as.push_back(r->regs[ins.a]);
            bs.push_back(r->regs[ins.b]);
            outs.push_back(r->regs[ins.dst]);
        }
        run_op_group(*g.lane[li], RCK_OP_MULT_RELIN_RESCALE, level + 1, g.keys[level], as, bs, outs, lanes[li]);
        for (Request* r : grp) {
            r->pc++;
            cheap(r, li);  // up to the next multiplication (next round) or the end
            if (r->pc >= r->prog.code.size()) finish(r, li);
        }
        mark(grp, li);
        mult_launches++;
        mults += grp.size();
        batch_hist[(int)grp.size()]++;
    }
    for (Request* r : tail) {
        const int li = r->last_lane >= 0 ? r->last_lane : 0;
        order_on(r, li);
        cheap(r, li);
        finish(r, li);
        mark({r}, li);
    }
    rounds++;
    inflight.push_back(rd);
}

// Retires the rounds whose events have all completed, oldest first (waits for the oldest when asked).
// Returns true when a round retired.
bool rck_server::retire_ready(bool wait_oldest) {
    bool any = false;
    while (!inflight.empty()) {
        Round* rd = inflight.front();
        bool done = true;
        for (hipEvent_t e : rd->events) {
            if (wait_oldest && !any) {
                RCK_CHECK_HIP(hipEventSynchronize(e));
                continue;
            }
            const hipError_t q = hipEventQuery(e);
            if (q == hipErrorNotReady) {
                done = false;
                break;
            }
            RCK_CHECK_HIP(q);
        }
        if (!done) break;
        inflight.pop_front();
        any = true;
        const Clock::time_point now = Clock::now();
        // a retired round's events are no longer pending work for anyone
        for (Request* r : active)
            if (r->last_round == rd) {
                r->last_event = nullptr;
                r->last_round = nullptr;
            }
        for (hipEvent_t e : rd->events) event_pool.push_back(e);
        {
            std::lock_guard<std::mutex> lk(mu);
            for (Request* r : rd->finishing) {
                if (r->has_block) alloc.release(r->pool_off, r->footprint);
                r->has_block = false;
                r->regs.clear();
                r->t_done = now;
                r->status = 0;
                used -= r->footprint;
                completed++;
                last_done = now;
                active.erase(std::find(active.begin(), active.end(), r));
            }
        }
        if (!rd->finishing.empty()) cv_done.notify_all();
        delete rd;
    }
    return any;
}

// A device or host failure in the worker: every request in flight or waiting fails (the device state
// of the active ones is unknown), their working sets are released, and later submissions are refused.
void rck_server::fail_all(const std::string& msg) {
    for (hipStream_t st : lanes) (void)hipStreamSynchronize(st);
    for (Round* rd : inflight) {
        for (hipEvent_t e : rd->events) event_pool.push_back(e);
        delete rd;
    }
    inflight.clear();
    std::lock_guard<std::mutex> lk(mu);
    fatal = msg;
    for (Request* r : active) {
        r->has_block = false;
        r->regs.clear();
        r->status = -1;
        r->error = msg;
        failed++;
    }
    for (Request* r : waiting) {
        r->status = -1;
        r->error = msg;
        failed++;
    }
    used = 0;
    alloc.init(budget);
    active.clear();
    waiting.clear();
    cv_done.notify_all();
}

void rck_server::run() {
    while (true) {
        std::string err;
        try {
            {
                std::lock_guard<std::mutex> lk(mu);
                if (stop) break;
            }
            admit();
            retire_ready(false);
            bool issuable = false;
            {
                std::lock_guard<std::mutex> lk(mu);
                for (Request* r : active) issuable |= !r->issued_all;
            }
            if (issuable && (int)inflight.size() < kMaxInflight) {
                issue_round();
                continue;
            }
            if (!inflight.empty()) {
                retire_ready(true);  // the queue is full or nothing new to issue: wait for the oldest round
                continue;
            }
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stop || !waiting.empty(); });
        } catch (const Error& e) {
            err = e.msg;
        } catch (const std::exception& e) {
            err = std::string("server worker: ") + e.what();
        } catch (...) {
            err = "server worker: unknown exception";
        }
        if (!err.empty()) {
            fail_all(err);
            return;
        }
    }
}

extern "C" {

int rck_server_create(uint64_t hbm_budget_bytes, uint32_t max_active, uint32_t lanes, rck_server** out) {
    return sguard([&] {
        RCK_REQUIRE(out != nullptr, "null output");
        RCK_REQUIRE(hbm_budget_bytes > 0, "the working-set budget must be positive");
        RCK_REQUIRE(lanes >= 1 && lanes <= 8, "lanes must be in 1..=8");
        rck_server* s = new rck_server();
        s->budget = hbm_budget_bytes;
        s->max_active = max_active;
        // the requests' working sets: one device pool of the budget (rounded to the sub-allocation
        // granule), allocated once; requests get blocks of it at admission
        const size_t pool_bytes = (hbm_budget_bytes + kPoolAlign - 1) / kPoolAlign * kPoolAlign;
        hipError_t e = hipMalloc(&s->pool, pool_bytes);
        for (uint32_t i = 0; e == hipSuccess && i < lanes; i++) {
            hipStream_t st = nullptr;
            e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
            if (e == hipSuccess) s->lanes.push_back(st);
        }
        if (e != hipSuccess) {
            for (hipStream_t x : s->lanes) (void)hipStreamDestroy(x);
            if (s->pool) (void)hipFree(s->pool);
            delete s;
            RCK_CHECK_HIP(e);
        }
        s->alloc.init(pool_bytes);
        s->worker = std::thread([s] { s->run(); });
        *out = s;
    });
}

int rck_server_register_ring(rck_server* s, rck_context* ctx, const rck_ksk* const* relin_keys, uint32_t n_keys) {
    return sguard([&] {
        RCK_REQUIRE(s && ctx && ctx->c, "null argument");
        require_fp64(*ctx->c);
        std::lock_guard<std::mutex> lk(s->mu);
        RCK_REQUIRE(s->requests.empty(), "rings must be registered before the first request");
        Ring& g = s->rings[ctx];
        while (g.lane.size() < s->lanes.size()) g.lane.push_back(ctx->c->scratch_clone());
        g.keys.assign(ctx->c->L, nullptr);
        for (uint32_t l = 0; l < n_keys && l < ctx->c->L; l++) {
            if (!relin_keys || !relin_keys[l]) continue;
            RCK_REQUIRE((int)relin_keys[l]->level == (int)l,
                        "relinearization key at index " + std::to_string(l) + " has another level");
            RCK_REQUIRE(relin_keys[l]->owner == ctx, "relinearization key belongs to another context");
            g.keys[l] = relin_keys[l];
        }
    });
}

int rck_server_submit(rck_server* s, rck_context* ctx, const uint64_t* ct, uint32_t limbs, uint32_t scale_bits,
                      const double* coeffs, uint32_t count, uint64_t* out, uint64_t* request_id) {
    return sguard([&] {
        RCK_REQUIRE(s && ctx && ct && out && request_id, "null argument");
        Request* r = new Request();
        try {
            std::vector<bool> has;
            {
                std::lock_guard<std::mutex> lk(s->mu);
                RCK_REQUIRE(s->fatal.empty(), "server stopped after a device error: " + s->fatal);
                auto it = s->rings.find(ctx);
                RCK_REQUIRE(it != s->rings.end(), "context not registered with the server");
                for (const KeyDev* k : it->second.keys) has.push_back(k != nullptr);
            }
            r->prog = ps_compile(*ctx->c, (int)limbs, scale_bits, coeffs, (int)count, has);
            r->footprint = r->prog.footprint(ctx->c->n);
            RCK_REQUIRE(r->footprint <= s->budget, "request working set exceeds the server budget");
        } catch (...) {
            delete r;
            throw;
        }
        r->ctx = ctx;
        r->in = ct;
        r->out = out;
        r->t_submit = Clock::now();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            r->id = s->next_id++;
            if (!s->any_submit) {
                s->first_submit = r->t_submit;
                s->any_submit = true;
            }
            s->requests[r->id] = r;
            s->waiting.push_back(r);
            *request_id = r->id;
        }
        s->cv_work.notify_one();
    });
}

int rck_server_wait(rck_server* s, uint64_t id, double timeout_s, int* done, uint32_t* out_limbs,
                    uint32_t* out_scale, double* latency_ms) {
    return sguard([&] {
        RCK_REQUIRE(s != nullptr, "null server");
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->requests.find(id);
        RCK_REQUIRE(it != s->requests.end(), "unknown request id");
        Request* r = it->second;
        const auto until = Clock::now() + std::chrono::duration<double>(std::max(0.0, timeout_s));
        s->cv_done.wait_until(lk, until, [&] { return r->status <= 0; });
        RCK_REQUIRE(r->status >= 0, "request failed: " + r->error);
        if (done) *done = r->status == 0;
        if (r->status == 0) {
            if (out_limbs) *out_limbs = (uint32_t)r->prog.reg_limbs[r->prog.result];
            if (out_scale) *out_scale = r->prog.reg_scale[r->prog.result];
            if (latency_ms) *latency_ms = ms_between(r->t_submit, r->t_done);
        }
    });
}

int rck_server_drain(rck_server* s, double timeout_s) {
    return sguard([&] {
        RCK_REQUIRE(s != nullptr, "null server");
        std::unique_lock<std::mutex> lk(s->mu);
        const auto until = Clock::now() + std::chrono::duration<double>(std::max(0.0, timeout_s));
        const bool ok =
            s->cv_done.wait_until(lk, until, [&] { return (s->waiting.empty() && s->active.empty()) || !s->fatal.empty(); });
        RCK_REQUIRE(s->fatal.empty(), "server stopped after a device error: " + s->fatal);
        RCK_REQUIRE(ok, "requests still in flight at the time-out");
    });
}

int rck_server_stats(rck_server* s, char* buf, size_t len) {
    return sguard([&] {
        RCK_REQUIRE(s && buf, "null argument");
        std::lock_guard<std::mutex> lk(s->mu);
        std::vector<double> lat, queue;
        for (auto& kv : s->requests) {
            const Request* r = kv.second;
            if (r->status != 0) continue;
            lat.push_back(ms_between(r->t_submit, r->t_done));
            queue.push_back(ms_between(r->t_submit, r->t_admit));
        }
        double mean = 0;
        for (double v : lat) mean += v;
        mean = lat.empty() ? 0 : mean / lat.size();
        const double span = s->any_submit && s->completed ? ms_between(s->first_submit, s->last_done) : 0.0;
        std::ostringstream o;
        o << "{\"completed\": " << s->completed << ", \"failed\": " << s->failed << ", \"waiting\": " << s->waiting.size()
          << ", \"active\": " << s->active.size() << ", \"p50_ms\": " << percentile(lat, 0.5)
          << ", \"p99_ms\": " << percentile(lat, 0.99) << ", \"mean_ms\": " << mean
          << ", \"max_ms\": " << (lat.empty() ? 0.0 : *std::max_element(lat.begin(), lat.end()))
          << ", \"queue_p50_ms\": " << percentile(queue, 0.5) << ", \"queue_p99_ms\": " << percentile(queue, 0.99)
          << ", \"span_ms\": " << span << ", \"throughput_rps\": " << (span > 0 ? s->completed * 1000.0 / span : 0.0)
          << ", \"rounds\": " << s->rounds << ", \"mult_launches\": " << s->mult_launches << ", \"mults\": " << s->mults
          << ", \"budget_bytes\": " << s->budget << ", \"peak_bytes\": " << s->peak << ", \"batch_hist\": {";
        bool first = true;
        for (auto& kv : s->batch_hist) {
            o << (first ? "" : ", ") << "\"" << kv.first << "\": " << kv.second;
            first = false;
        }
        o << "}}";
        const std::string str = o.str();
        RCK_REQUIRE(len > str.size(), "stats buffer too small");
        std::memcpy(buf, str.c_str(), str.size() + 1);
    });
}

int rck_server_destroy(rck_server* s) {
    return sguard([&] {
        if (!s) return;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        s->cv_work.notify_all();
        if (s->worker.joinable()) s->worker.join();
        for (hipStream_t st : s->lanes) (void)hipStreamSynchronize(st);
        for (Round* rd : s->inflight) {
            for (hipEvent_t e : rd->events) s->event_pool.push_back(e);
            delete rd;
        }
        for (hipEvent_t e : s->event_pool) (void)hipEventDestroy(e);
        for (auto& kv : s->requests) delete kv.second;
        for (auto& kv : s->rings)
            for (Context* c : kv.second.lane) delete c;
        for (hipStream_t st : s->lanes) (void)hipStreamDestroy(st);
        if (s->pool) (void)hipFree(s->pool);
        delete s;
    });
}

}  // extern "C"
