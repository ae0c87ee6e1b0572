// Request server: Paterson-Stockmeyer evaluations submitted at any time, admitted by device
// working-set size, and advanced together with continuous batching.
//
// A request is a compiled PS program (algorithms.hpp) over its own device registers.  A worker
// thread runs scheduling rounds: it admits waiting requests in arrival order while their working
// sets fit the byte budget, advances every admitted request through its cheap instructions (constant
// products and adds, truncations), then launches each set of ct x ct multiplications that share a
// ring, a level and a relinearization key as ONE batched mult + relinearize + rescale
// (run_op_group: the key is streamed once for all of them), and finishes requests whose program
// ended.  Requests that arrive while others run join at the next round.  The arithmetic of every
// request is exactly rck_ps_eval's (same instructions, same kernels), so a result does not depend
// on what it was batched with.
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

struct Ring {
    std::vector<Context*> lane;       // one scratch clone per stream lane: the server's own workspaces
    std::vector<const KeyDev*> keys;  // relinearization keys by level
};

struct Request {
    uint64_t id = 0;
    rck_context* ctx = nullptr;
    PsProgram prog;
    const uint64_t* in = nullptr;
    uint64_t* out = nullptr;
    size_t footprint = 0;
    uint64_t* pool = nullptr;
    std::vector<uint64_t*> regs;
    size_t pc = 0;
    Clock::time_point t_submit, t_admit, t_done;
    int status = 1;  // 1 waiting, 2 running, 0 done, <0 failed
    std::string error;
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
    std::vector<hipStream_t> lanes;  // independent multiplication groups run concurrently, one per lane
    std::thread worker;
    // statistics
    uint64_t rounds = 0, mult_launches = 0, mults = 0, completed = 0, failed = 0;
    std::map<int, uint64_t> batch_hist;  // batched multiplications per launch -> launches
    Clock::time_point first_submit, last_done;
    bool any_submit = false;
    std::string fatal;

    void run();
    bool round();
    void fail_all(const std::string& msg);
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

// One scheduling round.  Returns false when there was nothing to do.
bool rck_server::round() {
    std::vector<Request*> run;
    {
        std::lock_guard<std::mutex> lk(mu);
        // admission in arrival order (no overtaking: a large request is not starved)
        while (!waiting.empty() && (max_active == 0 || active.size() < max_active)) {
            Request* r = waiting.front();
            if (used + r->footprint > budget) break;
            waiting.pop_front();
            used += r->footprint;
            peak = std::max(peak, used);
            r->status = 2;
            r->t_admit = Clock::now();
            active.push_back(r);
        }
        run = active;
    }
    if (run.empty()) return false;
    // allocate the working sets of newly admitted requests (outside the lock)
    for (Request* r : run) {
        if (r->pool || r->footprint == 0) continue;
        RCK_CHECK_HIP(hipMalloc(&r->pool, r->footprint));
        const uint32_t n = rings.at(r->ctx).lane[0]->n;
        std::vector<uint64_t*> slot_ptr;
        size_t off = 0;
        for (int sl : r->prog.slot_limbs) {
            slot_ptr.push_back(r->pool + off);
            off += (size_t)sl * 2 * n;
        }
        r->regs.assign(r->prog.reg_limbs.size(), nullptr);
        r->regs[0] = const_cast<uint64_t*>(r->in);
        for (size_t k = 1; k < r->regs.size(); k++)
            if (r->prog.slot_of[k] >= 0) r->regs[k] = slot_ptr[r->prog.slot_of[k]];
    }
    // Group the requests by the (ring, level) of their next multiplication; each group is one
    // batched launch.  Groups go to the stream lanes (largest first, to the least loaded lane); a
    // request's cheap instructions before and after its multiplication run on its group's lane, so
    // every request's instructions stay in order on one stream within the round.
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
    const int nl = (int)lanes.size();
    std::vector<double> load(nl, 0.0);
    auto cheap = [&](Request* r, int li) {
        Ring& g = rings.at(r->ctx);
        while (r->pc < r->prog.code.size() && r->prog.code[r->pc].kind != kPsMultRelinRescale) {
            ps_exec(*g.lane[li], r->prog, r->prog.code[r->pc], r->regs, g.keys, lanes[li]);
            r->pc++;
        }
    };
    for (auto& o : order) {
        const int li = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[li] += o.first;
        std::vector<Request*>& grp = groups[o.second];
        Ring& g = rings.at(o.second.first);
        const int level = o.second.second;
        std::vector<const uint64_t*> as, bs;
        std::vector<uint64_t*> outs;
        for (Request* r : grp) {
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
        }
        mult_launches++;
        mults += grp.size();
        batch_hist[(int)grp.size()]++;
    }
    for (Request* r : tail) cheap(r, 0);
    std::vector<Request*> finished;
    for (Request* r : run) {
        if (r->pc < r->prog.code.size()) continue;
        // a finished request's last instruction ran on some lane: copy on every lane's order by
        // synchronizing first (the copy is small)
        finished.push_back(r);
    }
    for (hipStream_t st : lanes) RCK_CHECK_HIP(hipStreamSynchronize(st));
    for (Request* r : finished) {
        const uint32_t n = rings.at(r->ctx).lane[0]->n;
        const size_t rb = (size_t)2 * r->prog.reg_limbs[r->prog.result] * n * 8;
        RCK_CHECK_HIP(hipMemcpyAsync(r->out, r->regs[r->prog.result], rb, hipMemcpyDeviceToDevice, lanes[0]));
    }
    if (!finished.empty()) RCK_CHECK_HIP(hipStreamSynchronize(lanes[0]));
    rounds++;
    const Clock::time_point now = Clock::now();
    for (Request* r : finished) {
        if (r->pool) RCK_CHECK_HIP(hipFree(r->pool));
        r->pool = nullptr;
        r->regs.clear();
    }
    {
        std::lock_guard<std::mutex> lk(mu);
        for (Request* r : finished) {
            r->t_done = now;
            r->status = 0;
            used -= r->footprint;
            completed++;
            last_done = now;
            active.erase(std::find(active.begin(), active.end(), r));
        }
    }
    if (!finished.empty()) cv_done.notify_all();
    return true;
}

// A device or host failure in the worker: every request in flight or waiting fails (the device state
// of the active ones is unknown), their working sets are released, and later submissions are refused.
void rck_server::fail_all(const std::string& msg) {
    for (hipStream_t st : lanes) (void)hipStreamSynchronize(st);
    std::lock_guard<std::mutex> lk(mu);
    fatal = msg;
    for (Request* r : active) {
        if (r->pool) (void)hipFree(r->pool);
        r->pool = nullptr;
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
    active.clear();
    waiting.clear();
    cv_done.notify_all();
}

void rck_server::run() {
    while (true) {
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stop || !waiting.empty() || !active.empty(); });
            if (stop) return;
        }
        std::string err;
        try {
            round();
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
        for (uint32_t i = 0; i < lanes; i++) {
            hipStream_t st = nullptr;
            hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
            if (e != hipSuccess) {
                for (hipStream_t x : s->lanes) (void)hipStreamDestroy(x);
                delete s;
                RCK_CHECK_HIP(e);
            }
            s->lanes.push_back(st);
        }
        s->worker = std::thread([s] { s->run(); });
        *out = s;
    });
}

int rck_server_register_ring(rck_server* s, rck_context* ctx, const rck_ksk* const* relin_keys, uint32_t n_keys) {
    return sguard([&] {
        RCK_REQUIRE(s && ctx && ctx->c, "null argument");
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
        for (auto& kv : s->requests) {
            if (kv.second->pool) (void)hipFree(kv.second->pool);
            delete kv.second;
        }
        for (auto& kv : s->rings)
            for (Context* c : kv.second.lane) delete c;
        for (hipStream_t st : s->lanes) (void)hipStreamDestroy(st);
        delete s;
    });
}

}  // extern "C"
