// Batched executor: many independent ciphertext operations (at uneven levels, with mixed
// operation kinds) planned into grouped launches and captured in one hipGraph.
//
// Grouping: ops with the same (kind, limbs, key) form a group; a group runs the crate's
// operation for all of its B ciphertexts in the same launches (mult_batch / keyswitch_batch),
// so every kernel sees B ciphertexts at once and the key stream of a keyswitch is read from HBM
// once per group (the digit-MAC grid places the B users of a key tile on one XCD back to back).
// The arithmetic is exactly the per-op path's: the same kernels on the same operands.
//
// Tuning: the digit-MAC launch shape (eval limbs per group, digit splits) is chosen per
// (N, limbs, digits, B) by timing the candidates on the group's own operands (autotune), and
// cached in the context's tuning table.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <sstream>

#include "../../include/rnsckks.h"
#include "capi_types.hpp"

using namespace rck;

namespace {

struct BatchOp {
    int kind;
    uint32_t limbs;
    const uint64_t* a;
    const uint64_t* b;
    const rck_ksk* key;
    uint64_t* out;
};

struct BatchGroup {
    int kind;
    uint32_t limbs;
    const rck_ksk* key;
    std::vector<int> ops;
    KsShape shape;
    float tuned_ms = 0;
    std::string tune_log;
};

}  // namespace

// Independent groups are spread over up to kLanes capture lanes: each lane has its own stream and
// scratch clone (workspace; lane 0 uses the batch's own), and the graph forks into the lanes at its
// start and joins at the end, so small groups (few ciphertexts, short launches) run side by side and
// fill the GPU.
constexpr int kLanes = 4;

struct rck_batch {
    rck_context* ctx = nullptr;
    Context* sc = nullptr;  // scratch clone: own workspace, shared tables (tuning, eager runs)
    hipStream_t stream = nullptr;
    std::vector<Context*> lane_sc;         // per capture lane: scratch clone (lane 0: sc itself)
    std::vector<hipStream_t> lane_stream;  // per capture lane: stream (lane 0: stream itself)
    std::vector<std::vector<int>> lane_groups;
    std::vector<hipEvent_t> events;  // fork + one join per lane (capture dependencies only)
    std::vector<BatchOp> ops;
    std::vector<BatchGroup> groups;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool compiled = false;
    double capture_ms = 0;
};

template <typename F>
static int bguard(F&& f) {
    try {
        f();
        return RCK_OK;
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::exception& e) {
        set_error(e.what());
        return RCK_ERR_INTERNAL;
    }
}

static void run_group(rck_batch* bt, const BatchGroup& g, hipStream_t s, KsShape shape, Context* c = nullptr) {
    std::vector<const uint64_t*> as, bs;
    std::vector<uint64_t*> outs;
    for (int i : g.ops) {
        as.push_back(bt->ops[i].a);
        bs.push_back(bt->ops[i].b);
        outs.push_back(bt->ops[i].out);
    }
    run_op_group(c ? *c : *bt->sc, g.kind, (int)g.limbs, g.key, as, bs, outs, s, shape);
}

static bool has_keyswitch(int kind) { return kind != RCK_OP_MULT_PLAIN; }

// Work estimate of a group for the lane assignment: keyswitch digits x eval limbs x ciphertexts
// (the digit NTTs and key MAC dominate), or limbs x ciphertexts for a plaintext product.
static double group_cost(const Context& c, const BatchGroup& g) {
    const double B = (double)g.ops.size();
    if (!has_keyswitch(g.kind)) return 0.1 * B * g.limbs;
    const int E = (int)g.limbs + (int)g.key->aux;
    const int D = g.key->aux ? (int)((c.qbits[g.limbs] + 29) / 30) : (int)g.key->digits;
    return B * D * E + 0.05 * D * E * 2;  // + the key stream, read once per group
}

// Releases the lanes' own streams, scratch clones and events (lane 0 borrows the batch's).
static void release_lanes(rck_batch* bt) {
    for (size_t l = 1; l < bt->lane_stream.size(); l++) (void)hipStreamSynchronize(bt->lane_stream[l]);
    for (hipEvent_t ev : bt->events) (void)hipEventDestroy(ev);
    for (size_t l = 1; l < bt->lane_stream.size(); l++) (void)hipStreamDestroy(bt->lane_stream[l]);
    for (size_t l = 1; l < bt->lane_sc.size(); l++) delete bt->lane_sc[l];
    bt->events.clear();
    bt->lane_stream.clear();
    bt->lane_sc.clear();
    bt->lane_groups.clear();
}

// Longest-processing-time assignment of the groups to min(kLanes, groups) lanes (a compile retried
// after a failure starts from scratch).
static void assign_lanes(rck_batch* bt) {
    release_lanes(bt);
    const int nl = std::max(1, std::min<int>(kLanes, (int)bt->groups.size()));
    std::vector<int> order(bt->groups.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = (int)i;
    std::vector<double> cost(bt->groups.size());
    for (size_t i = 0; i < cost.size(); i++) cost[i] = group_cost(*bt->sc, bt->groups[i]);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    bt->lane_groups.assign(nl, {});
    std::vector<double> load(nl, 0.0);
    for (int gi : order) {
        const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        bt->lane_groups[l].push_back(gi);
        load[l] += cost[gi];
    }
    for (int l = 0; l < nl; l++) {
        std::sort(bt->lane_groups[l].begin(), bt->lane_groups[l].end());  // submission order within a lane
        if (l == 0) {  // lane 0 runs on the batch's own stream and scratch
            bt->lane_sc.push_back(bt->sc);
            bt->lane_stream.push_back(bt->stream);
            continue;
        }
        bt->lane_sc.push_back(bt->sc->scratch_clone());
        hipStream_t st;
        RCK_CHECK_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        bt->lane_stream.push_back(st);
    }
    for (int e = 0; e < nl; e++) {  // the fork, then one join per lane 1 .. nl-1
        hipEvent_t ev;
        RCK_CHECK_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        bt->events.push_back(ev);
    }
}

// Every lane runs its groups on its own stream with its own scratch clone; the lanes start after
// the work already queued on `s` and `s` continues after all of them (as graph dependencies when
// `s` is capturing).
static void run_lanes(rck_batch* bt, hipStream_t s) {
    const int nl = (int)bt->lane_stream.size();
    if (nl > 1) RCK_CHECK_HIP(hipEventRecord(bt->events[0], s));
    for (int l = 1; l < nl; l++) {
        RCK_CHECK_HIP(hipStreamWaitEvent(bt->lane_stream[l], bt->events[0], 0));
        for (int gi : bt->lane_groups[l]) {
            const BatchGroup& g = bt->groups[gi];
            run_group(bt, g, bt->lane_stream[l], g.shape, bt->lane_sc[l]);
        }
        RCK_CHECK_HIP(hipEventRecord(bt->events[l], bt->lane_stream[l]));
    }
    for (int gi : bt->lane_groups[0]) run_group(bt, bt->groups[gi], s, bt->groups[gi].shape, bt->lane_sc[0]);
    for (int l = 1; l < nl; l++) RCK_CHECK_HIP(hipStreamWaitEvent(s, bt->events[l], 0));
}

static void check_group(const Context& c, const BatchGroup& g) {
    RCK_REQUIRE(g.limbs >= 1 && g.limbs <= c.L, "batch op limb count out of range");
    if (g.kind == RCK_OP_MULT_RELIN_RESCALE) RCK_REQUIRE(g.limbs > 1, "rescale requires at least two levels");
    if (g.kind == RCK_OP_ROTATE || g.kind == RCK_OP_MULT_RELIN || g.kind == RCK_OP_MULT_RELIN_RESCALE) {
        RCK_REQUIRE(g.key != nullptr, "null key");
        if (g.kind == RCK_OP_ROTATE)
            RCK_REQUIRE((int)g.key->level == (int)g.limbs - 1, "rotation key level must match the ciphertext level");
    }
}

static std::string shape_json(const KsShape& sh, float ms) {
    std::ostringstream o;
    o << "{\"eval_limb_group\": " << sh.group << ", \"digit_splits\": " << sh.dsplit << ", \"mac_rows\": "
      << sh.rows << ", \"load_schedule\": " << sh.prefetch << ", \"mac_width\": " << sh.width << ", \"ms\": " << ms
      << "}";
    return o.str();
}

static void autotune(rck_batch* bt, BatchGroup& g) {
    Context& c = *bt->sc;
    const int E = (int)g.limbs + (int)g.key->aux;
    const int D = g.key->aux ? (int)((c.qbits[g.limbs] + 29) / 30) : (int)g.key->digits;
    const int B = (int)g.ops.size();
    // the key's methods: digit NTT (data), exact convolution (mdata), or both
    std::vector<bool> methods;
    if (g.key->data) methods.push_back(false);
    if (g.key->mdata) methods.push_back(true);
    {
        // every method already tuned for this shape: take the fastest entry
        bool all = true;
        float best = 1e30f;
        for (bool ex : methods) {
            KsShape sh;
            float ms = 0;
            if (!ks_tuned_find(c, (int)g.limbs, E, D, B, &sh, &ms, ex)) {
                all = false;
                break;
            }
            if (ms < best) {
                best = ms;
                g.shape = sh;
                g.tuned_ms = ms;
            }
        }
        if (all) return;
    }
    hipEvent_t e0, e1;
    RCK_CHECK_HIP(hipEventCreate(&e0));
    RCK_CHECK_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    KsShape best_shape{};
    std::vector<KsShape> cands;
    for (bool ex : methods)
        for (const KsShape& sh : ks_candidates(c, E, D, B, ex)) cands.push_back(sh);
    float best_m[2] = {1e30f, 1e30f};
    KsShape shape_m[2] = {};
    for (const KsShape& sh : cands) {
        run_group(bt, g, bt->stream, sh);  // warm-up (sizes the scratch)
        RCK_CHECK_HIP(hipEventRecord(e0, bt->stream));
        const int reps = 2;
        for (int r = 0; r < reps; r++) run_group(bt, g, bt->stream, sh);
        RCK_CHECK_HIP(hipEventRecord(e1, bt->stream));
        RCK_CHECK_HIP(hipEventSynchronize(e1));
        float ms = 0;
        RCK_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
        ms /= reps;
        g.tune_log += (g.tune_log.empty() ? "" : ", ") + shape_json(sh, ms);
        const int m = sh.prefetch == kExactShape ? 1 : 0;
        if (ms < best_m[m]) {
            best_m[m] = ms;
            shape_m[m] = sh;
        }
        if (ms < best) {
            best = ms;
            best_shape = sh;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    g.shape = best_shape;
    g.tuned_ms = best;
    for (bool ex : methods) ks_tuned_store(c, (int)g.limbs, E, D, B, shape_m[ex ? 1 : 0], best_m[ex ? 1 : 0], ex);
}

// Per-op tuning: time every candidate shape of keyswitch_batch for B ciphertexts at `limbs` limbs
// with `key` (on zero operands: the kernels' timing does not depend on the residues) and store the
// fastest in the process-wide table that per-op calls consult.
static void tune_keyswitch(Context& c, int limbs, const KeyDev* key, int B, KsShape* best_out, float* best_ms,
                           std::string* log) {
    const long long n = c.n, P = (long long)limbs * n;
    const int E = limbs + (int)key->aux;
    const int D = key->aux ? (int)((c.qbits[limbs] + 29) / 30) : (int)key->digits;
    uint64_t *in = nullptr, *out = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    try {
        RCK_CHECK_HIP(hipMalloc(&in, (size_t)B * P * 8));
        RCK_CHECK_HIP(hipMalloc(&out, (size_t)B * 2 * P * 8));
        RCK_CHECK_HIP(hipMemset(in, 0, (size_t)B * P * 8));
        RCK_CHECK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        RCK_CHECK_HIP(hipEventCreate(&e0));
        RCK_CHECK_HIP(hipEventCreate(&e1));
        std::vector<uint64_t*> outs;
        for (int b = 0; b < B; b++) outs.push_back(out + b * 2 * P);
        float best = 1e30f;
        KsShape bs{};
        std::vector<bool> methods;
        if (key->data) methods.push_back(false);
        if (key->mdata) methods.push_back(true);
        std::vector<KsShape> cands;
        for (bool ex : methods)
            for (const KsShape& sh : ks_candidates(c, E, D, B, ex)) cands.push_back(sh);
        float best_m[2] = {1e30f, 1e30f};
        KsShape shape_m[2] = {};
        for (const KsShape& sh : cands) {
            keyswitch_batch(c, in, P, B, limbs, key, nullptr, nullptr, 0, outs, s, sh);
            RCK_CHECK_HIP(hipEventRecord(e0, s));
            for (int r = 0; r < 2; r++) keyswitch_batch(c, in, P, B, limbs, key, nullptr, nullptr, 0, outs, s, sh);
            RCK_CHECK_HIP(hipEventRecord(e1, s));
            RCK_CHECK_HIP(hipEventSynchronize(e1));
            float ms = 0;
            RCK_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (log) *log += (log->empty() ? "" : ", ") + shape_json(sh, ms / 2);
            const int m = sh.prefetch == kExactShape ? 1 : 0;
            if (ms / 2 < best_m[m]) {
                best_m[m] = ms / 2;
                shape_m[m] = sh;
            }
            if (ms / 2 < best) {
                best = ms / 2;
                bs = sh;
            }
        }
        for (bool ex : methods) ks_tuned_store(c, limbs, E, D, B, shape_m[ex ? 1 : 0], best_m[ex ? 1 : 0], ex);
        *best_out = bs;
        *best_ms = best;
    } catch (...) {
        if (in) (void)hipFree(in);
        if (out) (void)hipFree(out);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (s) (void)hipStreamDestroy(s);
        throw;
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(in);
    (void)hipFree(out);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
}

extern "C" {

int rck_tune_keyswitch(rck_context* ctx, uint32_t limbs, const rck_ksk* key, uint32_t batch, char* buf, size_t len) {
    return bguard([&] {
        RCK_REQUIRE(ctx && ctx->c && key, "null argument");
        require_fp64(*ctx->c);
        RCK_REQUIRE(limbs >= 1 && limbs <= ctx->c->L && (int)key->level == (int)limbs - 1,
                    "keyswitch key level must match the limb count");
        RCK_REQUIRE(batch >= 1, "batch must be positive");
        const bool prof = prof_enabled();
        prof_set(false);
        KsShape sh;
        float ms = 0;
        std::string log;
        try {
            tune_keyswitch(*ctx->c, (int)limbs, key, (int)batch, &sh, &ms, &log);
        } catch (...) {
            prof_set(prof);
            throw;
        }
        prof_set(prof);
        if (buf) {
            std::ostringstream o;
            o << "{\"limbs\": " << limbs << ", \"batch\": " << batch << ", \"eval_limb_group\": " << sh.group
              << ", \"digit_splits\": " << sh.dsplit << ", \"mac_rows\": " << sh.rows << ", \"mac_width\": "
              << sh.width << ", \"key_prefetch\": "
              << sh.prefetch << ", \"keyswitch_ms\": " << ms << ", \"candidates\": [" << log << "]}";
            std::string r = o.str();
            RCK_REQUIRE(len > r.size(), "report buffer too small");
            std::memcpy(buf, r.c_str(), r.size() + 1);
        }
    });
}

int rck_tune_table_export(char* buf, size_t len, size_t* needed) {
    return bguard([&] {
        const std::string t = ks_tuned_export();
        if (needed) *needed = t.size() + 1;
        if (!buf) return;
        RCK_REQUIRE(len > t.size(), "tuning table buffer too small");
        std::memcpy(buf, t.c_str(), t.size() + 1);
    });
}

int rck_tune_table_import(const char* text, uint32_t* entries) {
    return bguard([&] {
        RCK_REQUIRE(text != nullptr, "null tuning table");
        const int k = ks_tuned_import(std::string(text));
        if (entries) *entries = (uint32_t)k;
    });
}

int rck_batch_create(rck_context* ctx, rck_batch** out) {
    return bguard([&] {
        RCK_REQUIRE(ctx && ctx->c && out, "null argument");
        require_fp64(*ctx->c);
        rck_batch* b = new rck_batch();
        b->ctx = ctx;
        b->sc = ctx->c->scratch_clone();
        RCK_CHECK_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        *out = b;
    });
}

int rck_batch_add(rck_batch* bt, int op, uint32_t limbs, const uint64_t* in0, const uint64_t* in1,
                  const rck_ksk* key, uint64_t* out) {
    return bguard([&] {
        RCK_REQUIRE(bt, "null batch");
        RCK_REQUIRE(!bt->compiled, "batch already compiled");
        RCK_REQUIRE(op >= RCK_OP_MULT_RELIN && op <= RCK_OP_MULT_RELIN_RESCALE, "unknown batch op kind");
        RCK_REQUIRE(in0 && out, "null operand");
        RCK_REQUIRE(op == RCK_OP_ROTATE || in1, "null second operand");
        bt->ops.push_back(BatchOp{op, limbs, in0, in1, key, out});
    });
}

int rck_batch_compile(rck_batch* bt, int tune) {
    return bguard([&] {
        RCK_REQUIRE(bt && !bt->compiled, "batch missing or already compiled");
        RCK_REQUIRE(!bt->ops.empty(), "batch has no operations");
        Context& c = *bt->sc;
        // a compile retried after a failure starts from no groups (every op joins exactly one group)
        bt->groups.clear();
        release_lanes(bt);
        std::map<std::tuple<int, uint32_t, const rck_ksk*>, int> index;
        for (size_t i = 0; i < bt->ops.size(); i++) {
            const BatchOp& o = bt->ops[i];
            const rck_ksk* k = o.kind == RCK_OP_MULT_PLAIN ? nullptr : o.key;
            auto key = std::make_tuple(o.kind, o.limbs, k);
            auto it = index.find(key);
            if (it == index.end()) {
                index[key] = (int)bt->groups.size();
                BatchGroup g;
                g.kind = o.kind;
                g.limbs = o.limbs;
                g.key = k;
                g.ops.push_back((int)i);
                bt->groups.push_back(g);
            } else {
                bt->groups[it->second].ops.push_back((int)i);
            }
        }
        try {
            for (auto& g : bt->groups) check_group(c, g);
        } catch (...) {
            bt->groups.clear();
            throw;
        }
        const bool prof = prof_enabled();
        prof_set(false);
        try {
            for (auto& g : bt->groups)
                if (tune && has_keyswitch(g.kind)) autotune(bt, g);
            assign_lanes(bt);
            // eager run: sizes every lane's scratch (streams, events, workspace) and leaves every
            // output computed
            run_lanes(bt, bt->stream);
            RCK_CHECK_HIP(hipStreamSynchronize(bt->stream));
            auto t0 = std::chrono::steady_clock::now();
            if (bt->exec) (void)hipGraphExecDestroy(bt->exec);  // left by a failed earlier compile
            if (bt->graph) (void)hipGraphDestroy(bt->graph);
            bt->exec = nullptr;
            bt->graph = nullptr;
            for (Context* lc : bt->lane_sc) lc->frozen = true;
            RCK_CHECK_HIP(hipStreamBeginCapture(bt->stream, hipStreamCaptureModeThreadLocal));
            try {
                run_lanes(bt, bt->stream);
            } catch (...) {
                hipGraph_t tmp;
                (void)hipStreamEndCapture(bt->stream, &tmp);
                if (tmp) (void)hipGraphDestroy(tmp);
                throw;
            }
            RCK_CHECK_HIP(hipStreamEndCapture(bt->stream, &bt->graph));
            RCK_CHECK_HIP(hipGraphInstantiate(&bt->exec, bt->graph, nullptr, nullptr, 0));
            for (Context* lc : bt->lane_sc) lc->frozen = false;
            bt->capture_ms =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        } catch (...) {
            for (Context* lc : bt->lane_sc) lc->frozen = false;
            prof_set(prof);
            (void)hipStreamSynchronize(bt->stream);
            release_lanes(bt);
            bt->groups.clear();
            throw;
        }
        prof_set(prof);
        bt->compiled = true;
    });
}

int rck_batch_run(rck_batch* bt, void* stream) {
    return bguard([&] {
        RCK_REQUIRE(bt && bt->compiled, "batch not compiled");
        hipStream_t s = stream ? static_cast<hipStream_t>(stream) : bt->stream;
        RCK_CHECK_HIP(hipGraphLaunch(bt->exec, s));
    });
}

int rck_batch_run_eager(rck_batch* bt) {
    return bguard([&] {
        RCK_REQUIRE(bt && bt->compiled, "batch not compiled");
        for (auto& g : bt->groups) run_group(bt, g, bt->stream, g.shape);
    });
}

int rck_batch_synchronize(rck_batch* bt) {
    return bguard([&] {
        RCK_REQUIRE(bt, "null batch");
        RCK_CHECK_HIP(hipStreamSynchronize(bt->stream));
    });
}

int rck_batch_report(rck_batch* bt, char* buf, size_t len) {
    return bguard([&] {
        RCK_REQUIRE(bt, "null batch");
        std::ostringstream o;
        o << "{\"ops\": " << bt->ops.size() << ", \"groups\": [";
        for (size_t i = 0; i < bt->groups.size(); i++) {
            const auto& g = bt->groups[i];
            o << (i ? ", " : "") << "{\"kind\": " << g.kind << ", \"limbs\": " << g.limbs
              << ", \"ops\": " << g.ops.size() << ", \"eval_limb_group\": " << g.shape.group
              << ", \"digit_splits\": " << g.shape.dsplit << ", \"mac_rows\": " << g.shape.rows
              << ", \"load_schedule\": " << g.shape.prefetch << ", \"tuned_ms\": " << g.tuned_ms
              << ", \"candidates\": [" << g.tune_log << "]}";
        }
        o << "], \"capture_ms\": " << bt->capture_ms << "}";
        std::string r = o.str();
        RCK_REQUIRE(buf && len > r.size(), "report buffer too small");
        std::memcpy(buf, r.c_str(), r.size() + 1);
    });
}

int rck_batch_destroy(rck_batch* bt) {
    return bguard([&] {
        if (!bt) return;
        if (bt->stream) (void)hipStreamSynchronize(bt->stream);
        for (size_t l = 1; l < bt->lane_stream.size(); l++) (void)hipStreamSynchronize(bt->lane_stream[l]);
        if (bt->exec) (void)hipGraphExecDestroy(bt->exec);
        if (bt->graph) (void)hipGraphDestroy(bt->graph);
        release_lanes(bt);
        delete bt->sc;
        if (bt->stream) (void)hipStreamDestroy(bt->stream);
        delete bt;
    });
}

}  // extern "C"
