// Compositions of the kernels into the crate's operations, for one ciphertext or a batch of B
// ciphertexts that share a level (and a key): negacyclic poly product, ct x ct mult, rescale and
// the keyswitch pipeline.  Shared by the per-op C ABI and the batched executor.
#include <algorithm>
#include <cstdio>
#include <map>
#include <mutex>
#include <tuple>

#include "../../include/rnsckks.h"
#include "pipeline.hpp"
#include "wide.hpp"

namespace rck {

enum { kSlotLhs = 10, kSlotRhs = 11, kSlotQuad = 12, kSlotTmp = 13 };

// negacyclic product of `polys` polynomial pairs a[z] * b[z or 0] -> out[z] (NTT resident).  The NTT
// forms live in the context's workspace, or in `scratch` when given (poly_mul_scratch_bytes bytes).
size_t poly_mul_scratch_bytes(const Context& c, long long b_ps, int limbs, int polys) {
    return (size_t)(polys + (b_ps ? polys : 1)) * limbs * c.n * 8;
}

void poly_mul(Context& c, const uint64_t* a, long long a_ps, const uint64_t* b, long long b_ps, uint64_t* out,
              long long out_ps, int limbs, int polys, PrimeMap pm, hipStream_t s, double* scratch) {
    if (c.wide) {
        RCK_REQUIRE(scratch == nullptr, "the caller-provided scratch serves the FP64 arithmetic only");
        return wide_poly_mul(c, a, a_ps, b, b_ps, out, out_ps, limbs, polys, pm, s);
    }
    const long long n = c.n, P = (long long)limbs * n;
    const int bpolys = b_ps ? polys : 1;
    double* t = scratch ? scratch
                        : static_cast<double*>(c.workspace(0, poly_mul_scratch_bytes(c, b_ps, limbs, polys)));
    double* tb = t + (size_t)polys * P;
    launch_fwd(c, kInU64, a, n, a_ps, t, n, P, limbs, polys, pm, s);
    launch_fwd(c, kInU64, b, n, b_ps, tb, n, P, limbs, bpolys, pm, s);
    launch_mul_f64(c, {t, n, P}, {tb, n, b_ps ? P : 0}, {t, n, P}, limbs, polys, pm, s);
    launch_inv(c, t, n, P, kOutU64, out, n, out_ps, limbs, polys, pm, s);
}

// B ciphertext products lhs[b] x rhs[b] (ciphertext strides in_ps) -> quadratic out[b] (stride out_ps).
// mode kMultFoldNtt / kMultFoldRows (two-pass rings): out[b] receives only d2 (coefficient form); d0, d1 stay
// in the returned buffer ([B][3][limbs][N] doubles, d0 at +0 and d1 at +limbs*N of each ciphertext), in NTT
// form or after the inverse row stages, for a keyswitch that folds them in.
double* mult_batch(Context& c, const uint64_t* lhs, const uint64_t* rhs, long long in_ps, int B, int limbs,
                   uint64_t* out, long long out_ps, hipStream_t s, int mode) {
    const bool fold_ntt = mode == kMultFoldNtt, d2_only = mode != kMultFull;
    const long long n = c.n, P = (long long)limbs * n;
    if (c.wide) {
        RCK_REQUIRE(!d2_only, "the folded product serves the FP64 arithmetic only");
        for (int b = 0; b < B; b++) wide_mult(c, lhs + b * in_ps, rhs + b * in_ps, limbs, out + b * out_ps, s);
        return nullptr;
    }
    PrimeMap pm{limbs, (int)c.L, 0};
    double* x = static_cast<double*>(c.workspace(0, (size_t)B * 4 * P * 8));
    double* d = static_cast<double*>(c.workspace(1, (size_t)B * 3 * P * 8));
    if (c.logn <= 11) {
        RCK_REQUIRE(!d2_only, "the folded product needs the two-pass NTT");
        for (int b = 0; b < B; b++) {
            launch_fwd(c, kInU64, lhs + b * in_ps, n, P, x, n, P, limbs, 2, pm, s);
            launch_fwd(c, kInU64, rhs + b * in_ps, n, P, x + 2 * P, n, P, limbs, 2, pm, s);
            launch_tensor(c, x, d, limbs, pm, s);
            launch_inv(c, d, n, P, kOutU64, out + b * out_ps, n, P, limbs, 3, pm, s);
        }
        return d;
    }
    // forward column pass: lhs[b] -> xl[b] = x[b][0..1], rhs[b] -> xr[b] = x[B + b][0..1]
    double* xl = x;
    double* xr = x + (long long)B * 2 * P;
    if (B == 1 || in_ps == 2 * P) {
        // one launch for both operands: xr follows xl, so output poly z is x + z * P
        launch_fwd_cols(c, kInU64, lhs, n, P, xl, n, P, limbs, 4 * B, pm, s, rhs, 2 * B);
    } else {
        for (int b = 0; b < B; b++) {
            launch_fwd_cols(c, kInU64, lhs + b * in_ps, n, P, xl + b * 2 * P, n, P, limbs, 2, pm, s);
            launch_fwd_cols(c, kInU64, rhs + b * in_ps, n, P, xr + b * 2 * P, n, P, limbs, 2, pm, s);
        }
    }
    launch_mult_rows(c, xl, xr, d, limbs, B, pm, s, fold_ntt);
    if (d2_only) {
        // inverse column pass of d2 only: d[b][2] -> out[b]
        launch_inv_cols(c, d + 2 * P, n, 3 * P, kOutU64, out, n, out_ps, limbs, B, pm, s);
        return d;
    }
    // inverse column pass of every d[b][k] -> out[b][k]
    if (B == 1 || out_ps == 3 * P) {
        launch_inv_cols(c, d, n, P, kOutU64, out, n, P, limbs, 3 * B, pm, s);
    } else {
        for (int b = 0; b < B; b++)
            launch_inv_cols(c, d + b * 3 * P, n, P, kOutU64, out + b * out_ps, n, P, limbs, 3, pm, s);
    }
    return d;
}

void mult_impl(Context& c, const uint64_t* lhs, const uint64_t* rhs, int limbs, uint64_t* out, hipStream_t s) {
    mult_batch(c, lhs, rhs, 0, 1, limbs, out, 0, s);
}

// Are ptr(op_0) .. ptr(op_{B-1}) equally spaced?  Returns the stride (elements) or -1.
template <typename P>
static long long common_stride(const std::vector<P>& ptrs) {
    if (ptrs.size() < 2) return 0;
    const long long st = (long long)(ptrs[1] - ptrs[0]);
    if (st <= 0) return -1;
    for (size_t i = 2; i < ptrs.size(); i++)
        if ((long long)(ptrs[i] - ptrs[i - 1]) != st) return -1;
    return st;
}


void rotate_batch(Context& c, const std::vector<const uint64_t*>& as, int limbs, const KeyDev* key,
                  const std::vector<uint64_t*>& outs, hipStream_t s, KsShape shape) {
    check_ptr(key, "key");
    const int B = (int)as.size();
    const long long n = c.n, P = (long long)limbs * n;
    const long long st = B > 1 ? common_stride(as) : 2 * P;
    if (c.logn >= 12 && key->aux > 0 && st > 0 && !c.wide) {
        // rotate = (sigma(c0), 0) + keyswitch(sigma(c1)): the decomposition reads c1 and the mod-down
        // epilogue reads c0 through the automorphism, so no rotated copy is written
        const uint64_t ginv = galois_inverse(c, key->galois);
        keyswitch_batch(c, as[0] + P, st, B, limbs, key, as[0], nullptr, st, outs, s, shape, nullptr, 0, ginv);
        return;
    }
    uint64_t* rot = static_cast<uint64_t*>(c.workspace(kSlotTmp, (size_t)B * 2 * P * 8));
    PrimeMap pm{limbs, (int)c.L, 0};
    if (B > 1 && st == 2 * P) {  // contiguous inputs: one launch for all B
        launch_galois(c, {as[0], n, P}, {rot, n, P}, limbs, 2 * B, key->galois, pm, s);
    } else {
        for (int i = 0; i < B; i++)
            launch_galois(c, {as[i], n, P}, {rot + i * 2 * P, n, P}, limbs, 2, key->galois, pm, s);
    }
    keyswitch_batch(c, rot + P, 2 * P, B, limbs, key, rot, nullptr, 2 * P, outs, s, shape);
}

// the relinearization fold needs the digit-NTT method on a hybrid key at a two-pass ring
static bool mult_relin_folds(const Context& c, const KeyDev* key, int limbs, int B, const KsShape& shape) {
    if (c.logn < 12 || !key || key->aux == 0 || key->data == nullptr || c.wide) return false;
    const int D = (int)((c.qbits[limbs] + 29) / 30);
    return !ks_use_exact(c, key, limbs, limbs + (int)key->aux, D, B, shape);
}

void mult_relin_batch(Context& c, const uint64_t* lhs, const uint64_t* rhs, long long in_ps, int B, int limbs,
                      const KeyDev* key, const std::vector<uint64_t*>& outs, hipStream_t s, KsShape shape) {
    const long long P = (long long)limbs * c.n;
    if (mult_relin_folds(c, key, limbs, B, shape)) {
        // relinearize = (d0, d1) + keyswitch(d2): d0, d1 are added to the Q limbs of the keyswitch
        // accumulators in NTT form, times P, so mod-down by P returns keyswitch + (d0, d1) exactly
        uint64_t* c2 = static_cast<uint64_t*>(c.workspace(kSlotQuad, (size_t)B * P * 8));
        const double* dn = mult_batch(c, lhs, rhs, in_ps, B, limbs, c2, P, s, kMultFoldNtt);
        keyswitch_batch(c, c2, P, B, limbs, key, nullptr, nullptr, 0, outs, s, shape, dn, 3 * P);
        return;
    }
    if (c.logn >= 12 && key && key->aux > 0 && key->mdata != nullptr && !c.wide) {
        // the exact convolution: d0, d1 leave the product after the inverse row stages, and their inverse
        // column pass runs inside the keyswitch's recombination, which adds them times P before mod-down
        uint64_t* c2 = static_cast<uint64_t*>(c.workspace(kSlotQuad, (size_t)B * P * 8));
        const double* dr = mult_batch(c, lhs, rhs, in_ps, B, limbs, c2, P, s, kMultFoldRows);
        keyswitch_batch(c, c2, P, B, limbs, key, nullptr, nullptr, 0, outs, s, shape, nullptr, 0, 0, dr, 3 * P);
        return;
    }
    uint64_t* quad = static_cast<uint64_t*>(c.workspace(kSlotQuad, (size_t)B * 3 * P * 8));
    mult_batch(c, lhs, rhs, in_ps, B, limbs, quad, 3 * P, s);
    keyswitch_batch(c, quad + 2 * P, 3 * P, B, limbs, key, quad, quad + P, 3 * P, outs, s, shape);
}

void rescale_impl(Context& c, const uint64_t* in, int limbs, int polys, uint64_t* out, hipStream_t s) {
    const long long n = c.n;
    launch_rescale(c, {in, n, (long long)limbs * n}, {out, n, (long long)(limbs - 1) * n}, limbs, polys, s);
}

void run_op_group(Context& c, int kind, int l, const KeyDev* key, const std::vector<const uint64_t*>& as,
                  const std::vector<const uint64_t*>& bs, const std::vector<uint64_t*>& outs, hipStream_t s,
                  KsShape shape) {
    const int B = (int)as.size();
    const long long n = c.n, P = (long long)l * n;
    switch (kind) {
        case RCK_OP_MULT_RELIN:
        case RCK_OP_MULT_RELIN_RESCALE: {
            const long long sa = common_stride(as), sb = common_stride(bs);
            const uint64_t *lhs = as[0], *rhs = bs[0];
            long long st = sa;
            if (B > 1 && !(sa == sb && sa >= 2 * P)) {  // stage both operands contiguously
                uint64_t* lb = static_cast<uint64_t*>(c.workspace(kSlotLhs, (size_t)B * 2 * P * 8));
                uint64_t* rb = static_cast<uint64_t*>(c.workspace(kSlotRhs, (size_t)B * 2 * P * 8));
                for (int i = 0; i < B; i++) {
                    RCK_CHECK_HIP(hipMemcpyAsync(lb + i * 2 * P, as[i], 2 * P * 8, hipMemcpyDeviceToDevice, s));
                    RCK_CHECK_HIP(hipMemcpyAsync(rb + i * 2 * P, bs[i], 2 * P * 8, hipMemcpyDeviceToDevice, s));
                }
                lhs = lb;
                rhs = rb;
                st = 2 * P;
            }
            if (kind == RCK_OP_MULT_RELIN) {
                mult_relin_batch(c, lhs, rhs, st, B, l, key, outs, s, shape);
            } else {
                uint64_t* rel = static_cast<uint64_t*>(c.workspace(kSlotTmp, (size_t)B * 2 * P * 8));
                std::vector<uint64_t*> rels;
                for (int i = 0; i < B; i++) rels.push_back(rel + i * 2 * P);
                mult_relin_batch(c, lhs, rhs, st, B, l, key, rels, s, shape);
                for (int i = 0; i < B; i++) rescale_impl(c, rels[i], l, 2, outs[i], s);
            }
            break;
        }
        case RCK_OP_ROTATE: {
            rotate_batch(c, as, l, key, outs, s, shape);
            break;
        }
        case RCK_OP_MULT_PLAIN: {
            PrimeMap pm{l, (int)c.L, 0};
            for (int i = 0; i < B; i++) poly_mul(c, as[i], P, bs[i], 0, outs[i], P, l, 2, pm, s);
            break;
        }
        default:
            RCK_REQUIRE(false, "unknown batch op kind");
    }
}

// Digit-MAC launch shape for B ciphertexts at `limbs` limbs: explicit tuning values when set,
// otherwise a heuristic that keeps >= ~2048 MAC workgroups in flight.
namespace {
std::mutex g_tune_mu;
// key: (logn, limbs, eval limbs, digits, batch, method: 0 digit NTT, 1 exact convolution)
std::map<std::tuple<int, int, int, int, int, int>, std::pair<KsShape, float>> g_tuned;
}  // namespace

void ks_tuned_store(const Context& c, int limbs, int E, int D, int B, KsShape sh, float ms, bool exact) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[std::make_tuple((int)c.logn, limbs, E, D, B, exact ? 1 : 0)] = {sh, ms};
}

bool ks_tuned_find(const Context& c, int limbs, int E, int D, int B, KsShape* sh, float* ms, bool exact) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(std::make_tuple((int)c.logn, limbs, E, D, B, exact ? 1 : 0));
    if (it == g_tuned.end()) return false;
    *sh = it->second.first;
    if (ms) *ms = it->second.second;
    return true;
}

std::string ks_tuned_export() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    std::string out;
    char line[160];
    for (const auto& kv : g_tuned) {
        const KsShape& sh = kv.second.first;
        std::snprintf(line, sizeof line, "%d %d %d %d %d %d %d %d %d %d %.6f\n", std::get<0>(kv.first),
                      std::get<1>(kv.first), std::get<2>(kv.first), std::get<3>(kv.first), std::get<4>(kv.first),
                      sh.group, sh.dsplit, sh.rows, sh.prefetch, sh.width, (double)kv.second.second);
        out += line;
    }
    return out;
}

int ks_tuned_import(const std::string& text) {
    // parse everything first: a malformed table changes nothing
    std::vector<std::pair<std::tuple<int, int, int, int, int, int>, std::pair<KsShape, float>>> rows;
    size_t pos = 0;
    while (pos < text.size()) {
        size_t end = text.find('\n', pos);
        if (end == std::string::npos) end = text.size();
        const std::string line = text.substr(pos, end - pos);
        pos = end + 1;
        if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
        int v[10];
        double ms = 0;
        char extra = 0;
        const int got = std::sscanf(line.c_str(), "%d %d %d %d %d %d %d %d %d %d %lf %c", &v[0], &v[1], &v[2],
                                    &v[3], &v[4], &v[5], &v[6], &v[7], &v[8], &v[9], &ms, &extra);
        RCK_REQUIRE(got == 11, "malformed tuning table line: " + line);
        RCK_REQUIRE(v[1] >= 1 && v[2] >= v[1] && v[3] >= 1 && v[4] >= 1, "tuning table entry out of range: " + line);
        const bool w16 = v[9] == 16 && (v[7] == 4 || v[7] == 16) && v[8] >= 0 && (v[8] <= 4 || ((v[8] == 5 || v[8] == 6) && v[7] == 4));
        const bool w8 = v[9] == 8 && (v[7] == 8 || v[7] == 16) && v[8] == 0;
        // exact-convolution entries: load schedule kExactShape, rows = slots per wave of the multiply-accumulate
        const bool ex = v[8] == kExactShape && v[5] == v[2] && v[6] == 1 && (v[7] == 1 || v[7] == 2 || v[7] == 4) &&
                        v[9] == 16;
        RCK_REQUIRE(v[5] >= 1 && v[5] <= v[2] && v[6] >= 1 && v[6] <= 8 && (w16 || w8 || ex),
                    "tuning table launch shape out of range: " + line);
        rows.push_back({std::make_tuple(v[0], v[1], v[2], v[3], v[4], ex ? 1 : 0),
                        {KsShape{v[5], v[6], v[7], v[8], v[9]}, (float)ms}});
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (const auto& r : rows) g_tuned[r.first] = r.second;
    return (int)rows.size();
}

bool ks_use_exact(const Context& c, const KeyDev* key, int limbs, int E, int D, int B, const KsShape& shape) {
    if (!key->mdata) return false;
    if (!key->data) return true;
    if (shape.prefetch == kExactShape) return true;
    if (shape.prefetch >= 0 || shape.group > 0 || shape.rows > 0) return false;  // an explicit digit-NTT shape
    KsShape se, sd;
    float me = 0, md = 0;
    if (ks_tuned_find(c, limbs, E, D, B, &se, &me, true) && ks_tuned_find(c, limbs, E, D, B, &sd, &md, false))
        return me < md;
    return B >= 4;
}

KsShape ks_auto_shape(const Context& c, int limbs, int E, int D, int B, bool exact) {
    KsShape r;
    if (ks_tuned_find(c, limbs, E, D, B, &r, nullptr, exact)) return r;
    if (exact) {  // two slots per wave of the matrix-core multiply-accumulate
        r.group = E;
        r.dsplit = 1;
        r.rows = 2;
        r.prefetch = kExactShape;
        r.width = 16;
        return r;
    }
    r.group = c.tune_group > 0 ? c.tune_group : E;
    r.prefetch = 0;
    if (c.logn <= 11) {
        r.rows = 16;
        r.dsplit = 1;
        return r;
    }
    // the streamed MAC: four ciphertexts per workgroup from four up, four digit ranges of one
    // ciphertext below; one digit range, so its epilogue also runs the inverse rows (and the fold)
    r.rows = 4;
    r.prefetch = B >= 4 ? 5 : 6;
    r.dsplit = c.tune_dsplit > 0 ? std::min(c.tune_dsplit, std::min(8, D)) : 1;
    return r;
}

std::vector<KsShape> ks_candidates(const Context& c, int E, int D, int B, bool exact) {
    std::vector<KsShape> out;
    if (exact) {
        for (int spw : {1, 2, 4}) out.push_back({E, 1, spw, kExactShape, 16});
        return out;
    }
    if (c.logn <= 11) {
        out.push_back({E, 1, 4, 0});
        return out;
    }
    std::vector<int> gs = {E, (E + 1) / 2, (E + 3) / 4, 4, 1};
    std::sort(gs.begin(), gs.end());
    gs.erase(std::unique(gs.begin(), gs.end()), gs.end());
    for (int g : gs)
        for (int ds : {1, 2, 4}) {
            if (ds > D) continue;
            for (int rows : {4, 16})
                for (int pf : {0, 1, 2, 3, 4}) out.push_back({g, ds, rows, pf, 16});
            if (B >= 2) out.push_back({g, ds, 4, 5, 16});  // four ciphertexts per workgroup, LDS-streamed rows
            out.push_back({g, ds, 4, 6, 16});              // four digit ranges of one ciphertext per workgroup
            if (g == E && ds <= 2)
                for (int rows : {8, 16}) out.push_back({g, ds, rows, 0, 8});
        }
    return out;
}

// Keyswitch of B polynomials `poly` (limbs Q-limbs each, poly stride in_ps) with one key; adds
// add0 / add1 (nullable; element offsets relative to each ciphertext, stride add_ps) and writes
// out[b] = [c0][c1] at stride out_ps.  Mirrors keyswitch_rns (keyswitching.rs:130-157) and
// keyswitch_rns_hybrid (:159-194).
void keyswitch_batch(Context& c, const uint64_t* poly, long long in_ps, int B, int limbs, const KeyDev* key,
                     const uint64_t* add0, const uint64_t* add1, long long add_ps, const std::vector<uint64_t*>& outs,
                     hipStream_t s, KsShape shape, const double* addn, long long addn_ps, unsigned long long ginv,
                     const double* addr, long long addr_ps) {
    RCK_REQUIRE((int)outs.size() == B, "one output per ciphertext");
    check_ptr(key, "key");
    const long long n = c.n, P = (long long)limbs * n;
    const uint32_t active = (c.qbits[limbs] + 29) / 30;
    int levels;
    if (key->aux == 0) {
        RCK_REQUIRE((int)key->level == limbs - 1, "keyswitch key level must match the active modulus chain");
        RCK_REQUIRE(key->digits >= active, "keyswitch key must cover the active modulus chain");
        levels = key->digits;  // keyswitching.rs:138 decomposes into ksk.ksk.len() digits
    } else {
        RCK_REQUIRE((int)key->level == limbs - 1, "keyswitch key level must match the active Q chain");
        RCK_REQUIRE(key->aux == c.K, "eval-key context must match the keyswitch key auxiliary basis");
        RCK_REQUIRE(key->digits >= active, "keyswitch key must cover the active modulus chain");
        RCK_REQUIRE(c.qbits[limbs] > 30, "hybrid keyswitching needs a ciphertext modulus above 2^30");
        levels = active;  // keyswitching.rs:170
    }
    const int naux = key->aux;
    if (c.wide) {
        RCK_REQUIRE(addn == nullptr && ginv == 0, "folded addends and the fused automorphism serve the FP64 arithmetic only");
        for (int b = 0; b < B; b++)
            wide_keyswitch(c, poly + b * in_ps, limbs, key, add0 ? add0 + b * add_ps : nullptr,
                           add1 ? add1 + b * add_ps : nullptr, outs[b], s);
        return;
    }
    RCK_REQUIRE(key->data != nullptr || key->mdata != nullptr,
                "keyswitch key was created for another arithmetic of this context");
    if (addr != nullptr) {
        RCK_REQUIRE(add0 == nullptr && add1 == nullptr && addn == nullptr && ginv == 0 && key->mdata != nullptr &&
                        naux > 0 && c.logn >= 12,
                    "row-form addends serve the exact-convolution keyswitch only");
        const KsShape sh = shape.prefetch == kExactShape ? shape : ks_auto_shape(c, limbs, limbs + naux, levels, B, true);
        keyswitch_exact(c, poly, in_ps, B, limbs, key, levels, nullptr, nullptr, 0, outs, s, 0, sh.rows, addr, addr_ps);
        return;
    }
    if (addn == nullptr && ks_use_exact(c, key, limbs, limbs + naux, levels, B, shape)) {
        const KsShape sh = shape.prefetch == kExactShape ? shape : ks_auto_shape(c, limbs, limbs + naux, levels, B, true);
        keyswitch_exact(c, poly, in_ps, B, limbs, key, levels, add0, add1, add_ps, outs, s, ginv, sh.rows);
        return;
    }
    RCK_REQUIRE(key->data != nullptr, "folded addends need a key with its digit-NTT form");
    RCK_REQUIRE(addn == nullptr || (naux > 0 && c.logn >= 12), "folded addends need hybrid keys and the two-pass NTT");
    RCK_REQUIRE(ginv == 0 || (naux > 0 && c.logn >= 12), "the fused automorphism needs hybrid keys and the two-pass NTT");
    const int E = limbs + naux;
    const long long EN = (long long)E * n;
    const KsShape aut = ks_auto_shape(c, limbs, E, levels, B);
    const int group = shape.group > 0 ? std::min(shape.group, E) : aut.group;
    const int dsplit = c.logn <= 11 ? 1 : (shape.dsplit > 0 ? std::min(shape.dsplit, std::min(8, levels)) : aut.dsplit);
    const int rows = shape.rows > 0 ? shape.rows : aut.rows;
    const int prefetch = shape.prefetch >= 0 ? shape.prefetch : aut.prefetch;
    int32_t* digits = static_cast<int32_t*>(c.workspace(5, (size_t)B * levels * n * 4));
    double* part = static_cast<double*>(c.workspace(6, (size_t)dsplit * B * 2 * EN * 8));
    launch_decompose(c, poly, limbs, levels, digits, s, B, in_ps, ginv);
    const int width = shape.group > 0 || shape.rows > 0 ? shape.width : aut.width;
    // the streamed MAC with one digit range also runs the accumulators' inverse row stages (and the fold)
    const bool fuse_inv = naux && c.logn >= 12 && (prefetch == 5 || prefetch == 6) && rows == 4 && dsplit == 1;
    MacFold mf{addn, addn_ps, P, limbs, c.d_pmodq};
    launch_digit_mac(c, digits, B, levels, key->data, limbs, naux, part, dsplit, group, rows, prefetch,
                     rows == 4 ? 16 : width, s, fuse_inv ? &mf : nullptr);
    PrimeMap pm{limbs, (int)c.L, 0};
    const long long sstride = (long long)B * 2 * EN;
    if (naux && c.logn >= 12) {
        // inverse rows of every limb; inverse columns of the aux limbs to coefficient form; the Q-limb
        // inverse columns run fused with mod-down and the adds
        if (!fuse_inv) launch_inv_rows(c, part, n, EN, E, 2 * B, pm, s, dsplit, sstride, addn, addn_ps, P, limbs);
        uint64_t* accP = static_cast<uint64_t*>(c.workspace(4, (size_t)B * 2 * naux * n * 8));
        launch_inv_cols(c, part + (long long)limbs * n, n, EN, kOutU64, accP, n, (long long)naux * n, naux, 2 * B,
                        PrimeMap{limbs, (int)c.L, limbs}, s);
        launch_inv_cols_moddown(c, part, limbs, B, accP, add0, add1, add_ps, outs, s, ginv);
    } else if (naux) {
        uint64_t* accu = static_cast<uint64_t*>(c.workspace(4, (size_t)B * 2 * EN * 8));
        launch_inv(c, part, n, EN, kOutU64, accu, n, EN, E, 2 * B, pm, s, dsplit, sstride);
        for (int b = 0; b < B; b++)
            launch_mod_down(c, accu + b * 2 * EN, limbs, add0 ? add0 + b * add_ps : nullptr,
                            add1 ? add1 + b * add_ps : nullptr, outs[b], P, s);
    } else {
        uint64_t* accu = static_cast<uint64_t*>(c.workspace(4, (size_t)B * 2 * P * 8));
        launch_inv(c, part, n, P, kOutU64, accu, n, P, limbs, 2 * B, pm, s, dsplit, sstride);
        for (int b = 0; b < B; b++) {
            uint64_t* ab = accu + b * 2 * P;
            uint64_t* ob = outs[b];
            if (add0)
                launch_addsub(c, {add0 + b * add_ps, n, 0}, {ab, n, 0}, {ob, n, 0}, limbs, 1, 0, pm, s);
            else
                RCK_CHECK_HIP(hipMemcpyAsync(ob, ab, P * 8, hipMemcpyDeviceToDevice, s));
            if (add1)
                launch_addsub(c, {add1 + b * add_ps, n, 0}, {ab + P, n, 0}, {ob + P, n, 0}, limbs, 1, 0, pm, s);
            else
                RCK_CHECK_HIP(hipMemcpyAsync(ob + P, ab + P, P * 8, hipMemcpyDeviceToDevice, s));
        }
    }
}

void keyswitch_impl(Context& c, const uint64_t* poly, int limbs, const KeyDev* key, const uint64_t* add0,
                           const uint64_t* add1, uint64_t* out, hipStream_t s) {
    keyswitch_batch(c, poly, 0, 1, limbs, key, add0, add1, 0, {out}, s);
}

}  // namespace rck
