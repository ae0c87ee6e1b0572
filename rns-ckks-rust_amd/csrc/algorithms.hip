// Paterson-Stockmeyer evaluation (src/algorithms/paterson_stockmeyer.rs:28-350): the crate's
// algorithm step for step — baby powers x^1..x^m at the input scale, blocks of m terms at their
// target levels, the giant-step Horner loop — compiled into a straight-line program over
// ciphertext registers and run with the per-op kernels.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "../../include/rnsckks.h"
#include "algorithms.hpp"
#include "capi_types.hpp"

namespace rck {

namespace {

uint32_t floor_log2_u64(uint64_t q) { return 63 - __builtin_clzll(q); }

struct Consts {
    uint64_t k[kMaxQ];
};

Consts consts_of(const std::vector<uint64_t>& k) {
    Consts c{};
    for (size_t i = 0; i < k.size() && i < (size_t)kMaxQ; i++) c.k[i] = k[i];
    return c;
}

// out[z][l] = in[z][l] * k_l mod q_l (z = c0, c1); in may hold more limbs than out (truncation)
__global__ void k_scalar_mul(const uint64_t* __restrict__ in, long long in_ps, uint64_t* __restrict__ out,
                             long long out_ps, int n, Consts kc, const DevPrime* __restrict__ primes) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int l = blockIdx.y, z = blockIdx.z;
    if (i >= n) return;
    const DevPrime P = primes[l];
    const double x = load_centred(in[z * in_ps + (long long)l * n + i], P.q);
    const double w = load_centred(kc.k[l], P.q);
    out[z * out_ps + (long long)l * n + i] = (uint64_t)canon(mulmod(x, w, P.q, P.qinv), P.q, P.qinv);
}

// out = in, plus k_l on coefficient 0 of c0
__global__ void k_add_const0(const uint64_t* __restrict__ in, long long in_ps, uint64_t* __restrict__ out,
                             long long out_ps, int n, Consts kc, const DevPrime* __restrict__ primes) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int l = blockIdx.y, z = blockIdx.z;
    if (i >= n) return;
    uint64_t v = in[z * in_ps + (long long)l * n + i];
    if (z == 0 && i == 0) {
        const uint64_t q = (uint64_t)primes[l].q;
        v += kc.k[l];
        if (v >= q) v -= q;
    }
    out[z * out_ps + (long long)l * n + i] = v;
}

__global__ void k_zero(uint64_t* __restrict__ out, long long count) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = 0;
}

// ---------------------------------------------------------------- program builder
struct Ct {
    int reg;
    int limbs;
    uint32_t scale;
};

struct Builder {
    const Context& c;
    PsProgram p;
    const std::vector<bool>& has_key;

    Builder(const Context& c_, const std::vector<bool>& hk) : c(c_), has_key(hk) {}

    Ct reg(int limbs, uint32_t scale) {
        p.reg_limbs.push_back(limbs);
        p.reg_scale.push_back(scale);
        return Ct{(int)p.reg_limbs.size() - 1, limbs, scale};
    }
    void emit(int kind, const Ct& dst, int a, int b, std::vector<uint64_t> k = {}, int key_level = -1) {
        PsInstr ins;
        ins.kind = kind;
        ins.dst = dst.reg;
        ins.a = a;
        ins.b = b;
        ins.limbs = dst.limbs;
        ins.key_level = key_level;
        ins.k = std::move(k);
        p.code.push_back(std::move(ins));
    }
    // constant_ciphertext (:297-312): zero ciphertext plus a constant term
    Ct constant(int level, uint32_t scale, double value) {
        Ct z = reg(level + 1, scale);
        emit(kPsZero, z, -1, -1);
        return add_constant(z, value);
    }
    // add_constant_term (:314-323)
    Ct add_constant(const Ct& x, double value) {
        if (value == 0.0) return x;
        Ct d = reg(x.limbs, x.scale);
        emit(kPsAddConst, d, x.reg, -1, const_residues(c, value, x.scale, x.limbs));
        return d;
    }
    // align_ciphertext_to_level (:218-225): keep the first target_level + 1 limbs
    Ct truncate(const Ct& x, int target_level) {
        RCK_REQUIRE(target_level <= x.limbs - 1, "target level cannot exceed ciphertext level");
        if (target_level + 1 == x.limbs) return x;
        Ct d = reg(target_level + 1, x.scale);
        emit(kPsTruncate, d, x.reg, -1);
        return d;
    }
    // lift_ciphertext_scale (:227-251): multiply by the constant 1 encoded at the scale difference
    Ct lift(const Ct& x, uint32_t target) {
        RCK_REQUIRE(x.scale <= target, "cannot lower ciphertext scale with a plaintext-one lift");
        if (x.scale == target) return x;
        Ct d = reg(x.limbs, target);
        emit(kPsScalarMul, d, x.reg, -1, const_residues(c, 1.0, target - x.scale, x.limbs));
        return d;
    }
    // mult_plain_slots_real_at_scale (:253-273) of a constant slot vector, on the first `limbs` limbs
    Ct mul_const(const Ct& x, int limbs, double value, uint32_t plain_scale) {
        Ct d = reg(limbs, x.scale + plain_scale);
        emit(kPsScalarMul, d, x.reg, -1, const_residues(c, value, plain_scale, limbs));
        return d;
    }
    Ct add(const Ct& a, const Ct& b) {
        RCK_REQUIRE(a.scale == b.scale, "ciphertexts must share the same scale to add");
        RCK_REQUIRE(a.limbs == b.limbs, "ciphertexts must be at the same level to add");
        Ct d = reg(a.limbs, a.scale);
        emit(kPsAdd, d, a.reg, b.reg);
        return d;
    }
    // RnsCkksContext::mult_relinearized (context/ops.rs:124-133) with relin_keys[level]
    Ct mult_relin(const Ct& a, const Ct& b) {
        RCK_REQUIRE(a.limbs == b.limbs, "ciphertexts must be at the same level to multiply");
        RCK_REQUIRE(a.scale == b.scale, "ciphertexts must share the same scale to multiply");
        const int level = a.limbs - 1;
        RCK_REQUIRE(level >= 1, "rescale requires at least two levels");
        RCK_REQUIRE(level < (int)has_key.size() && has_key[level],
                    "missing Paterson-Stockmeyer relinearization key at level " + std::to_string(level));
        const uint32_t s = a.scale + b.scale, drop = floor_log2_u64(c.hp[level].q);
        Ct d = reg(a.limbs - 1, s > drop ? s - drop : 0);
        emit(kPsMultRelinRescale, d, a.reg, b.reg, {}, level);
        return d;
    }
};

void assign_slots(PsProgram& p) {
    const int R = (int)p.reg_limbs.size();
    std::vector<int> last(R, -1);
    for (int i = 0; i < (int)p.code.size(); i++) {
        const PsInstr& ins = p.code[i];
        if (ins.a >= 0) last[ins.a] = i;
        if (ins.b >= 0) last[ins.b] = i;
        last[ins.dst] = std::max(last[ins.dst], i);
    }
    last[p.result] = (int)p.code.size();
    p.slot_of.assign(R, -1);
    std::vector<int> free_slots;
    for (int i = 0; i < (int)p.code.size(); i++) {
        const PsInstr& ins = p.code[i];
        const int need = p.reg_limbs[ins.dst];
        int best = -1;
        for (int s : free_slots)
            if (p.slot_limbs[s] >= need && (best < 0 || p.slot_limbs[s] < p.slot_limbs[best])) best = s;
        if (best < 0) {
            best = (int)p.slot_limbs.size();
            p.slot_limbs.push_back(need);
        } else {
            free_slots.erase(std::find(free_slots.begin(), free_slots.end(), best));
        }
        p.slot_of[ins.dst] = best;
        // operands (and a dst nobody reads) whose last use is this instruction free their slots afterwards
        for (int r : {ins.a, ins.b, ins.dst})
            if (r > 0 && last[r] == i && p.slot_of[r] >= 0 &&
                std::find(free_slots.begin(), free_slots.end(), p.slot_of[r]) == free_slots.end())
                free_slots.push_back(p.slot_of[r]);
    }
}

}  // namespace

std::vector<uint64_t> const_residues(const Context& c, double value, uint32_t scale_bits, int limbs) {
    // encode_integer_coeffs coefficient 0 of a constant vector: round(value * 2^s) as i128
    const double x = std::round(std::fma(value, std::ldexp(1.0, (int)scale_bits), 0.0));
    RCK_REQUIRE(std::isfinite(x) && std::fabs(x) < 1.7e38, "encoded constant out of range");
    const __int128 v = (__int128)x;
    std::vector<uint64_t> k(limbs);
    for (int l = 0; l < limbs; l++) {
        const __int128 q = (__int128)c.hp[l].q;
        __int128 r = v % q;
        if (r < 0) r += q;
        k[l] = (uint64_t)r;
    }
    return k;
}

size_t PsProgram::footprint(uint32_t n) const {
    size_t limbs = 0;
    for (int l : slot_limbs) limbs += (size_t)l;
    return limbs * 2 * n * 8;
}

int PsProgram::mults() const {
    int m = 0;
    for (const auto& i : code) m += i.kind == kPsMultRelinRescale;
    return m;
}

int ps_block_size(int term_count) {
    RCK_REQUIRE(term_count > 0, "term_count must be positive");
    int block = 1;
    while (block * block < term_count) block++;
    return block;
}

std::vector<int> ps_required_levels(int ciphertext_level, int term_count) {
    RCK_REQUIRE(term_count > 0, "term_count must be positive");
    std::vector<int> out;
    if (term_count == 1) return out;
    const int block = ps_block_size(term_count);
    const int blocks = (term_count + block - 1) / block;
    const char* msg = "polynomial degree is too large for the available CKKS level budget";
    RCK_REQUIRE(ciphertext_level >= block - 1, msg);
    const int top = ciphertext_level - (block - 1);
    RCK_REQUIRE(top >= blocks - 1, msg);
    const int final_level = top - (blocks - 1);
    for (int l = ciphertext_level; l >= final_level + 1; l--) out.push_back(l);
    return out;
}

PsProgram ps_compile(const Context& c, int limbs, uint32_t scale_bits, const double* coeffs, int count,
                     const std::vector<bool>& has_key) {
    RCK_REQUIRE(count > 0 && coeffs != nullptr, "polynomial must contain at least one coefficient");
    RCK_REQUIRE(limbs >= 1 && limbs <= (int)c.L, "ciphertext level context must exist");
    Builder B(c, has_key);
    const Ct x = B.reg(limbs, scale_bits);  // register 0: the input
    const int level = limbs - 1;
    const int m = ps_block_size(count);
    const int blocks = (count + m - 1) / m;
    if (count == 1) {
        B.p.result = B.constant(level, scale_bits, coeffs[0]).reg;
        assign_slots(B.p);
        return B.p;
    }
    // paterson_stockmeyer_working_scale_bits (:275-295)
    uint32_t working = UINT32_MAX;
    for (int l : ps_required_levels(level, count)) working = std::min(working, floor_log2_u64(c.hp[l].q));
    RCK_REQUIRE(working >= scale_bits, "Paterson-Stockmeyer working scale must dominate the input ciphertext scale");
    const uint32_t coeff_scale = working - scale_bits;
    // precompute_baby_powers (:146-174) at the input scale
    std::vector<Ct> powers;
    const Ct base = B.lift(x, scale_bits);
    powers.push_back(base);
    for (int e = 2; e <= m; e++) {
        const Ct prev = powers.back();
        const Ct aligned_prev = B.lift(prev, scale_bits);
        const Ct aligned_base = B.truncate(base, prev.limbs - 1);
        const Ct next = B.mult_relin(aligned_prev, aligned_base);
        powers.push_back(B.lift(next, scale_bits));
    }
    const Ct giant = powers.back();
    RCK_REQUIRE(giant.limbs >= blocks, "polynomial degree is too large for the available CKKS level budget");
    const int top = giant.limbs - 1;
    // evaluate_block (:176-216) for every block at its target level
    std::vector<Ct> bl;
    for (int bi = 0; bi < blocks; bi++) {
        const int target = top - (blocks - 1 - bi);
        Ct acc = B.constant(target, working, 0.0);
        for (int off = 0; off < m; off++) {
            const int idx = bi * m + off;
            if (idx >= count) break;
            const double coef = coeffs[idx];
            if (coef == 0.0) continue;
            if (off == 0) {
                acc = B.add_constant(acc, coef);
                continue;
            }
            const Ct& pw = powers[off - 1];
            RCK_REQUIRE(target <= pw.limbs - 1, "target level cannot exceed ciphertext level");
            // align_block_term_to_level (truncate, lift to its own scale) fused into the constant product
            const Ct term = B.mul_const(pw, target + 1, coef, coeff_scale);
            acc = B.add(acc, term);
        }
        bl.push_back(acc);
    }
    // giant-step Horner loop (:87-99)
    Ct result = bl.back();
    for (int bi = blocks - 2; bi >= 0; bi--) {
        const Ct aligned_giant = B.lift(B.truncate(giant, result.limbs - 1), result.scale);
        const Ct mult = B.mult_relin(result, aligned_giant);
        const Ct lifted = B.lift(mult, result.scale);
        result = B.add(lifted, bl[bi]);
    }
    B.p.result = result.reg;
    assign_slots(B.p);
    return B.p;
}

void ps_exec(Context& c, const PsProgram& p, const PsInstr& ins, const std::vector<uint64_t*>& regs,
             const std::vector<const KeyDev*>& keys, hipStream_t s) {
    const long long n = c.n;
    const int l = ins.limbs;
    uint64_t* dst = regs[ins.dst];
    const long long dps = (long long)l * n;
    const dim3 grid((c.n + 255) / 256, l, 2);
    switch (ins.kind) {
        case kPsMultRelinRescale: {
            const int li = p.reg_limbs[ins.a];
            run_op_group(c, RCK_OP_MULT_RELIN_RESCALE, li, keys[ins.key_level], {regs[ins.a]}, {regs[ins.b]},
                         {dst}, s);
            break;
        }
        case kPsScalarMul: {
            const long long aps = (long long)p.reg_limbs[ins.a] * n;
            hipLaunchKernelGGL(k_scalar_mul, grid, dim3(256), 0, s, regs[ins.a], aps, dst, dps, (int)c.n,
                               consts_of(ins.k), c.d_primes);
            RCK_CHECK_HIP(hipGetLastError());
            break;
        }
        case kPsAddConst: {
            const long long aps = (long long)p.reg_limbs[ins.a] * n;
            hipLaunchKernelGGL(k_add_const0, grid, dim3(256), 0, s, regs[ins.a], aps, dst, dps, (int)c.n,
                               consts_of(ins.k), c.d_primes);
            RCK_CHECK_HIP(hipGetLastError());
            break;
        }
        case kPsAdd: {
            const long long aps = (long long)p.reg_limbs[ins.a] * n, bps = (long long)p.reg_limbs[ins.b] * n;
            launch_addsub(c, {regs[ins.a], n, aps}, {regs[ins.b], n, bps}, {dst, n, dps}, l, 2, 0,
                          PrimeMap{l, (int)c.L, 0}, s);
            break;
        }
        case kPsTruncate: {
            const long long aps = (long long)p.reg_limbs[ins.a] * n;
            for (int z = 0; z < 2; z++)
                RCK_CHECK_HIP(hipMemcpyAsync(dst + z * dps, regs[ins.a] + z * aps, (size_t)dps * 8,
                                             hipMemcpyDeviceToDevice, s));
            break;
        }
        case kPsZero: {
            const long long cnt = 2 * dps;
            hipLaunchKernelGGL(k_zero, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, dst, cnt);
            RCK_CHECK_HIP(hipGetLastError());
            break;
        }
        default:
            RCK_REQUIRE(false, "unknown program instruction");
    }
}

}  // namespace rck

using namespace rck;

namespace {

template <typename F>
int aguard(F&& f) {
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

std::vector<bool> key_mask(const Context& c, const rck_ksk* const* keys, uint32_t n_keys,
                           std::vector<const KeyDev*>* devs) {
    std::vector<bool> has(c.L, false);
    if (devs) devs->assign(c.L, nullptr);
    for (uint32_t l = 0; l < n_keys && l < c.L; l++) {
        if (!keys || !keys[l]) continue;
        RCK_REQUIRE((int)keys[l]->level == (int)l, "relinearization key at index " + std::to_string(l) +
                                                      " has level " + std::to_string(keys[l]->level));
        has[l] = true;
        if (devs) (*devs)[l] = keys[l];
    }
    return has;
}

}  // namespace

extern "C" {

int rck_ps_block_size(uint32_t term_count, uint32_t* block_size) {
    return aguard([&] {
        RCK_REQUIRE(block_size != nullptr, "null output");
        *block_size = (uint32_t)ps_block_size((int)term_count);
    });
}

int rck_ps_required_levels(uint32_t ciphertext_level, uint32_t term_count, uint32_t* levels, uint32_t cap,
                           uint32_t* count) {
    return aguard([&] {
        const std::vector<int> v = ps_required_levels((int)ciphertext_level, (int)term_count);
        RCK_REQUIRE(count != nullptr, "null output");
        *count = (uint32_t)v.size();
        RCK_REQUIRE(v.size() <= cap || levels == nullptr, "levels buffer too small");
        for (size_t i = 0; levels && i < v.size(); i++) levels[i] = (uint32_t)v[i];
    });
}

int rck_ps_plan(rck_context* ctx, uint32_t limbs, uint32_t scale_bits, const double* coeffs, uint32_t count,
                uint32_t* out_limbs, uint32_t* out_scale, uint32_t* mults, uint64_t* footprint) {
    return aguard([&] {
        RCK_REQUIRE(ctx && ctx->c, "null context");
        const Context& c = *ctx->c;
        require_fp64(c);
        const PsProgram p = ps_compile(c, (int)limbs, scale_bits, coeffs, (int)count, std::vector<bool>(c.L, true));
        if (out_limbs) *out_limbs = (uint32_t)p.reg_limbs[p.result];
        if (out_scale) *out_scale = p.reg_scale[p.result];
        if (mults) *mults = (uint32_t)p.mults();
        if (footprint) *footprint = p.footprint(c.n);
    });
}

int rck_ps_eval(rck_context* ctx, const uint64_t* ct, uint32_t limbs, uint32_t scale_bits, const double* coeffs,
                uint32_t count, const rck_ksk* const* relin_keys, uint32_t n_keys, uint64_t* out,
                uint32_t* out_limbs, uint32_t* out_scale, void* stream) {
    return aguard([&] {
        RCK_REQUIRE(ctx && ctx->c, "null context");
        Context& c = *ctx->c;
        require_fp64(c);
        check_ptr(ct, "ciphertext");
        check_ptr(out, "out");
        std::vector<const KeyDev*> keys;
        const std::vector<bool> has = key_mask(c, relin_keys, n_keys, &keys);
        const PsProgram p = ps_compile(c, (int)limbs, scale_bits, coeffs, (int)count, has);
        hipStream_t s = static_cast<hipStream_t>(stream);
        uint64_t* pool = nullptr;
        const size_t bytes = p.footprint(c.n);
        RCK_CHECK_HIP(hipMalloc(&pool, bytes ? bytes : 8));
        try {
            std::vector<uint64_t*> slot_ptr;
            size_t off = 0;
            for (int sl : p.slot_limbs) {
                slot_ptr.push_back(pool + off);
                off += (size_t)sl * 2 * c.n;
            }
            std::vector<uint64_t*> regs(p.reg_limbs.size(), nullptr);
            regs[0] = const_cast<uint64_t*>(ct);  // never written: no instruction has dst 0
            for (size_t r = 1; r < regs.size(); r++)
                if (p.slot_of[r] >= 0) regs[r] = slot_ptr[p.slot_of[r]];
            for (const PsInstr& ins : p.code) ps_exec(c, p, ins, regs, keys, s);
            const size_t rb = (size_t)2 * p.reg_limbs[p.result] * c.n * 8;
            RCK_CHECK_HIP(hipMemcpyAsync(out, regs[p.result], rb, hipMemcpyDeviceToDevice, s));
            RCK_CHECK_HIP(hipStreamSynchronize(s));
        } catch (...) {
            (void)hipStreamSynchronize(s);
            (void)hipFree(pool);
            throw;
        }
        RCK_CHECK_HIP(hipFree(pool));
        if (out_limbs) *out_limbs = (uint32_t)p.reg_limbs[p.result];
        if (out_scale) *out_scale = p.reg_scale[p.result];
    });
}

}  // extern "C"
