// Baby-step giant-step diagonal linear transform (src/ckks/bootstrapping/linear_transform.rs:78-253,
// via src/algorithms/packed_diagonal.rs:33-77) for a batch of B ciphertexts sharing one matrix.
//
// The crate computes, per ciphertext,
//   baby_b  = rotate(ct, b)                                b = 1 .. bs-1   (baby_0 = ct)
//   inner_g = sum_b mult_plain(baby_b, diag[g][b])         (scale doubles; left unrescaled)
//   acc     = inner_0 + sum_{g>=1} rotate(inner_g, g * bs)
//   out     = rescale(acc)
// Here the same values are computed for the whole batch:
//   * each baby rotation is ONE batched keyswitch over the B ciphertexts (one key stream);
//   * the baby rotations are forward-transformed once, and every inner sum of every giant step for
//     every ciphertext comes out of one multiply-accumulate kernel in the NTT domain:
//       inner[g][v][z] = sum_b NTT(baby_b[v][z]) (.) NTT(diag[g][b])
//     (a 64 x 64 x 64 product per coefficient at the crate's sizes), followed by one inverse NTT —
//     the exact negacyclic products and their sum mod q, i.e. the crate's mult_plain + add values;
//   * each giant rotation is one batched keyswitch; the sums and the final rescale are elementwise.
// The diagonals are encoded once (encode_at_level_and_scale of packed_bsgs_diagonal_slots, :255-280)
// and kept NTT-resident.
#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/rnsckks.h"
#include "algorithms.hpp"
#include "capi_types.hpp"

struct rck_diag {
    rck_context* ctx = nullptr;
    uint32_t dim = 0, baby = 0, giants = 0, limbs = 0, scale_bits = 0;
    std::vector<uint8_t> present;  // [giants][baby]: diagonal step < dim
    double* ntt = nullptr;         // [giants][baby][limbs][N] NTT-form doubles (absent entries zero)
};

namespace rck {
namespace {

uint32_t floor_log2_u64(uint64_t q) { return 63 - __builtin_clzll(q); }

// diagonal_transform_bsgs_baby_step (:146-152)
uint32_t bsgs_baby_step(uint32_t dim) {
    uint32_t b = 1;
    while ((uint64_t)b * b < dim) b++;
    return b;
}

// inner[g][vz][l][x] = sum_b diag[g][b][l][x] * baby[b][vz][l][x]  (NTT domain, reduced doubles):
// per coefficient (l, x) a (giants x bs) by (bs x vz) product mod q_l.  Lanes run over x (coalesced
// loads and stores); each thread owns a TG x TV register tile of (giant, vz) outputs, so every loaded
// diagonal value feeds TV products and every baby value TG products.  Absent diagonals are stored as
// zeros and contribute exact zeros.  Both operands are reduced (NTT outputs, |.| <= q/2 + 1), so
// each product is <= 0.57q (modarith.hpp: 0.5 + 0.133 A at A = 0.5) and the running sums are
// reduced every eighth term (|sum| <= 0.5q + 8 * 0.57q < 5.1q < 2^53 / q: every addition stays exact).
template <int TG, int TV>
__global__ __launch_bounds__(256) void k_diag_mac(const double* __restrict__ baby, const double* __restrict__ diag,
                                                  double* __restrict__ inner, int n, int limbs, int vz_count,
                                                  int giants, int bs, const DevPrime* __restrict__ primes) {
    // tiles fastest: the workgroups sharing a coefficient block run back to back (cache reuse)
    const int x = blockIdx.z * 256 + threadIdx.x;
    const int l = blockIdx.y;
    const int ngb = (giants + TG - 1) / TG;
    const int g0 = (blockIdx.x % ngb) * TG, v0 = (blockIdx.x / ngb) * TV;
    if (x >= n) return;
    const DevPrime P = primes[l];
    const long long lim = (long long)limbs * n;  // stride of one polynomial
    const long long off = (long long)l * n + x;
    double acc[TG][TV];
#pragma unroll
    for (int t = 0; t < TG; t++)
#pragma unroll
        for (int u = 0; u < TV; u++) acc[t][u] = 0.0;
    for (int b = 0; b < bs; b++) {
        double d[TG], w[TV];
#pragma unroll
        for (int t = 0; t < TG; t++)
            d[t] = g0 + t < giants ? diag[((long long)(g0 + t) * bs + b) * lim + off] : 0.0;
#pragma unroll
        for (int u = 0; u < TV; u++)
            w[u] = v0 + u < vz_count ? baby[((long long)b * vz_count + v0 + u) * lim + off] : 0.0;
#pragma unroll
        for (int t = 0; t < TG; t++)
#pragma unroll
            for (int u = 0; u < TV; u++) acc[t][u] += mulmod(d[t], w[u], P.q, P.qinv);
        if ((b & 7) == 7) {
#pragma unroll
            for (int t = 0; t < TG; t++)
#pragma unroll
                for (int u = 0; u < TV; u++) acc[t][u] = reduce(acc[t][u], P.q, P.qinv);
        }
    }
#pragma unroll
    for (int t = 0; t < TG; t++)
#pragma unroll
        for (int u = 0; u < TV; u++)
            if (g0 + t < giants && v0 + u < vz_count)
                inner[((long long)(g0 + t) * vz_count + v0 + u) * lim + off] = reduce(acc[t][u], P.q, P.qinv);
}

template <typename F>
int bguard2(F&& f) {
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

void launch_diag_mac(const Context& c, int bs, const double* baby, const double* diag, double* inner, int limbs,
                     int vz, int giants, hipStream_t s) {
    ProfScope prof(s, "diag_mac");
    RCK_REQUIRE(bs >= 1 && bs <= 64, "baby step must be at most 64 (matrix dimension at most 4096)");
    if (giants >= 8 && vz >= 8) {
        const dim3 grid(((giants + 7) / 8) * ((vz + 7) / 8), limbs, (c.n + 255) / 256);
        hipLaunchKernelGGL((k_diag_mac<8, 8>), grid, dim3(256), 0, s, baby, diag, inner, (int)c.n, limbs, vz, giants,
                           bs, c.d_primes);
    } else {
        const dim3 grid(((giants + 3) / 4) * ((vz + 3) / 4), limbs, (c.n + 255) / 256);
        hipLaunchKernelGGL((k_diag_mac<4, 4>), grid, dim3(256), 0, s, baby, diag, inner, (int)c.n, limbs, vz, giants,
                           bs, c.d_primes);
    }
    RCK_CHECK_HIP(hipGetLastError());
}

template <typename T>
struct DevArr {
    T* p = nullptr;
    explicit DevArr(size_t count) { RCK_CHECK_HIP(hipMalloc(&p, count ? count * sizeof(T) : 8)); }
    ~DevArr() {
        if (p) (void)hipFree(p);
    }
    DevArr(const DevArr&) = delete;
    DevArr& operator=(const DevArr&) = delete;
};

}  // namespace
}  // namespace rck

using namespace rck;

extern "C" {

int rck_bsgs_rotation_steps(uint32_t dim, uint32_t* steps, uint32_t cap, uint32_t* count) {
    return bguard2([&] {
        RCK_REQUIRE(dim > 0, "transform dimension must be positive");
        RCK_REQUIRE(count != nullptr, "null output");
        std::vector<uint32_t> v;
        if (dim > 1) {
            const uint32_t bs = bsgs_baby_step(dim), gc = (dim + bs - 1) / bs;
            for (uint32_t b = 1; b < std::min(bs, dim); b++) v.push_back(b);
            for (uint32_t g = 1; g < gc; g++)
                if (g * bs < dim) v.push_back(g * bs);
            std::sort(v.begin(), v.end());
            v.erase(std::unique(v.begin(), v.end()), v.end());
        }
        *count = (uint32_t)v.size();
        RCK_REQUIRE(steps == nullptr || v.size() <= cap, "steps buffer too small");
        for (size_t i = 0; steps && i < v.size(); i++) steps[i] = v[i];
    });
}

int rck_diag_pack(rck_context* ctx, const double* matrix, uint32_t dim, uint32_t level, uint32_t scale_bits,
                  rck_diag** out) {
    return bguard2([&] {
        RCK_REQUIRE(ctx && ctx->c && matrix && out, "null argument");
        Context& c = *ctx->c;
        require_fp64(c);
        const uint32_t slots = c.n / 2;
        RCK_REQUIRE(dim > 0, "matrix cannot be empty");
        RCK_REQUIRE((dim & (dim - 1)) == 0, "matrix dimension must be a power of two");
        RCK_REQUIRE(dim <= slots, "matrix dimension cannot exceed slot count");
        RCK_REQUIRE(level < c.L, "requested level context must exist for encoding");
        const uint32_t bs = bsgs_baby_step(dim), gc = (dim + bs - 1) / bs, limbs = level + 1;
        rck_diag* d = new rck_diag();
        d->ctx = ctx;
        d->dim = dim;
        d->baby = bs;
        d->giants = gc;
        d->limbs = limbs;
        d->scale_bits = scale_bits;
        d->present.assign((size_t)gc * bs, 0);
        const long long n = c.n, P = (long long)limbs * n;
        try {
            RCK_CHECK_HIP(hipMalloc(&d->ntt, (size_t)gc * bs * P * 8));
            RCK_CHECK_HIP(hipMemset(d->ntt, 0, (size_t)gc * bs * P * 8));
            // encode every present diagonal on host threads, then upload and transform in chunks
            std::vector<std::pair<uint32_t, uint32_t>> todo;
            for (uint32_t g = 0; g < gc; g++)
                for (uint32_t b = 0; b < bs; b++)
                    if (g * bs + b < dim) {
                        todo.push_back({g, b});
                        d->present[(size_t)g * bs + b] = 1;
                    }
            const size_t chunk = 64;
            std::vector<uint64_t> host(chunk * P);
            DevArr<uint64_t> stage(chunk * P);
            for (size_t t0 = 0; t0 < todo.size(); t0 += chunk) {
                const size_t cnt = std::min(chunk, todo.size() - t0);
                const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
                std::vector<std::thread> ts;
                for (unsigned w = 0; w < hw; w++)
                    ts.emplace_back([&, w] {
                        std::vector<double> re(slots);
                        for (size_t k = w; k < cnt; k += hw) {
                            const uint32_t g = todo[t0 + k].first, b = todo[t0 + k].second;
                            const uint32_t step = g * bs + b, shift = (g * bs) % dim;
                            // packed_bsgs_diagonal_slots (:255-280): period-dim tiling of the shifted diagonal
                            for (uint32_t slot = 0; slot < slots; slot++) {
                                const uint32_t row = slot % dim;
                                const uint32_t src = (row + dim - shift) % dim;
                                re[slot] = matrix[(size_t)src * dim + (step + src) % dim];
                            }
                            encode_host(c, re.data(), nullptr, limbs, scale_bits, host.data() + k * P, false);
                        }
                    });
                for (auto& t : ts) t.join();
                RCK_CHECK_HIP(hipMemcpy(stage.p, host.data(), cnt * P * 8, hipMemcpyHostToDevice));
                for (size_t k = 0; k < cnt; k++) {
                    const uint32_t g = todo[t0 + k].first, b = todo[t0 + k].second;
                    launch_fwd(c, kInU64, stage.p + k * P, n, P, d->ntt + ((size_t)g * bs + b) * P, n, P, (int)limbs,
                               1, PrimeMap{(int)limbs, (int)c.L, 0}, nullptr);
                }
                RCK_CHECK_HIP(hipDeviceSynchronize());
            }
        } catch (...) {
            if (d->ntt) (void)hipFree(d->ntt);
            delete d;
            throw;
        }
        *out = d;
    });
}

int rck_diag_info(const rck_diag* d, uint32_t* dim, uint32_t* baby_step, uint32_t* giant_steps, uint32_t* level,
                  uint32_t* scale_bits) {
    return bguard2([&] {
        RCK_REQUIRE(d != nullptr, "null diagonals");
        if (dim) *dim = d->dim;
        if (baby_step) *baby_step = d->baby;
        if (giant_steps) *giant_steps = d->giants;
        if (level) *level = d->limbs - 1;
        if (scale_bits) *scale_bits = d->scale_bits;
    });
}

int rck_diag_export(const rck_diag* d, uint64_t* out) {
    return bguard2([&] {
        RCK_REQUIRE(d && out, "null argument");
        Context& c = *d->ctx->c;
        const long long n = c.n, P = (long long)d->limbs * n;
        DevArr<double> t(P);
        DevArr<uint64_t> u(P);
        for (uint32_t k = 0; k < d->giants * d->baby; k++) {
            if (!d->present[k]) {
                std::memset(out + (size_t)k * P, 0, P * 8);
                continue;
            }
            RCK_CHECK_HIP(hipMemcpy(t.p, d->ntt + (size_t)k * P, P * 8, hipMemcpyDeviceToDevice));
            launch_inv(c, t.p, n, P, kOutU64, u.p, n, P, (int)d->limbs, 1, PrimeMap{(int)d->limbs, (int)c.L, 0},
                       nullptr);
            RCK_CHECK_HIP(hipMemcpy(out + (size_t)k * P, u.p, P * 8, hipMemcpyDeviceToHost));
        }
    });
}

int rck_diag_destroy(rck_diag* d) {
    return bguard2([&] {
        if (!d) return;
        RCK_CHECK_HIP(hipDeviceSynchronize());
        if (d->ntt) RCK_CHECK_HIP(hipFree(d->ntt));
        delete d;
    });
}

int rck_bsgs_matvec(rck_context* ctx, const uint64_t* cts, uint32_t batch, uint32_t limbs, uint32_t scale_bits,
                    const rck_diag* d, const rck_ksk* const* keys_by_step, uint32_t n_keys, uint64_t* out,
                    uint32_t* out_scale, void* stream) {
    return bguard2([&] {
        RCK_REQUIRE(ctx && ctx->c && cts && d && out, "null argument");
        Context& c = *ctx->c;
        require_fp64(c);
        RCK_REQUIRE(batch >= 1, "batch must be positive");
        RCK_REQUIRE(d->ctx == ctx, "packed diagonals belong to another context");
        RCK_REQUIRE(d->limbs == limbs, "packed diagonals must match the ciphertext level");
        RCK_REQUIRE(d->scale_bits == scale_bits, "packed diagonals must match the ciphertext scale");
        RCK_REQUIRE(limbs >= 2, "rescale requires at least two levels");
        hipStream_t s = static_cast<hipStream_t>(stream);
        const uint32_t bs = d->baby, gc = d->giants, nb = std::min(bs, d->dim);
        const long long n = c.n, P = (long long)limbs * n, CT = 2 * P;
        const PrimeMap pm{(int)limbs, (int)c.L, 0};
        auto key_for = [&](uint32_t step, const char* what) -> const KeyDev* {
            const rck_ksk* k = step < n_keys && keys_by_step ? keys_by_step[step] : nullptr;
            RCK_REQUIRE(k != nullptr, std::string("missing input rotation key for BSGS ") + what + " step " +
                                          std::to_string(step));
            RCK_REQUIRE((int)k->level == (int)limbs - 1, "rotation key level must match the ciphertext level");
            return k;
        };
        // baby rotations of every ciphertext (baby_0 = the ciphertexts themselves)
        // persistent scratch of the context (slots 20-26), grown on demand
        uint64_t* baby_p = static_cast<uint64_t*>(c.workspace(20, (size_t)nb * batch * CT * 8));
        double* bntt_p = static_cast<double*>(c.workspace(21, (size_t)nb * batch * CT * 8));
        double* inner_p = static_cast<double*>(c.workspace(22, (size_t)gc * batch * CT * 8));
        uint64_t* coef_p = static_cast<uint64_t*>(c.workspace(23, (size_t)gc * batch * CT * 8));
        uint64_t* acc_p = static_cast<uint64_t*>(c.workspace(24, (size_t)batch * CT * 8));
        uint64_t* rot_p = static_cast<uint64_t*>(c.workspace(25, (size_t)batch * CT * 8));
        RCK_CHECK_HIP(hipMemcpyAsync(baby_p, cts, (size_t)batch * CT * 8, hipMemcpyDeviceToDevice, s));
        std::vector<const uint64_t*> ins;
        for (uint32_t v = 0; v < batch; v++) ins.push_back(cts + v * CT);
        for (uint32_t b = 1; b < nb; b++) {
            std::vector<uint64_t*> outs;
            for (uint32_t v = 0; v < batch; v++) outs.push_back(baby_p + ((size_t)b * batch + v) * CT);
            run_op_group(c, RCK_OP_ROTATE, (int)limbs, key_for(b, "baby"), ins, ins, outs, s);
        }
        // NTT of every baby rotation: [b][v][z][l][N]
        launch_fwd(c, kInU64, baby_p, n, P, bntt_p, n, P, (int)limbs, (int)(nb * batch * 2), pm, s);
        // every inner sum of every giant step, then back to coefficient form: [g][v][z][l][N]
        const int vz = (int)batch * 2;
        RCK_REQUIRE(nb == bs, "matrix dimension below the baby step is not supported");
        launch_diag_mac(c, (int)bs, bntt_p, d->ntt, inner_p, (int)limbs, vz, (int)gc, s);
        launch_inv(c, inner_p, n, P, kOutU64, coef_p, n, P, (int)limbs, (int)(gc * batch * 2), pm, s);
        // giant rotations and the sum (a giant step whose diagonals are all absent contributes nothing)
        bool have = false;
        for (uint32_t g = 0; g < gc; g++) {
            bool any = false;
            for (uint32_t b = 0; b < bs; b++) any |= d->present[(size_t)g * bs + b] != 0;
            if (!any) continue;
            const uint64_t* src = coef_p + (size_t)g * batch * CT;
            if (g > 0) {
                std::vector<const uint64_t*> gi;
                std::vector<uint64_t*> go;
                for (uint32_t v = 0; v < batch; v++) {
                    gi.push_back(src + v * CT);
                    go.push_back(rot_p + v * CT);
                }
                run_op_group(c, RCK_OP_ROTATE, (int)limbs, key_for(g * bs, "giant"), gi, gi, go, s);
                src = rot_p;
            }
            if (!have) {
                RCK_CHECK_HIP(hipMemcpyAsync(acc_p, src, (size_t)batch * CT * 8, hipMemcpyDeviceToDevice, s));
                have = true;
            } else {
                launch_addsub(c, {acc_p, n, P}, {src, n, P}, {acc_p, n, P}, (int)limbs, (int)(2 * batch), 0, pm, s);
            }
        }
        RCK_REQUIRE(have, "diagonal transform must contain at least one non-empty giant-step block");
        launch_rescale(c, {acc_p, n, P}, {out, n, (long long)(limbs - 1) * n}, (int)limbs, (int)(2 * batch), s);
        RCK_CHECK_HIP(hipStreamSynchronize(s));
        if (out_scale) {
            const uint32_t sc = 2 * scale_bits, drop = floor_log2_u64(c.hp[limbs - 1].q);
            *out_scale = sc > drop ? sc - drop : 0;
        }
    });
}

}  // extern "C"
