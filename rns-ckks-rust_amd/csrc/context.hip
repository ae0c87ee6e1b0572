// Context construction: validation (src/rns/context/params.rs:48-73, src/rns/basis.rs:32-59,
// src/math/ntt.rs:17-60), twiddle tables and basis-conversion constants uploaded to HBM.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>

#include "context.hpp"
#include "wide.hpp"

namespace rck {

typedef unsigned __int128 u128;

uint64_t h_mulmod(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)(((u128)a * b) % q); }

uint64_t h_powmod(uint64_t a, uint64_t e, uint64_t q) {
    uint64_t r = 1 % q;
    a %= q;
    while (e) {
        if (e & 1) r = h_mulmod(r, a, q);
        a = h_mulmod(a, a, q);
        e >>= 1;
    }
    return r;
}

bool h_invmod(uint64_t a, uint64_t q, uint64_t* out) {
    __int128 t = 0, nt = 1, r = q, nr = a % q;
    if (nr == 0) return false;
    while (nr != 0) {
        __int128 quo = r / nr, tmp = t - quo * nt;
        t = nt;
        nt = tmp;
        tmp = r - quo * nr;
        r = nr;
        nr = tmp;
    }
    if (r != 1) return false;
    if (t < 0) t += q;
    *out = (uint64_t)t;
    return true;
}

double h_centred(uint64_t x, uint64_t q) {
    x %= q;
    return x > q / 2 ? (double)x - (double)q : (double)x;
}

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }

// ---- per-kernel timing ----
namespace {
struct ProfRec {
    const char* name;
    hipEvent_t a, b;
};
std::mutex g_prof_mu;
bool g_prof_on = false;
std::vector<ProfRec> g_prof_recs;
std::vector<hipEvent_t> g_prof_pool;
hipEvent_t prof_event() {
    if (!g_prof_pool.empty()) {
        hipEvent_t e = g_prof_pool.back();
        g_prof_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    RCK_CHECK_HIP(hipEventCreate(&e));
    return e;
}
}  // namespace

bool prof_enabled() { return g_prof_on; }

ProfScope::ProfScope(hipStream_t s_, const char* n_) : s(s_), name(n_) {
    if (!g_prof_on) return;
    std::lock_guard<std::mutex> lk(g_prof_mu);
    ProfRec r{name, prof_event(), prof_event()};
    RCK_CHECK_HIP(hipEventRecord(r.a, s));
    slot = (int)g_prof_recs.size();
    g_prof_recs.push_back(r);
}

ProfScope::~ProfScope() {
    if (slot < 0) return;
    std::lock_guard<std::mutex> lk(g_prof_mu);
    (void)hipEventRecord(g_prof_recs[slot].b, s);
}

void prof_set(bool on) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof_on = on;
}

std::string prof_report() {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    RCK_CHECK_HIP(hipDeviceSynchronize());
    std::map<std::string, std::pair<long, double>> acc;
    for (auto& r : g_prof_recs) {
        float ms = 0;
        RCK_CHECK_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        auto& e = acc[r.name];
        e.first += 1;
        e.second += ms;
        g_prof_pool.push_back(r.a);
        g_prof_pool.push_back(r.b);
    }
    g_prof_recs.clear();
    std::string out = "{";
    bool first = true;
    for (auto& kv : acc) {
        char buf[256];
        snprintf(buf, sizeof buf, "%s\"%s\": [%ld, %.6f]", first ? "" : ", ", kv.first.c_str(), kv.second.first,
                 kv.second.second);
        out += buf;
        first = false;
    }
    return out + "}";
}

static uint32_t bitrev(uint32_t x, int bits) {
    uint32_t r = 0;
    for (int i = 0; i < bits; i++) {
        r = (r << 1) | (x & 1);
        x >>= 1;
    }
    return r;
}

// NttPlan::new validation, same messages (src/math/ntt.rs:17-60)
static std::string plan_check(uint64_t size, uint64_t q, uint64_t g, HostPrime* hp) {
    if (size == 0 || (size & (size - 1))) return "NTT size must be a non-zero power of two";
    if (q < 3) return "modulus must be an odd prime larger than 2";
    if ((q - 1) % size) return "size must divide modulus - 1";
    if ((q - 1) % (2 * size)) return "2 * size must divide modulus - 1 for negacyclic NTTs";
    uint64_t root = h_powmod(g % q, (q - 1) / size, q);
    if (h_powmod(root, size, q) != 1) return "primitive_root does not generate a size-th root of unity";
    if (size > 1 && h_powmod(root, size / 2, q) == 1) return "root is not primitive for the requested size";
    uint64_t root_inv, n_inv;
    if (!h_invmod(root, q, &root_inv)) return "root is not invertible";
    if (!h_invmod(size % q, q, &n_inv)) return "size is not invertible modulo modulus";
    uint64_t psi = h_powmod(g % q, (q - 1) / (2 * size), q);
    if (h_powmod(psi, 2 * size, q) != 1 || h_powmod(psi, size, q) != q - 1)
        return "primitive_root does not generate a valid negacyclic twiddle";
    uint64_t psi_inv;
    if (!h_invmod(psi, q, &psi_inv)) return "negacyclic twiddle is not invertible";
    *hp = HostPrime{q, g, root, root_inv, psi, psi_inv, n_inv};
    return "";
}

// words of a product of u64 moduli, little-endian 32-bit
static std::vector<uint32_t> product_words(const uint64_t* m, size_t k) {
    std::vector<uint32_t> w(1, 1);
    for (size_t i = 0; i < k; i++) {
        u128 carry = 0;
        for (auto& x : w) {
            u128 v = (u128)x * m[i] + carry;
            x = (uint32_t)v;
            carry = v >> 32;
        }
        while (carry) {
            w.push_back((uint32_t)carry);
            carry >>= 32;
        }
    }
    return w;
}

static std::vector<uint32_t> half_words(std::vector<uint32_t> w) {
    uint32_t carry = 0;
    for (size_t k = w.size(); k-- > 0;) {
        uint64_t v = ((uint64_t)carry << 32) | w[k];
        w[k] = (uint32_t)(v >> 1);
        carry = (uint32_t)(v & 1);
    }
    return w;
}

// 32-bit little-endian words -> 30-bit little-endian words (at most kWords30)
static std::vector<uint32_t> words30(const std::vector<uint32_t>& w) {
    std::vector<uint32_t> out;
    const size_t bits = w.size() * 32;
    for (size_t b = 0; b < bits; b += 30) {
        uint32_t v = 0;
        for (int k = 0; k < 30 && b + k < bits; k++) v |= ((w[(b + k) / 32] >> ((b + k) % 32)) & 1u) << k;
        out.push_back(v);
    }
    while (!out.empty() && out.back() == 0) out.pop_back();
    RCK_REQUIRE(out.size() <= (size_t)kWords30, "ciphertext modulus too large for the decomposition kernel");
    return out;
}

static uint32_t word_bits(const std::vector<uint32_t>& w) {
    size_t n = w.size();
    while (n > 1 && w[n - 1] == 0) n--;
    uint32_t top = w[n - 1], b = 0;
    while (top) {
        b++;
        top >>= 1;
    }
    return (uint32_t)(n - 1) * 32 + b;
}

// deterministic Miller-Rabin for 64-bit n (these bases suffice below 3.3 * 10^24)
static bool h_is_prime(uint64_t n) {
    if (n < 2) return false;
    static const uint64_t bases[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    for (uint64_t p : bases)
        if (n % p == 0) return n == p;
    uint64_t d = n - 1;
    int r = 0;
    while ((d & 1) == 0) {
        d >>= 1;
        r++;
    }
    for (uint64_t a : bases) {
        uint64_t x = h_powmod(a, d, n);
        if (x == 1 || x == n - 1) continue;
        bool comp = true;
        for (int i = 1; i < r && comp; i++) {
            x = h_mulmod(x, x, n);
            if (x == n - 1) comp = false;
        }
        if (comp) return false;
    }
    return true;
}

template <typename T>
static T* upload(const std::vector<T>& v) {
    T* d = nullptr;
    RCK_CHECK_HIP(hipMalloc(&d, std::max<size_t>(1, v.size()) * sizeof(T)));
    if (!v.empty()) RCK_CHECK_HIP(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

Context* context_create(uint32_t n, const uint64_t* q, const uint64_t* g, uint32_t L, const uint64_t* p,
                        const uint64_t* gp, uint32_t K, uint32_t scale_bits) {
    // RnsCkksParams::validate (params.rs:48-73)
    RCK_REQUIRE(scale_bits > 0, "scale_bits must be greater than zero");
    RCK_REQUIRE(scale_bits < 60, "scale_bits must stay below 60 bits per rescale step in this RNS scaffold");
    for (uint32_t k = 0; k < K; k++)
        for (uint32_t i = 0; i < L; i++)
            RCK_REQUIRE(q[i] != p[k], "auxiliary eval-key modulus " + std::to_string(p[k]) +
                                          " must be distinct from the ciphertext modulus chain");
    // RnsContext::new (basis.rs:32-59)
    RCK_REQUIRE(n > 0 && (n & (n - 1)) == 0, "poly_degree must be a non-zero power of two");
    RCK_REQUIRE(n >= 2, "poly_degree must be at least 2");
    RCK_REQUIRE(L > 0, "RNS basis must contain at least one prime");
    RCK_REQUIRE(n <= (1u << 16), "poly_degree above 2^16 is not supported by the NTT kernels");
    RCK_REQUIRE(L <= (uint32_t)kMaxQ, "at most 64 ciphertext primes are supported");
    RCK_REQUIRE(K <= (uint32_t)kMaxP, "at most 4 auxiliary primes are supported");
    std::vector<uint64_t> all(q, q + L), roots(g, g + L);
    all.insert(all.end(), p, p + K);
    roots.insert(roots.end(), gp, gp + K);
    for (size_t i = 0; i < all.size(); i++)
        for (size_t j = 0; j < i; j++)
            RCK_REQUIRE(all[i] != all[j], "duplicate RNS modulus " + std::to_string(all[i]));
    // Two auxiliary NTT primes for the exact-convolution keyswitch (hybrid keys, N >= 2^12): the largest
    // primes below 2^50 + 2^46 with s = 1 mod 2N outside Q || P, each with its smallest quadratic
    // non-residue as generator (a non-residue g makes g^((s-1)/2N) a primitive 2N-th root of unity).
    // the FP64-exact products need |a*w/q| < 2^51 with lazily reduced inputs (modarith.hpp); a larger
    // modulus selects the 64-bit integer path for the whole context
    bool wide = false;
    for (uint64_t x : all) wide |= x >= (1ull << 50) + (1ull << 46);
    uint32_t S = 0;
    if (K > 0 && n >= (1u << 12) && !wide) {
        const uint64_t step = 2ull * n;
        uint64_t cand = (((1ull << 50) + (1ull << 46) - 2) / step) * step + 1;
        while (S < 2 && cand > step) {
            if (h_is_prime(cand) && std::find(all.begin(), all.end(), cand) == all.end()) {
                uint64_t g2 = 2;
                while (h_powmod(g2, (cand - 1) / 2, cand) == 1) g2++;
                all.push_back(cand);
                roots.push_back(g2);
                S++;
            }
            cand -= step;
        }
        RCK_REQUIRE(S == 2, "no auxiliary NTT primes for the exact-convolution keyswitch");
    }
    Context* c = new Context();
    try {
        c->n = n;
        c->logn = 0;
        while ((1u << c->logn) < n) c->logn++;
        c->L = L;
        c->K = K;
        c->S = S;
        c->scale_bits = scale_bits;
        const uint32_t T = L + K, TS = T + S;
        c->hp.resize(TS);
        for (uint32_t i = 0; i < TS; i++) {
            std::string err = plan_check(n, all[i], roots[i], &c->hp[i]);
            RCK_REQUIRE(err.empty(), "invalid RNS prime " + std::to_string(all[i]) + ": " + err);
        }
        // twiddle tables per prime: fw, iv (merged negacyclic) and fwc, ivc (cyclic), all in the
        // Cooley-Tukey group order k = 2^s + g.  fw[k] = psi^{brv_logN(k)} with
        // brv_logN(k) = (2 brv_s(g) + 1) N / 2^{s+1}; the cyclic table drops the odd part:
        // fwc[k] = psi^{2 brv_s(g) N / 2^{s+1}} = psi^{brv_logN(k) - N / 2^{s+1}}.  The same merged
        // loops then compute the cyclic DFT in bit-reversed order (and its inverse from that order).
        std::vector<double> tables((size_t)TS * 4 * n);
        for (uint32_t i = 0; i < TS; i++) {
            const HostPrime& h = c->hp[i];
            double* fw = &tables[(size_t)i * 4 * n];
            double *iv = fw + n, *fwc = iv + n, *ivc = fwc + n;
            std::vector<uint64_t> pw(2 * n);  // psi^e for e < 2N (psi^{2N} = 1)
            uint64_t a = 1;
            for (uint32_t k = 0; k < 2 * n; k++) {
                pw[k] = a;
                a = h_mulmod(a, h.psi, h.q);
            }
            for (uint32_t k = 0; k < n; k++) {
                const uint32_t r = bitrev(k, c->logn);
                fw[k] = h_centred(pw[r], h.q);
                iv[k] = h_centred(pw[(2 * n - r) % (2 * n)], h.q);
                uint32_t s = 0;
                while (k >> (s + 1)) s++;
                const uint32_t e = k == 0 ? 0 : r - (n >> (s + 1));
                fwc[k] = h_centred(pw[e], h.q);
                ivc[k] = h_centred(pw[(2 * n - e) % (2 * n)], h.q);
            }
        }
        c->d_tables = upload(tables);
        std::vector<DevPrime> dp(TS);
        for (uint32_t i = 0; i < TS; i++) {
            const HostPrime& h = c->hp[i];
            double* base = c->d_tables + (size_t)i * 4 * n;
            dp[i] = DevPrime{(double)h.q, 1.0 / (double)h.q, h_centred(h.n_inv, h.q), h.q, base, base + n,
                             base + 2 * n, base + 3 * n};
        }
        c->d_primes = upload(dp);
        if (S == 2) {
            // exact-convolution constants: s1^{-1} mod s2, and s1 mod p_j, s1 s2 mod p_j per eval prime
            const uint64_t s1 = all[T], s2 = all[T + 1];
            c->sp[0] = s1;
            c->sp[1] = s2;
            uint64_t inv;
            RCK_REQUIRE(h_invmod(s1 % s2, s2, &inv), "auxiliary NTT primes must be coprime");
            c->sp_inv12 = h_centred(inv, s2);
            std::vector<double> spm((size_t)T * 2);
            for (uint32_t j = 0; j < T; j++) {
                spm[2 * j] = h_centred(s1 % all[j], all[j]);
                spm[2 * j + 1] = h_centred(h_mulmod(s1 % all[j], s2 % all[j], all[j]), all[j]);
            }
            c->d_spmod = upload(spm);
        }
        // inverse table q_j^{-1} mod q_i
        std::vector<double> invq((size_t)T * T, 0.0);
        for (uint32_t i = 0; i < T; i++)
            for (uint32_t j = 0; j < T; j++) {
                if (i == j) continue;
                uint64_t v;
                RCK_REQUIRE(h_invmod(all[j] % all[i], all[i], &v), "RNS moduli must be pairwise coprime");
                invq[(size_t)i * T + j] = h_centred(v, all[i]);
            }
        c->d_invq = upload(invq);
        // Q_l words for l = 0..L (30-bit words on the device)
        // [l][0][:] = words of Q_l, [l][1][:] = words of floor(Q_l / 2)
        std::vector<uint32_t> qw((size_t)(L + 1) * 2 * kWords30, 0);
        c->qwords.resize(L + 1);
        c->qbits.resize(L + 1);
        {
            uint64_t qmin = all[0], qmax = all[0];
            for (uint32_t i = 1; i < L; i++) qmin = std::min(qmin, all[i]), qmax = std::max(qmax, all[i]);
            c->garner_lazy = (double)qmax <= 2.5 * (double)qmin;
        }
        for (uint32_t l = 0; l <= L; l++) {
            auto w = product_words(all.data(), l);
            RCK_REQUIRE(w.size() <= (size_t)kWords - 2, "ciphertext modulus too large for the decomposition kernel");
            c->qwords[l] = w;
            c->qbits[l] = word_bits(w);
            auto w30 = words30(w), h30 = words30(half_words(w));
            std::memcpy(&qw[(size_t)l * 2 * kWords30], w30.data(), w30.size() * 4);
            std::memcpy(&qw[(size_t)l * 2 * kWords30 + kWords30], h30.data(), h30.size() * 4);
        }
        c->d_qwords = upload(qw);
        if (K > 0) {
            auto pw = product_words(all.data() + L, K);
            std::vector<uint32_t> pwv(16, 0);  // [0..8) = P, [8..16) = floor(P / 2)
            RCK_REQUIRE(pw.size() <= 7 || wide, "aux modulus too large");
            std::memcpy(pwv.data(), pw.data(), pw.size() * 4);
            auto ph = half_words(pw);
            std::memcpy(pwv.data() + 8, ph.data(), ph.size() * 4);
            c->d_pwords = upload(pwv);
            std::vector<double> pmodq(L), pinvq(L), pmodaux((size_t)K * L);
            for (uint32_t i = 0; i < L; i++) {
                uint64_t m = 1 % all[i];
                for (uint32_t k = 0; k < K; k++) {
                    pmodaux[(size_t)k * L + i] = h_centred(all[L + k] % all[i], all[i]);
                    m = h_mulmod(m, all[L + k] % all[i], all[i]);
                }
                uint64_t inv;
                RCK_REQUIRE(h_invmod(m, all[i], &inv), "P must be invertible modulo each q_i");
                pmodq[i] = h_centred(m, all[i]);
                pinvq[i] = h_centred(inv, all[i]);
            }
            c->d_pmodq = upload(pmodq);
            c->d_pinvq = upload(pinvq);
            c->d_pmod_aux = upload(pmodaux);
        }
        if (wide) context_make_wide(*c);
    } catch (...) {
        delete c;
        throw;
    }
    return c;
}

void context_make_wide(Context& c) {
    if (!c.d_wp) {
        std::vector<uint64_t> all;
        for (uint32_t i = 0; i < c.L + c.K; i++) all.push_back(c.hp[i].q);
        wide_tables_create(c, all, c.L + c.K);
    }
    c.wide = true;
}

void* Context::workspace(int slot, size_t bytes) {
    if ((size_t)slot >= ws_ptr.size()) {
        ws_ptr.resize(slot + 1, nullptr);
        ws_size.resize(slot + 1, 0);
    }
    if (ws_size[slot] < bytes) {
        RCK_REQUIRE(!frozen, "workspace growth during graph capture");
        if (ws_ptr[slot]) {
            RCK_CHECK_HIP(hipDeviceSynchronize());
            RCK_CHECK_HIP(hipFree(ws_ptr[slot]));
            ws_ptr[slot] = nullptr;
            ws_size[slot] = 0;
        }
        RCK_CHECK_HIP(hipMalloc(&ws_ptr[slot], bytes));
        ws_size[slot] = bytes;
    }
    return ws_ptr[slot];
}

hipStream_t Context::aux_stream() {
    if (!aux) {
        RCK_REQUIRE(!frozen, "stream creation during graph capture");
        RCK_CHECK_HIP(hipStreamCreateWithFlags(&aux, hipStreamNonBlocking));
    }
    return aux;
}

hipEvent_t Context::event(int i) {
    while ((int)events.size() <= i) {
        RCK_REQUIRE(!frozen, "event creation during graph capture");
        hipEvent_t e;
        RCK_CHECK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        events.push_back(e);
    }
    return events[i];
}

Context* Context::scratch_clone() const {
    Context* c = new Context();
    c->n = n;
    c->logn = logn;
    c->L = L;
    c->K = K;
    c->scale_bits = scale_bits;
    c->hp = hp;
    c->d_primes = d_primes;
    c->d_tables = d_tables;
    c->d_invq = d_invq;
    c->d_qwords = d_qwords;
    c->qwords = qwords;
    c->qbits = qbits;
    c->garner_lazy = garner_lazy;
    c->d_pwords = d_pwords;
    c->d_pmodq = d_pmodq;
    c->d_pinvq = d_pinvq;
    c->d_pmod_aux = d_pmod_aux;
    c->S = S;
    c->sp[0] = sp[0];
    c->sp[1] = sp[1];
    c->sp_inv12 = sp_inv12;
    c->d_spmod = d_spmod;
    c->wide = wide;
    c->d_wp = d_wp;
    c->d_wtables = d_wtables;
    c->d_winvq = d_winvq;
    c->d_wpmodq = d_wpmodq;
    c->d_wpinvq = d_wpinvq;
    c->d_wpaux = d_wpaux;
    c->d_wphalf = d_wphalf;
    c->ks_method = ks_method;
    c->tune_group = tune_group;
    c->tune_dsplit = tune_dsplit;
    c->owns_tables = false;
    return c;
}

Context::~Context() {
    (void)hipDeviceSynchronize();
    for (void* p : ws_ptr)
        if (p) (void)hipFree(p);
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
    if (aux) (void)hipStreamDestroy(aux);
    if (!owns_tables) return;
    void* ptrs[] = {d_primes, d_tables,  d_invq,    d_qwords,  d_pwords,  d_pmodq, d_pinvq,
                    d_pmod_aux, d_spmod, d_wp, d_wtables, d_winvq, d_wpmodq, d_wpinvq, d_wpaux, d_wphalf};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
}

}  // namespace rck
