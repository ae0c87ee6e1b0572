// Versioned limb-major binary format for polynomials / ciphertexts and keyswitching keys.
//
// The crate has no serialization; this format carries device data between processes (fixtures,
// key upload, request payloads).  Layout (little-endian):
//   offset 0   char[4]  magic "RCK1"
//          4   u32      version (1)
//          8   u32      kind (1 = polynomials, 2 = keyswitching key)
//         12   u32      log2 N
//         16   u32      count (polynomials, or key digits)
//         20   u32      Q limbs
//         24   u32      aux limbs
//         28   u32      scale_bits (polynomials) / key level (keys)
//         32   u64      galois element (keys; 0 otherwise)
//         40   u64      fingerprint of the prime chain the residues belong to (FNV-1a of the moduli)
//         48   u64      payload bytes
//         56   payload  u64 residues, [count][(2 for keys)][Q limbs + aux limbs][N], canonical
// Keys are written in the crate's coefficient form (what rck_ksk_create takes).
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rnsckks.h"
#include "capi_types.hpp"

namespace rck {
namespace {

constexpr uint32_t kVersion = 1;
constexpr size_t kHeader = 56;

template <typename F>
int zguard(F&& f) {
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

uint64_t fingerprint(const Context& c, uint32_t nq, uint32_t np) {
    uint64_t h = 1469598103934665603ULL;
    auto mix = [&](uint64_t v) {
        for (int b = 0; b < 8; b++) {
            h ^= (v >> (8 * b)) & 0xff;
            h *= 1099511628211ULL;
        }
    };
    mix(c.n);
    for (uint32_t l = 0; l < nq; l++) mix(c.hp[l].q);
    for (uint32_t k = 0; k < np; k++) mix(c.hp[c.L + k].q);
    return h;
}

struct Header {
    uint32_t kind, logn, count, nq, np, scale;
    uint64_t galois, fp, payload;
};

void put32(uint8_t* p, uint32_t v) { std::memcpy(p, &v, 4); }
void put64(uint8_t* p, uint64_t v) { std::memcpy(p, &v, 8); }
uint32_t get32(const uint8_t* p) {
    uint32_t v;
    std::memcpy(&v, p, 4);
    return v;
}
uint64_t get64(const uint8_t* p) {
    uint64_t v;
    std::memcpy(&v, p, 8);
    return v;
}

void write_header(uint8_t* b, const Header& h) {
    std::memcpy(b, "RCK1", 4);
    put32(b + 4, kVersion);
    put32(b + 8, h.kind);
    put32(b + 12, h.logn);
    put32(b + 16, h.count);
    put32(b + 20, h.nq);
    put32(b + 24, h.np);
    put32(b + 28, h.scale);
    put64(b + 32, h.galois);
    put64(b + 40, h.fp);
    put64(b + 48, h.payload);
}

Header read_header(const Context& c, const uint8_t* b, size_t len, uint32_t kind) {
    RCK_REQUIRE(b != nullptr && len >= kHeader, "serialized data is truncated");
    RCK_REQUIRE(std::memcmp(b, "RCK1", 4) == 0, "not serialized library data (bad magic)");
    RCK_REQUIRE(get32(b + 4) == kVersion, "unsupported serialized data version");
    Header h{get32(b + 8), get32(b + 12), get32(b + 16), get32(b + 20), get32(b + 24), get32(b + 28),
             get64(b + 32), get64(b + 40), get64(b + 48)};
    RCK_REQUIRE(h.kind == kind, kind == 1 ? "serialized data is not a polynomial set" : "serialized data is not a key");
    RCK_REQUIRE(h.logn == c.logn, "serialized data has another ring degree");
    RCK_REQUIRE(h.nq >= 1 && h.nq <= c.L && (h.np == 0 || h.np == c.K), "serialized limb counts do not fit the context");
    RCK_REQUIRE(h.fp == fingerprint(c, h.nq, h.np), "serialized data belongs to another prime chain");
    const uint64_t per = (uint64_t)(h.nq + h.np) * c.n * 8 * (kind == 2 ? 2 : 1);
    RCK_REQUIRE(h.count >= 1 && h.payload == per * h.count, "serialized payload size does not match its header");
    RCK_REQUIRE(len >= kHeader + h.payload, "serialized data is truncated");
    return h;
}

// Every residue of the payload ([groups][Q limbs + aux limbs][N]) must be canonical, below its limb's
// modulus: the kernels' exact arithmetic assumes reduced inputs, so a damaged payload is rejected here
// instead of producing wrong results later.
void check_canonical(const Context& c, const uint8_t* payload, uint64_t groups, uint32_t nq, uint32_t np) {
    const uint32_t limbs = nq + np;
    for (uint64_t g = 0; g < groups; g++) {
        for (uint32_t l = 0; l < limbs; l++) {
            const uint64_t q = l < nq ? c.hp[l].q : c.hp[c.L + (l - nq)].q;
            const uint8_t* row = payload + ((g * limbs + l) * (uint64_t)c.n) * 8;
            for (uint32_t i = 0; i < c.n; i++) {
                const uint64_t v = get64(row + (uint64_t)i * 8);
                RCK_REQUIRE(v < q, "serialized residue " + std::to_string(v) + " at limb " + std::to_string(l) +
                                       " is not below its modulus");
            }
        }
    }
}

}  // namespace
}  // namespace rck

using namespace rck;

extern "C" {

int rck_serialize_polys(rck_context* ctx, const uint64_t* polys, uint32_t count, uint32_t nq, uint32_t np,
                        uint32_t scale_bits, uint8_t* buf, size_t len, size_t* needed) {
    return zguard([&] {
        RCK_REQUIRE(ctx && ctx->c, "null context");
        Context& c = *ctx->c;
        RCK_REQUIRE(nq >= 1 && nq <= c.L && (np == 0 || np == c.K), "limb counts do not fit the context");
        RCK_REQUIRE(count >= 1, "nothing to serialize");
        const uint64_t payload = (uint64_t)count * (nq + np) * c.n * 8;
        if (needed) *needed = kHeader + payload;
        if (!buf) return;
        RCK_REQUIRE(polys != nullptr, "null polynomials");
        RCK_REQUIRE(len >= kHeader + payload, "serialization buffer too small");
        write_header(buf, Header{1, c.logn, count, nq, np, scale_bits, 0, fingerprint(c, nq, np), payload});
        RCK_CHECK_HIP(hipMemcpy(buf + kHeader, polys, payload, hipMemcpyDeviceToHost));
    });
}

int rck_deserialize_polys(rck_context* ctx, const uint8_t* buf, size_t len, uint64_t* polys, size_t capacity,
                          uint32_t* count, uint32_t* nq, uint32_t* np, uint32_t* scale_bits) {
    return zguard([&] {
        RCK_REQUIRE(ctx && ctx->c, "null context");
        const Header h = read_header(*ctx->c, buf, len, 1);
        if (count) *count = h.count;
        if (nq) *nq = h.nq;
        if (np) *np = h.np;
        if (scale_bits) *scale_bits = h.scale;
        if (!polys) return;
        RCK_REQUIRE(capacity >= h.payload, "destination buffer too small");
        check_canonical(*ctx->c, buf + kHeader, h.count, h.nq, h.np);
        RCK_CHECK_HIP(hipMemcpy(polys, buf + kHeader, h.payload, hipMemcpyHostToDevice));
    });
}

int rck_ksk_serialize(rck_context* ctx, const rck_ksk* key, uint8_t* buf, size_t len, size_t* needed) {
    return zguard([&] {
        RCK_REQUIRE(ctx && ctx->c && key, "null argument");
        Context& c = *ctx->c;
        const uint32_t nq = key->level + 1, np = key->aux;
        const uint64_t payload = (uint64_t)key->digits * 2 * (nq + np) * c.n * 8;
        if (needed) *needed = kHeader + payload;
        if (!buf) return;
        RCK_REQUIRE(len >= kHeader + payload, "serialization buffer too small");
        write_header(buf, Header{2, c.logn, key->digits, nq, np, key->level, key->galois, fingerprint(c, nq, np),
                                 payload});
        RCK_REQUIRE(rck_ksk_export(ctx, key, reinterpret_cast<uint64_t*>(buf + kHeader)) == RCK_OK,
                    last_error());
    });
}

int rck_ksk_deserialize(rck_context* ctx, const uint8_t* buf, size_t len, rck_ksk** out) {
    return zguard([&] {
        RCK_REQUIRE(ctx && ctx->c && out, "null argument");
        const Header h = read_header(*ctx->c, buf, len, 2);
        RCK_REQUIRE(h.scale == h.nq - 1, "serialized key level does not match its limb count");
        check_canonical(*ctx->c, buf + kHeader, (uint64_t)h.count * 2, h.nq, h.np);
        // the payload is the crate's coefficient form: upload through rck_ksk_create (host memory;
        // the payload is 8-byte aligned when buf is)
        std::vector<uint64_t> tmp;
        const uint64_t* src = reinterpret_cast<const uint64_t*>(buf + kHeader);
        if (reinterpret_cast<uintptr_t>(src) % alignof(uint64_t)) {
            tmp.resize(h.payload / 8);
            std::memcpy(tmp.data(), buf + kHeader, h.payload);
            src = tmp.data();
        }
        RCK_REQUIRE(rck_ksk_create(ctx, src, h.count, h.nq - 1, h.np, h.galois, out) == RCK_OK, last_error());
    });
}

}  // extern "C"
