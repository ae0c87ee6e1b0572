// Allocation-free negacyclic polynomial product for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops): the NTT forms of both operands live in a
// fixed file-scope device array instead of the context's grow-on-demand workspace.
//
// The array is kNoAllocScratchBytes (64 MiB) on every GPU the library is loaded on, held until the
// process ends.  Its device address is looked up once per device with hipGetSymbolAddress (at
// context creation, or at the first call on a device) and only that address is ever handed to
// kernels.  Every call on one GPU shares that GPU's array: calls that overlap on one GPU (two host
// threads, or two streams without ordering) race on it and must be serialized by the caller.
#include <mutex>

#include "../../include/rnsckks.h"
#include "capi_types.hpp"

namespace rck {
namespace {

constexpr size_t kNoAllocScratchBytes = 64ull << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

std::mutex g_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// device address of this device's scratch array (looked up once per device, with it current)
void* scratch_for_current_device() {
    int dev = 0;
    RCK_CHECK_HIP(hipGetDevice(&dev));
    RCK_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index out of range for the allocation-free scratch");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        RCK_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        RCK_REQUIRE(p != nullptr, "allocation-free scratch has no device address");
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

template <typename F>
int nguard(F&& f) {
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

}  // namespace

// set-up at context creation: the current device's scratch address
void noalloc_setup_current_device() { (void)scratch_for_current_device(); }

}  // namespace rck

using namespace rck;

extern "C" {

int rck_poly_multiply_ntt_noalloc(rck_context* ctx, const uint64_t* a, const uint64_t* b, uint64_t* out,
                                  uint32_t nq, uint32_t np, void* stream) {
    return nguard([&] {
        RCK_REQUIRE(ctx && ctx->c, "null context");
        Context& c = *ctx->c;
        require_fp64(c);
        RCK_REQUIRE(nq >= 1 && nq <= c.L, "limb count must be in 1..=" + std::to_string(c.L) + ", got " +
                                              std::to_string(nq));
        RCK_REQUIRE(np == 0 || np == c.K, "aux limb count must be 0 or the aux basis size");
        RCK_REQUIRE(a && b && out, "null polynomial");
        const int limbs = (int)(nq + np);
        const long long P = (long long)limbs * c.n;
        const size_t need = poly_mul_scratch_bytes(c, P, limbs, 1);
        RCK_REQUIRE(need <= kNoAllocScratchBytes,
                    "allocation-free product needs " + std::to_string(need) + " scratch bytes (2 x limbs x N x 8), "
                    "above the fixed " + std::to_string(kNoAllocScratchBytes) + "-byte scratch");
        double* t = static_cast<double*>(scratch_for_current_device());
        poly_mul(c, a, P, b, P, out, P, limbs, 1, PrimeMap{(int)nq, (int)c.L, 0}, static_cast<hipStream_t>(stream),
                 t);
    });
}

}  // extern "C"
