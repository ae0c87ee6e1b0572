// Paterson-Stockmeyer polynomial evaluation (src/algorithms/paterson_stockmeyer.rs:28-350) as a
// straight-line program over ciphertext registers, shared by the per-request entry point and the
// request scheduler.
//
// Every plaintext the crate's algorithm builds is the encoding of a constant slot vector
// (constant_plaintext, constant_slots, paterson_stockmeyer.rs:325-336), whose canonical-embedding
// inverse is exactly the constant polynomial round(c * 2^s) (the crate's FFT keeps every non-DC
// entry at exactly zero for a constant input).  Multiplying by such a plaintext is therefore the
// scalar product by round(c * 2^s) mod q_i on every limb, and adding it touches coefficient 0
// only: the program uses those exact shortcuts.
#pragma once
#include <string>
#include <vector>

#include "pipeline.hpp"

namespace rck {

enum PsKind {
    kPsMultRelinRescale = 0,  // dst = mult_relinearized(a, b, key)        (context/ops.rs:124-133)
    kPsScalarMul = 1,         // dst = a[..limbs] * k_i on limb i           (multiply_plain_rns by a constant)
    kPsAddConst = 2,          // dst = a; dst.c0[i][0] += k_i               (add_constant_term, :314-323)
    kPsAdd = 3,               // dst = a + b                                (RnsCiphertext::add)
    kPsTruncate = 4,          // dst = a[..limbs]                           (truncate_levels)
    kPsZero = 5,              // dst = 0
};

struct PsInstr {
    int kind = 0;
    int dst = -1, a = -1, b = -1;
    int limbs = 0;             // limbs of dst (and of the operation)
    int key_level = -1;        // relinearization key (MultRelinRescale)
    std::vector<uint64_t> k;   // per-limb constants (ScalarMul, AddConst)
};

struct PsProgram {
    // registers: limbs and scale of the ciphertext each holds; reg 0 is the input
    std::vector<int> reg_limbs;
    std::vector<uint32_t> reg_scale;
    std::vector<PsInstr> code;
    int result = 0;
    // physical buffers after liveness-based reuse: slot of every register, capacity (limbs) per slot
    std::vector<int> slot_of;
    std::vector<int> slot_limbs;
    // device bytes of all slots except the input's (the request's working set)
    size_t footprint(uint32_t n) const;
    int mults() const;
};

// paterson_stockmeyer_block_size (:18-26)
int ps_block_size(int term_count);
// paterson_stockmeyer_required_levels (:108-126); throws the crate's message when the degree is too large
std::vector<int> ps_required_levels(int ciphertext_level, int term_count);
// Builds the program for one evaluation; has_key[level] says which relinearization keys exist.
// Throws the crate's panic messages for the same conditions.
PsProgram ps_compile(const Context& c, int limbs, uint32_t scale_bits, const double* coeffs, int count,
                     const std::vector<bool>& has_key);

// Runs instruction `ins` of a program whose registers live at regs[r] (2 polys of reg_limbs[r] limbs
// each).  keys[level] are the relinearization keys.
void ps_exec(Context& c, const PsProgram& p, const PsInstr& ins, const std::vector<uint64_t*>& regs,
             const std::vector<const KeyDev*>& keys, hipStream_t s);

// encode_integer_coeffs + residues of the first `limbs` primes into host memory out[limbs][N]
// (threads: split the coefficient loop over host threads)
void encode_host(const Context& c, const double* re, const double* im, uint32_t limbs, uint32_t scale_bits,
                 uint64_t* out, bool threads);

// round(c * 2^s) mod q_i for the first `limbs` primes (the encoding of a constant slot vector)
std::vector<uint64_t> const_residues(const Context& c, double value, uint32_t scale_bits, int limbs);

}  // namespace rck
