// Negacyclic NTT, same class interface as the reference's
// axhel/include/axhel/ntt/ntt.hpp:14-65 (pImpl over a shared implementation,
// shallow copies).  The implementation keeps the twiddle tables resident in
// GPU memory; operands may be host or device pointers.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <memory>

namespace unipi {
namespace axhel {

// Twiddle operand with its Shoup quotient floor(operand * 2^64 / modulus).
struct NTTMultiplyOperand {
    uint64_t operand;
    uint64_t quotient;
};

class NTT {
public:
    // Generates and owns the tables for degree (a power of two), modulus
    // (prime, == 1 mod 2*degree) and a primitive 2*degree-th root of unity.
    NTT(uint64_t degree, uint64_t modulus, uint64_t root_of_unity);

    // Uses caller tables (degree entries each); they are copied to the device
    // once, so the caller's arrays need not outlive the call.
    NTT(uint64_t degree, uint64_t modulus, const NTTMultiplyOperand *root_powers,
        const NTTMultiplyOperand *inv_root_powers, NTTMultiplyOperand inv_degree_modulo);

    ~NTT();

    NTT(const NTT &) noexcept;
    NTT &operator=(const NTT &) noexcept;

    NTT(NTT &&) noexcept;
    NTT &operator=(NTT &&) noexcept;

    // Forward transform; output_mod_factor 4 -> lazy [0,4q) output, else [0,q).
    // result and operand may be the same buffer.
    void ComputeForward(uint64_t *result, const uint64_t *operand, uint64_t input_mod_factor = 1,
                        uint64_t output_mod_factor = 1) const;

    // Inverse transform; output_mod_factor 2 -> lazy [0,2q) output, else [0,q).
    void ComputeInverse(uint64_t *result, const uint64_t *operand, uint64_t input_mod_factor = 1,
                        uint64_t output_mod_factor = 1) const;

    uint64_t Degree() const noexcept;

    uint64_t Modulus() const noexcept;

private:
    class Impl;

    std::shared_ptr<Impl> impl_;
};

}  // namespace axhel
}  // namespace unipi
