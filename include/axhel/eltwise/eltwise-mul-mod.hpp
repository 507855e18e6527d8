// EltwiseMulMod, same declaration as the reference's
// axhel/include/axhel/eltwise/eltwise-mul-mod.hpp:14.
#pragma once

#include <stdint.h>

namespace unipi {
namespace axhel {

// result[i] = operand1[i] * operand2[i] mod modulus for inputs in
// [0, input_mod_factor * modulus); input_mod_factor is 1, 2 or 4 (any other
// value leaves result untouched).
void EltwiseMulMod(uint64_t *result, const uint64_t *operand1, const uint64_t *operand2, uint64_t n, uint64_t modulus,
                   uint64_t input_mod_factor);

}  // namespace axhel
}  // namespace unipi
