// EltwiseReduceMod, same declaration as the reference's
// axhel/include/axhel/eltwise/eltwise-reduce-mod.hpp:14.
#pragma once

#include <stdint.h>

namespace unipi {
namespace axhel {

// res[i] = op[i] reduced from [0, in_mod_factor * mod) to
// [0, out_mod_factor * mod); in_mod_factor == mod reduces any 64-bit word.
void EltwiseReduceMod(uint64_t *res, const uint64_t *op, uint64_t n, uint64_t mod, uint64_t in_mod_factor,
                      uint64_t out_mod_factor);

}  // namespace axhel
}  // namespace unipi
