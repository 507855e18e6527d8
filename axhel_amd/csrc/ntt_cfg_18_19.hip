// NTT configurations for 2^18 and 2^19 (see ntt_registry.hpp): two passes
// of 9 or 10 bits.  A COL pass of 10 bits holds 8 KiB per column, so 8 or 16
// columns per workgroup (64 or 128 contiguous bytes per row access) fill
// most of a CU's LDS; 4 columns (32 bytes per access) ran at half the rate.
#include "ntt_registry.hpp"

namespace axhel_host {

void add_configs_18_19(Registry &r) {
    r[18] = {two<9, 3, 8, 9, 3, 8>(), two<9, 4, 8, 9, 4, 8>(), two<9, 3, 16, 9, 3, 8>()};
    r[19] = {two<10, 4, 4, 9, 3, 8>(), two<9, 3, 8, 10, 4, 4>(), two<10, 4, 8, 9, 3, 8>(),
             two<9, 3, 16, 10, 3, 4>()};
}

}  // namespace axhel_host
