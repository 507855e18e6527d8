// NTT configurations for 2^14 and 2^15 (see ntt_registry.hpp); index 4 of
// each spreads a single limb over small workgroups (single calls).
#include "ntt_registry.hpp"

namespace axhel_host {

void add_configs_14_15(Registry &r) {
    r[14] = {two<7, 4, 32, 7, 4, 32>(), two<7, 4, 16, 7, 4, 16>(), two<7, 3, 16, 7, 3, 16>(),
             two<7, 3, 32, 7, 3, 16>(), two<7, 3, 4, 7, 3, 4>()};
    r[15] = {two<8, 4, 16, 7, 4, 32>(), two<7, 4, 32, 8, 4, 16>(), two<8, 3, 16, 7, 3, 16>(),
             two<7, 3, 16, 8, 3, 8>(), two<7, 3, 4, 8, 3, 2>()};
}

}  // namespace axhel_host
