// NTT configurations for 2^17 (see ntt_registry.hpp); index 4 spreads a
// single limb over small workgroups (single calls).
#include "ntt_registry.hpp"

namespace axhel_host {

void add_configs_17(Registry &r) {
    r[17] = {two<9, 3, 8, 8, 4, 16>(), two<8, 4, 16, 9, 3, 8>(), two<9, 3, 8, 8, 3, 8>(),
             two<8, 3, 16, 9, 3, 4>(), two<8, 3, 4, 9, 3, 1>()};
}

}  // namespace axhel_host
