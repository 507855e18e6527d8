// NTT configurations for 2^16, indices 5-9 (see ntt_registry.hpp); 9
// spreads a single limb over 128 workgroups of 64 threads per pass (single
// calls).
#include "ntt_registry.hpp"

namespace axhel_host {

void add_configs_16b(Registry &r) {
    r[16].insert(r[16].end(), {two<8, 3, 8, 8, 3, 8>(), two<8, 3, 16, 8, 3, 16>(), two<7, 3, 16, 9, 3, 4>(),
                               two<9, 3, 8, 7, 3, 16>(), two<7, 3, 4, 9, 3, 1>()});
}

}  // namespace axhel_host
