// NTT configurations for 2^16, indices 0-4 (see ntt_registry.hpp).
#include "ntt_registry.hpp"

namespace axhel_host {

void add_configs_16a(Registry &r) {
    r[16].insert(r[16].end(), {two<8, 4, 16, 8, 4, 16>(), two<8, 4, 8, 8, 4, 8>(), two<8, 3, 16, 8, 3, 8>(),
                               two<8, 3, 16, 8, 4, 8>(), two<8, 4, 16, 8, 3, 8>()});
}

}  // namespace axhel_host
