// NTT configurations for 2^20 (see ntt_registry.hpp): two passes of 10 bits.
#include "ntt_registry.hpp"

namespace axhel_host {

void add_configs_20(Registry &r) {
    r[20] = {two<10, 4, 4, 10, 4, 4>(), two<10, 4, 8, 10, 4, 4>(), two<10, 4, 16, 10, 4, 4>(),
             two<10, 3, 8, 10, 3, 4>(), two<10, 4, 8, 10, 3, 4>()};
}

}  // namespace axhel_host
