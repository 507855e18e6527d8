// NTT configurations for 2^1 .. 2^13 (see ntt_registry.hpp).
#include "ntt_registry.hpp"

namespace axhel_host {

void add_configs_1_13(Registry &r) {
    r[1] = {single<1, 1, 256>()};
    r[2] = {single<2, 2, 256>()};
    r[3] = {single<3, 3, 256>()};
    r[4] = {single<4, 4, 256>()};
    r[5] = {single<5, 4, 128>(), single<5, 3, 64>()};
    r[6] = {single<6, 4, 64>(), single<6, 3, 32>()};
    r[7] = {single<7, 4, 32>(), single<7, 3, 16>()};
    r[8] = {single<8, 4, 16>(), single<8, 3, 8>()};
    r[9] = {single<9, 4, 8>(), single<9, 3, 4>()};
    r[10] = {single<10, 4, 4>(), single<10, 3, 2>()};
    r[11] = {single<11, 4, 2>(), single<11, 3, 1>()};
    // two-pass forms of 2^12 and 2^13: more, shorter workgroups, for single calls
    r[12] = {single<12, 4, 1>(), single<12, 3, 1>(), two<6, 3, 8, 6, 3, 8>()};
    r[13] = {single<13, 4, 1>(), single<13, 3, 1>(), two<6, 3, 16, 7, 3, 16>(), two<6, 3, 8, 7, 3, 8>()};
}

}  // namespace axhel_host
