// Umbrella header of the AXHEL drop-in API (reference: axhel/include/axhel/axhel.hpp:6-12).
// Implemented by axhel_amd/lib/libaxhel.so on top of the gfx950 C ABI in axhel_hip.h.
#pragma once

#include "axhel/eltwise/eltwise-add-mod.hpp"
#include "axhel/eltwise/eltwise-fma-mod.hpp"
#include "axhel/eltwise/eltwise-mul-mod.hpp"
#include "axhel/eltwise/eltwise-reduce-mod.hpp"
#include "axhel/eltwise/eltwise-sub-mod.hpp"
#include "axhel/ntt/ntt.hpp"
