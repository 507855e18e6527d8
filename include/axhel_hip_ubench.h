/*
 * axhel_hip_ubench.h -- measurement hooks of the gfx950 library (no reference
 * counterpart: they define the ceilings the kernels are reported against).
 */
#ifndef AXHEL_HIP_UBENCH_H
#define AXHEL_HIP_UBENCH_H

#include "axhel_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Rate of register-resident forward Harvey butterflies (the NTT kernels'
   exact arithmetic, no memory traffic) over the whole device, and the timed
   kernel's duration. */
axhel_hip_status axhel_hip_ubench_butterfly_peak(double *butterflies_per_second, double *seconds);

#ifdef __cplusplus
}
#endif

#endif /* AXHEL_HIP_UBENCH_H */
