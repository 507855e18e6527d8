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

/* Butterfly schedules of the NTT pass kernels.  One iteration of a schedule
   is the stage sequence of one 8-bit pass:
     FWD_EXACT          8 forward Harvey butterflies (ntt-internal.hpp:20-31):
                        every forward pass of a lazy transform, and of any
                        transform whose modulus is 2^60 or more;
     FWD_RELAXED        relaxed forward schedule of a first pass: six stages
                        without subtraction, one subtracting 8q, one without:
                        the COL pass of a normalized forward transform;
     FWD_RELAXED_LAST   relaxed schedule of a last pass: 8q subtracted at its
                        stages 0 and 4, bit 0 on the final-stage form and both
                        words normalized to [0, q): the ROW pass of a
                        normalized forward transform;
     INV                8 inverse butterflies (ntt-internal.hpp:37-49): the
                        inverse ROW pass;
     INV_FINAL          7 inverse butterflies then the scaled final stage
                        (ntt-native.cpp:126-141): the inverse COL pass, lazy;
     INV_FINAL_NORM     the same with the [0, 2q) -> [0, q) normalization. */
enum {
    AXHEL_HIP_UBENCH_FWD_EXACT = 0,
    AXHEL_HIP_UBENCH_FWD_RELAXED = 1,
    AXHEL_HIP_UBENCH_FWD_RELAXED_LAST = 2,
    AXHEL_HIP_UBENCH_INV = 3,
    AXHEL_HIP_UBENCH_INV_FINAL = 4,
    AXHEL_HIP_UBENCH_INV_FINAL_NORM = 5,
    AXHEL_HIP_UBENCH_SCHEDULES = 6
};

/* Rate of register-resident butterflies of one schedule (the NTT kernels'
   exact arithmetic, no memory traffic) over the whole device, and the timed
   kernel's duration. */
axhel_hip_status axhel_hip_ubench_schedule_peak(int schedule, double *butterflies_per_second, double *seconds);

/* The FWD_EXACT schedule. */
axhel_hip_status axhel_hip_ubench_butterfly_peak(double *butterflies_per_second, double *seconds);

#ifdef __cplusplus
}
#endif

#endif /* AXHEL_HIP_UBENCH_H */
