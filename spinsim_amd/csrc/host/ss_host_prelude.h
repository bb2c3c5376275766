// Host build of the integration path for the CPU devices (Device.CPU,
// CPU_SINGLE, PYTHON): the definitions that let the device headers
// (ss_args.h, ss_complex.h, ss_prelude.h, ss_integrator.h) compile as plain
// C++17 under g++.  Included first; SS_HOST makes ss_integrator.h stop
// before its GPU kernels.
#pragma once

#include <math.h>
#include <stdint.h>
#include <string.h>

#define SS_HOST 1
#define __device__
#define __host__
#define __forceinline__ inline __attribute__((always_inline))
#define __constant__

struct double2 {
    double x, y;
};
static inline double2 make_double2(double x, double y)
{
    double2 r;
    r.x = x;
    r.y = y;
    return r;
}
