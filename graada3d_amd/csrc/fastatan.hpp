// atan(n / d) (principal value) with one division and no library call:
// with a = |n|, b = |d|, t = min(a, b) / max(a, b) in [0, 1] and
// atan(t) = t P(t^2) (atan_poly.hpp, degree 20, within 1 ulp on [0, 1]);
// a > b takes pi/2 - atan(t) (pi/2 as hi + lo); the sign is that of n / d,
// signed zeros included (d = +-0 gives +-pi/2 like atan(n / +-0)).  NaN and
// 0/0, inf/inf give NaN as n / d does.  About 40 % fewer instructions than a
// division followed by the library atan, which divides again for |x| > 1.
#pragma once

#include <math.h>

#ifndef GRA_HD
#define GRA_HD __host__ __device__
#endif

#include "atan_poly.hpp"

namespace gra {

#if defined(__HIP_DEVICE_COMPILE__)
// The coefficients in constant memory for device code: each use reads them
// with scalar loads into SGPRs, where the FMAs take them directly; as
// literals the compiler materialised them in VGPRs with two moves per
// coefficient and use (about 13 % of the assembly kernels' instructions).
static __constant__ double kAtanPc[kAtanDeg + 1] = {
    kAtanP[0],  kAtanP[1],  kAtanP[2],  kAtanP[3],  kAtanP[4],  kAtanP[5],  kAtanP[6],
    kAtanP[7],  kAtanP[8],  kAtanP[9],  kAtanP[10], kAtanP[11], kAtanP[12], kAtanP[13],
    kAtanP[14], kAtanP[15], kAtanP[16], kAtanP[17], kAtanP[18], kAtanP[19], kAtanP[20]};
static_assert(kAtanDeg == 20, "one initialiser per coefficient");
#endif

GRA_HD inline double atan_ratio(double n, double d)
{
    const double a = fabs(n), b = fabs(d);
    const bool swap = a > b;
    const double t = (swap ? b : a) / (swap ? a : b);
    const double s = t * t;
#if defined(__HIP_DEVICE_COMPILE__)
    const double *P = kAtanPc;
#else
    const double *P = kAtanP;
#endif
    double p = P[kAtanDeg];
#pragma unroll
    for (int k = kAtanDeg - 1; k >= 0; --k) p = fma(p, s, P[k]);
    double r = t * p;
    if (swap) r = (1.5707963267948966 - r) + 6.123233995736766e-17;
    return copysign(r, n) * copysign(1.0, d);
}

}  // namespace gra
