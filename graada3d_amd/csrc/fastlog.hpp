// Natural logarithm of a positive normal double by table reduction (Tang's
// method with left table points).  v = 2^e m, m in [1, 2); for the top 7
// mantissa bits j, F_j = 1 + j/128, G_j = 1/F_j (rounded) and
// T_j = -ln G_j (log_table.hpp):
//     ln v = e ln2 + T_j + ln(1 + r),  r = m G_j - 1 in (-2^-52, 2^-7),
// ln(1 + r) by its Taylor series to r^8 (truncation below r^9/9 < 2^-66);
// ln2 split as fdlibm's ln2_hi + ln2_lo so that e ln2_hi is exact.  Within
// 1.5 ulp of the exact value for |ln v| >= 0.5 (1.002 ulp measured) and
// within 1.2e-16 absolutely on [0.5, 2]; about a quarter of the instructions of the
// general library log.  Arguments are clamped to [2^-1000, 2^1000]: callers
// pass distances or sums of distances, far inside that range.
#pragma once

#include <math.h>
#include <stdint.h>

#ifndef GRA_HD
#define GRA_HD __host__ __device__
#endif

#include "log_table.hpp"

namespace gra {

constexpr int kLogBits = 7;

// Series coefficients 1/8, 1/7, -1/6, 1/5, -1/4, 1/3 and ln2 as hi + lo.  In
// device code they sit in constant memory, so that each use reads them with
// scalar loads into SGPRs instead of materialising every one in VGPRs with
// two moves per use (a quarter of the log's instructions otherwise).
constexpr double kLogCoef[8] = {1.0 / 8.0, 1.0 / 7.0, -1.0 / 6.0, 1.0 / 5.0, -0.25, 1.0 / 3.0,
                                6.93147180369123816490e-01, 1.90821492927058770002e-10};
#if defined(__HIP_DEVICE_COMPILE__)
static __constant__ double kLogC[8] = {kLogCoef[0], kLogCoef[1], kLogCoef[2], kLogCoef[3],
                                kLogCoef[4], kLogCoef[5], kLogCoef[6], kLogCoef[7]};
#endif

GRA_HD inline double table_log(double v, const double *G, const double *T)
{
    v = fmin(fmax(v, 9.33263618503218879e-302), 1.07150860718626732e301);
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const int e = (int)(b >> 52) - 1023;
    const int j = (int)((b >> (52 - kLogBits)) & ((1u << kLogBits) - 1u));
    const double m = __builtin_bit_cast(double, (b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    const double r = fma(m, G[j], -1.0);
#if defined(__HIP_DEVICE_COMPILE__)
    const double *C = kLogC;
#else
    const double *C = kLogCoef;
#endif
    double p = C[0];
    p = fma(p, -r, C[1]);
    p = fma(p, r, C[2]);
    p = fma(p, r, C[3]);
    p = fma(p, r, C[4]);
    p = fma(p, r, C[5]);
    p = fma(p, r, -0.5);
    p = fma(p * r, r, r);
    const double ed = (double)e;
    return fma(ed, C[6], T[j]) + fma(ed, C[7], p);
}

// table_log with the constant-memory table.
GRA_HD inline double fast_log(double v) { return table_log(v, kLogG, kLogT); }

}  // namespace gra
