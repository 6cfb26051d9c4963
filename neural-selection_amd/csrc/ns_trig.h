// Sine and cosine of a float angle for the world step (ray casting and
// movement), identical on the device and in a host restatement: Cody-Waite
// reduction by pi/2 in double (fdlibm's three-part pi/2) and Taylor
// polynomials in Horner form on [-pi/4, pi/4], every step an IEEE double
// operation (no contraction), rounded once to float.  Within one float ulp of
// the true values for |angle| < 2^20.
#pragma once
#include <hip/hip_runtime.h>

namespace nsa {

__host__ __device__ inline void world_sincos(float angle, float *s, float *c) {
    const double x = (double)angle;
    const double k = __builtin_rint(x * 0.63661977236758134308);
    double r = x - k * 1.57079632673412561417e+00;
    r = r - k * 6.07710050650619224932e-11;
    r = r - k * 2.02226624879595063154e-21;
    const double r2 = r * r;
    const double sr = r + r * r2 * (-1.0 / 6 + r2 * (1.0 / 120 + r2 * (-1.0 / 5040 + r2 * (1.0 / 362880 +
                      r2 * (-1.0 / 39916800 + r2 * (1.0 / 6227020800.0))))));
    const double cr = 1.0 + r2 * (-0.5 + r2 * (1.0 / 24 + r2 * (-1.0 / 720 + r2 * (1.0 / 40320 +
                      r2 * (-1.0 / 3628800 + r2 * (1.0 / 479001600.0))))));
    const long long q = (long long)k & 3;
    const double sv = q == 0 ? sr : q == 1 ? cr : q == 2 ? -sr : -cr;
    const double cv = q == 0 ? cr : q == 1 ? -sr : q == 2 ? -cr : sr;
    *s = (float)sv;
    *c = (float)cv;
}

}  // namespace nsa
