// Exact single-precision math for the device tick path.
//
// The reference's arithmetic (src/physiology.cuh) calls the host libm: glibc
// 2.35's expf on x86-64 with FMA.  That expf is not correctly rounded, so a
// device expf must reproduce glibc's own evaluation to keep spike rasters
// bit-exact.  ns_expf follows the published ARM optimized-routines expf that
// glibc ships (MIT licensed): x*N/ln2 = k + r with N = 32 and |r| <= 1/2,
// exp(x) = 2^(k/N) * 2^(r/N), 2^(r/N) by a degree-3 polynomial in double,
// one final rounding to float.  The FMA build contracts kd = x*N/ln2 + shift
// and r = x*N/ln2 - kd; that form agrees with this image's libm expf on all
// 2^32 inputs (checked exhaustively on the host, and against the device in
// tests/test_gpu_math.py).
//
// Compile this code with -ffp-contract=off: every other float expression on
// the tick path must round exactly as the host's uncontracted SSE code does.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP_DEVICE_COMPILE__)
#define NS_HOST_DEVICE __host__ __device__
#else
#define NS_HOST_DEVICE
#endif

namespace nsa {

// 2^(i/32) as the nearest double, minus i << 47 (so that adding k << 47
// restores the exponent).  Generated with exp2l and rounded to double.
#define NS_EXP2_TABLE                                                                     \
    {0x3ff0000000000000ull, 0x3fefd9b0d3158574ull, 0x3fefb5586cf9890full,                  \
     0x3fef9301d0125b51ull, 0x3fef72b83c7d517bull, 0x3fef54873168b9aaull,                  \
     0x3fef387a6e756238ull, 0x3fef1e9df51fdee1ull, 0x3fef06fe0a31b715ull,                  \
     0x3feef1a7373aa9cbull, 0x3feedea64c123422ull, 0x3feece086061892dull,                  \
     0x3feebfdad5362a27ull, 0x3feeb42b569d4f82ull, 0x3feeab07dd485429ull,                  \
     0x3feea47eb03a5585ull, 0x3feea09e667f3bcdull, 0x3fee9f75e8ec5f74ull,                  \
     0x3feea11473eb0187ull, 0x3feea589994cce13ull, 0x3feeace5422aa0dbull,                  \
     0x3feeb737b0cdc5e5ull, 0x3feec49182a3f090ull, 0x3feed503b23e255dull,                  \
     0x3feee89f995ad3adull, 0x3feeff76f2fb5e47ull, 0x3fef199bdd85529cull,                  \
     0x3fef3720dcef9069ull, 0x3fef5818dcfba487ull, 0x3fef7c97337b9b5full,                  \
     0x3fefa4afa2a490daull, 0x3fefd0765b6e4540ull}

constexpr double kInvLn2N = 0x1.71547652b82fep+5;     // 32/ln2
constexpr double kShift = 0x1.8p+52;
constexpr double kC0 = 0x1.c6af84b912394p-20;         // poly / N^3
constexpr double kC1 = 0x1.ebfce50fac4f3p-13;         // poly / N^2
constexpr double kC2 = 0x1.62e42ff0c52d6p-6;          // poly / N

NS_HOST_DEVICE inline uint32_t f2u(float f) {
    union { float f; uint32_t u; } c; c.f = f; return c.u;
}
NS_HOST_DEVICE inline uint64_t d2u(double d) {
    union { double d; uint64_t u; } c; c.d = d; return c.u;
}
NS_HOST_DEVICE inline double u2d(uint64_t u) {
    union { uint64_t u; double d; } c; c.u = u; return c.d;
}

// ns_expf for |x| < 88 (no overflow, underflow, infinity or nan to handle):
// the same evaluation as ns_expf on those inputs.  `tab` points at
// NS_EXP2_TABLE (LDS or global).
NS_HOST_DEVICE inline float ns_expf_finite(float x, const uint64_t *tab) {
    double xd = (double)x;
    double kd = __builtin_fma(kInvLn2N, xd, kShift);
    uint64_t ki = d2u(kd);
    kd -= kShift;
    double r = __builtin_fma(kInvLn2N, xd, -kd);
    uint64_t t = tab[ki & 31];
    t += ki << 47;
    double s = u2d(t);
    double z = __builtin_fma(kC0, r, kC1);
    double r2 = r * r;
    double y = __builtin_fma(kC2, r, 1.0);
    y = __builtin_fma(z, r2, y);
    y = y * s;
    return (float)y;
}

// `tab` points at NS_EXP2_TABLE (LDS or global).
NS_HOST_DEVICE inline float ns_expf(float x, const uint64_t *tab) {
    uint32_t ux = f2u(x);
    uint32_t abstop = (ux >> 20) & 0x7ff;
    if (abstop >= 0x42b) {                       // |x| >= 88 or nan
        if (ux == 0xff800000u) return 0.0f;      // -inf
        if (abstop >= 0x7f8) return x + x;       // inf / nan
        if (x > 0x1.62e42ep6f) return __builtin_inff();
        if (x < -0x1.9fe368p6f) return 0.0f;
    }
    return ns_expf_finite(x, tab);
}

// logf: the published ARM optimized-routines logf that glibc ships.  x =
// 2^k z with z in [0x1.6p-1, 0x1.6p+0) split into 16 subintervals of centre
// c; log(x) = k ln2 + log(c) + log1p(z/c - 1) with a degree-4 polynomial in
// double.  Table pairs {1/c, log c}.  Agrees with this image's libm logf on
// all 2^31 non-negative inputs (host check; tests/test_gpu_math.py on the
// device).
#define NS_LOGF_TABLE                                                                     \
    {0x1.661ec79f8f3bep+0, -0x1.57bf7808caadep-2, 0x1.571ed4aaf883dp+0,                    \
     -0x1.2bef0a7c06ddbp-2, 0x1.49539f0f010bp+0, -0x1.01eae7f513a67p-2,                    \
     0x1.3c995b0b80385p+0, -0x1.b31d8a68224e9p-3, 0x1.30d190c8864a5p+0,                    \
     -0x1.6574f0ac07758p-3, 0x1.25e227b0b8eap+0, -0x1.1aa2bc79c81p-3,                      \
     0x1.1bb4a4a1a343fp+0, -0x1.a4e76ce8c0e5ep-4, 0x1.12358f08ae5bap+0,                    \
     -0x1.1973c5a611cccp-4, 0x1.0953f419900a7p+0, -0x1.252f438e10c1ep-5,                   \
     0x1p+0, 0x0p+0, 0x1.e608cfd9a47acp-1, 0x1.aa5aa5df25984p-5,                          \
     0x1.ca4b31f026aap-1, 0x1.c5e53aa362eb4p-4, 0x1.b2036576afce6p-1,                      \
     0x1.526e57720db08p-3, 0x1.9c2d163a1aa2dp-1, 0x1.bc2860d22477p-3,                      \
     0x1.886e6037841edp-1, 0x1.1058bc8a07ee1p-2, 0x1.767dcf5534862p-1,                     \
     0x1.4043057b6ee09p-2}

constexpr double kLn2 = 0x1.62e42fefa39efp-1;
constexpr double kLogA0 = -0x1.00ea348b88334p-2;
constexpr double kLogA1 = 0x1.5575b0be00b6ap-2;
constexpr double kLogA2 = -0x1.ffffef20a4123p-2;

NS_HOST_DEVICE inline float u2f(uint32_t u) {
    union { uint32_t u; float f; } c; c.u = u; return c.f;
}

// `tab` points at NS_LOGF_TABLE (32 doubles: invc, logc pairs).
NS_HOST_DEVICE inline float ns_logf(float x, const double *tab) {
    uint32_t ix = f2u(x);
    if (ix == 0x3f800000u) return 0.0f;
    if (ix - 0x00800000u >= 0x7f800000u - 0x00800000u) {
        if (ix * 2 == 0) return -__builtin_inff();
        if (ix == 0x7f800000u) return x;
        if ((ix & 0x80000000u) || ix * 2 >= 0xff000000u) return __builtin_nanf("");
        ix = f2u(x * 0x1p23f);                   // subnormal: normalise
        ix -= 23u << 23;
    }
    uint32_t tmp = ix - 0x3f330000u;
    int i = (tmp >> 19) & 15;
    int k = (int32_t)tmp >> 23;
    uint32_t iz = ix - (tmp & (0x1ffu << 23));
    double invc = tab[2 * i], logc = tab[2 * i + 1];
    double z = (double)u2f(iz);
    double r = __builtin_fma(z, invc, -1.0);
    double y0 = __builtin_fma((double)k, kLn2, logc);
    double r2 = r * r;
    double y = __builtin_fma(kLogA1, r, kLogA2);
    y = __builtin_fma(kLogA0, r2, y);
    y = __builtin_fma(y, r2, y0 + r);
    return (float)y;
}

// Decay and offset constants the host compiler folds from constant libm
// calls (values equal to glibc's at run time; pinned by tests/test_constants.py).
constexpr float kDecayExcQuarter = 0x1.d7101p-1f;     // expf(-0.25f/3)
constexpr float kDecayInhQuarter = 0x1.f35bd2p-1f;    // expf(-0.25f/10)
constexpr float kDecayExcMs = 0x1.6edd3p-1f;          // expf(-1.0f/3)
constexpr float kDecayInhMs = 0x1.cf46dap-1f;         // expf(-1.0f/10)
constexpr float kBranchOffset = 0x1.dc07fcp-13f;      // 5/(1+expf(10))
constexpr float kDecayLtp = 0x1.e2c044p-1f;           // expf(-1.0f/17)
constexpr float kDecayLtd = 0x1.f12914p-1f;           // expf(-1.0f/34)
constexpr float kDecayEligLtp = 0x1.ffedccp-1f;       // expf(-1.0f/7200)
constexpr float kDecayEligLtd = 0x1.ffdb98p-1f;       // expf(-1.0f/3600)
constexpr float kDecayOpen40 = 0x1.f35bd2p-1f;        // expf(-1.0f/40)
constexpr float kDecayOpen750 = 0x1.ff515ap-1f;       // expf(-1.0f/750)
constexpr float kLog192 = 0x1.4dfd9p-1f;              // logf(1.92f)
constexpr float kLog070 = -0x1.6d3c34p-2f;            // logf(0.70f)

}  // namespace nsa
