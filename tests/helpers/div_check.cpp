// Checks the device's division shortcuts against IEEE float division over
// every float numerator (all 2^32 bit patterns) for the given divisors:
//   const divisor:    x / c  ==  div_const(x, c, 1.0 / (double)c)
//   variable divisor: 5 / d  ==  div5(d) for d = 1 + e, e >= 0
// The device's reciprocal estimate is modelled as RN(1/d) perturbed by up
// to +-4 float ulps (the Newton steps must absorb any such start).
// Prints "<mismatches> <first bad input hex>".
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

static inline float div_const(float x, float c, double rc) {
    const double q = (double)x * rc;
    float r = (float)q;
    if (std::fabs(q) < 0x1p-126 && x != 0.0f) r = x / c;
    return r;
}

static inline float div5(float d, int perturb) {
    float r0f = 1.0f / d;
    uint32_t u;
    std::memcpy(&u, &r0f, 4);
    u += perturb;
    std::memcpy(&r0f, &u, 4);
    const double dd = d;
    double r = r0f;
    double e = std::fma(-dd, r, 1.0);
    r = std::fma(r, e, r);
    e = std::fma(-dd, r, 1.0);
    r = std::fma(r, e, r);
    const double q = 5.0 * r;
    const double a = std::fabs(q);
    if (a != 0.0 && a < 0x1p-126) return 5.0f / d;
    return (float)q;
}

int main(int argc, char **argv) {
    std::vector<float> divisors;
    for (int i = 1; i < argc; ++i) divisors.push_back(std::strtof(argv[i], nullptr));
    const int threads = 16;
    std::vector<long long> bad(threads, 0);
    std::vector<uint32_t> first(threads, 0);
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            for (uint64_t u = t; u < (1ull << 32); u += threads) {
                float x;
                uint32_t v = uint32_t(u);
                std::memcpy(&x, &v, 4);
                for (float c : divisors) {
                    const float a = x / c, b = div_const(x, c, 1.0 / (double)c);
                    if (std::memcmp(&a, &b, 4) && !(std::isnan(a) && std::isnan(b))) {
                        if (!bad[t]) first[t] = v;
                        ++bad[t];
                    }
                }
                // variable divisor: d = 1 + e over every float e in [0, 2.3e4]
                if (!(x >= 0.0f) || x > 23000.0f) continue;
                const float d = 1.0f + x;
                const float a = 5.0f / d;
                for (int p = -4; p <= 4; p += 4) {
                    const float b = div5(d, p);
                    if (std::memcmp(&a, &b, 4)) {
                        if (!bad[t]) first[t] = v;
                        ++bad[t];
                    }
                }
            }
        });
    for (auto &th : pool) th.join();
    long long total = 0;
    uint32_t f = 0;
    for (int t = 0; t < threads; ++t) {
        total += bad[t];
        if (bad[t] && !f) f = first[t];
    }
    std::printf("%lld 0x%08x\n", total, f);
    return 0;
}
