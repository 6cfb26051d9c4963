// Host build of the device math (neural-selection_amd/csrc/ns_math.h)
// compared with this machine's libm over every float (expf) or every
// non-negative float (logf).  Prints "<mismatches> <first bad input hex>".
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

#include "ns_math.h"

static const uint64_t kExp[32] = NS_EXP2_TABLE;
static const double kLog[32] = NS_LOGF_TABLE;

// digest mode: "<fn> digest <lo> <hi> <chunk>" prints one digest per chunk:
// sum of bits(f(u)) * (u | 1) mod 2^64, NaN results as 0x7fc00000.
static int digest_mode(bool use_log, uint64_t lo, uint64_t hi, uint64_t chunk) {
    const uint64_t n = (hi - lo + chunk - 1) / chunk;
    std::vector<uint64_t> d(n, 0);
    const int threads = 16;
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            for (uint64_t c = t; c < n; c += threads) {
                uint64_t acc = 0;
                const uint64_t end = std::min(hi, lo + (c + 1) * chunk);
                for (uint64_t u = lo + c * chunk; u < end; ++u) {
                    float x;
                    uint32_t v = uint32_t(u);
                    std::memcpy(&x, &v, 4);
                    float y = use_log ? logf(x) : expf(x);
                    uint32_t b;
                    std::memcpy(&b, &y, 4);
                    if (std::isnan(y)) b = 0x7fc00000u;
                    acc += uint64_t(b) * (u | 1u);
                }
                d[c] = acc;
            }
        });
    for (auto &th : pool) th.join();
    for (uint64_t c = 0; c < n; ++c) std::printf("%llu\n", (unsigned long long)d[c]);
    return 0;
}

int main(int argc, char **argv) {
    const bool use_log = argc > 1 && std::strcmp(argv[1], "logf") == 0;
    if (argc > 5 && std::strcmp(argv[2], "digest") == 0)
        return digest_mode(use_log, std::strtoull(argv[3], nullptr, 0),
                           std::strtoull(argv[4], nullptr, 0), std::strtoull(argv[5], nullptr, 0));
    const uint64_t lo = 0, hi = use_log ? 0x80000000ull : 0x100000000ull;
    const uint64_t stride = argc > 2 ? std::strtoull(argv[2], nullptr, 0) : 1;
    const int threads = 8;
    std::vector<long long> bad(threads, 0);
    std::vector<uint32_t> first(threads, 0);
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            for (uint64_t u = lo + t * stride; u < hi; u += threads * stride) {
                float x;
                uint32_t v = uint32_t(u);
                std::memcpy(&x, &v, 4);
                float a = use_log ? logf(x) : expf(x);
                float b = use_log ? nsa::ns_logf(x, kLog) : nsa::ns_expf(x, kExp);
                if (std::isnan(a) && std::isnan(b)) continue;
                uint32_t ua, ub;
                std::memcpy(&ua, &a, 4);
                std::memcpy(&ub, &b, 4);
                if (ua != ub) {
                    if (!bad[t]) first[t] = v;
                    ++bad[t];
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
