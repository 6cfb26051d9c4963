// Host build of the device math (neural-selection_amd/csrc/ns_math.h)
// compared with this machine's libm over every float (expf) or every
// non-negative float (logf).  Prints "<mismatches> <first bad input hex>".
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "ns_math.h"

static const uint64_t kExp[32] = NS_EXP2_TABLE;
static const double kLog[32] = NS_LOGF_TABLE;

int main(int argc, char **argv) {
    const bool use_log = argc > 1 && std::strcmp(argv[1], "logf") == 0;
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
