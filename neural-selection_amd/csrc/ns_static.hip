// ns_step_static: ns_step with the one-brain pool laid out in a file-scope
// device array instead of device memory of the handle's own, for callers
// that may not allocate (code under graph capture, tight serving loops,
// harnesses that forbid it); no counterpart in the reference.
//
// Each device holds its own copy of the array, whose address
// hipGetSymbolAddress gives once per device (with that device current); host
// code uses that address only, never the array's name.  Calls on one device
// share the array: a mutex per device serialises this process's calls (each
// lays its pool out afresh: genes, state, cleared counters), and the array
// keeps the last call's contents.
#include <mutex>

#include "ns_host.h"

namespace nsa {

void set_error(const std::string &msg);

constexpr size_t kStaticPoolBytes = NS_STEP_STATIC_BYTES;
constexpr int kStaticDevices = 64;
static __device__ __align__(256) unsigned char scratch[kStaticPoolBytes];
static unsigned char *g_scratch[kStaticDevices];
static std::mutex g_scratch_init;
static std::mutex g_scratch_use[kStaticDevices];

static unsigned char *scratch_on(int dev) {
    std::lock_guard<std::mutex> lock(g_scratch_init);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        NS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[dev] = static_cast<unsigned char *>(p);
    }
    return g_scratch[dev];
}

static void step_static(GenomeHandle &h, const float *input, int ticks, float gain, float *output) {
    check_step(h.brain(), input, ticks, gain);
    int dev = 0;
    NS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kStaticDevices) throw std::runtime_error("Device index past the static pools");
    unsigned char *base = scratch_on(dev);
    std::lock_guard<std::mutex> lock(g_scratch_use[dev]);
    DevicePool pool;
    pool.arena.carve(base, kStaticPoolBytes);
    float *d_in = nullptr, *d_out = nullptr;
    try {
        build_one_brain(pool, h.brain(), d_in, d_out);
    } catch (const CarveFull &) {
        throw std::runtime_error("Brain of " + std::to_string(h.brain().genes.size()) + " neurons and " +
                                 std::to_string(h.brain().edges.size()) +
                                 " synapses exceeds the static device pool of " +
                                 std::to_string(kStaticPoolBytes) + " bytes");
    }
    step_one_brain(pool, h, d_in, d_out, input, ticks, gain, output);
}

}  // namespace nsa

using namespace nsa;

extern "C" {

int ns_step_static(void *p, const float *input, int ticks, float gain, float *output) {
    try {
        step_static(*static_cast<GenomeHandle *>(p), input, ticks, gain, output);
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

}  // extern "C"
