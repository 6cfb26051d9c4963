// Genome handles with state: the reference's single-brain CPU entry points
// (gym_adapter/native.cpp:89-109 ns_reset / ns_step) on the device.
//
// A handle keeps the state the reference's CPU handle keeps (neuron and
// synapse records, the delay ring, the tick) on the host.  ns_reset writes the
// initial states there; ns_step moves them through a one-brain device pool
// for `ticks` ticks of the gym rules (gym_adapter/core.cuh:30-75, the same
// kernels and arithmetic as ns_gpu_step) and back.
#include <cmath>

#include "ns_host.h"

namespace nsa {

void set_error(const std::string &msg);

GenomeHandle *make_handle(HostBrain *parsed) {
    std::unique_ptr<HostBrain> b(parsed);
    auto *h = new GenomeHandle;
    b->slow_cell = int8_t(has_slow_cell(*b));
    h->genome = std::shared_ptr<HostBrain>(std::move(b));
    // native.cpp:46-82 sizes the state vectors value-initialised (zeros)
    h->neurons.assign(h->brain().genes.size(), ns_neuron_state{});
    h->synapses.assign(h->brain().edges.size(), ns_synapse_state{});
    h->hist.assign(h->brain().genes.size(), 0);
    return h;
}

// native.cpp:89-94 with core.cuh:25-29: tick 0, the initial state of every
// neuron and synapse (physiology.cuh:179-202, 393-404) seeded seed +
// original * 7919 / + original * 104729, an empty ring.
static void reset_handle(GenomeHandle &h, uint32_t seed) {
    const HostBrain &b = h.brain();
    h.tick = 0;
    for (size_t i = 0; i < b.genes.size(); ++i) {
        const ns_neuron_gene &g = b.genes[i];
        const uint32_t orig = b.orig_n.empty() ? uint32_t(i) : b.orig_n[i];
        ns_neuron_state s;
        std::memset(&s, 0, sizeof s);
        s.voltage_mv = initial_cell(g.type, g.rest_mv, g.threshold_mv, s.cell);
        s.receptor_rng = initial_rng(seed + orig * 7919u);
        h.neurons[i] = s;
        h.hist[i] = 0;
    }
    for (size_t j = 0; j < b.edges.size(); ++j) {
        const ns_synapse_gene &g = b.edges[j];
        const uint32_t orig = b.orig_s.empty() ? uint32_t(j) : b.orig_s[j];
        ns_synapse_state s;
        std::memset(&s, 0, sizeof s);
        s.weight_ns = g.initial_weight_ns;
        s.da_age_ms = -1;
        s.spine_volume = 1;
        s.structural_age_ms = -1;
        s.pre_age_ms = 1000000;
        s.post_age_ms = 1000000;
        s.last_arrival_tick = -1000000;
        s.ready_mask = initial_ready(site_count(g));
        s.rng_state = initial_rng(seed + orig * 104729u);
        s.slow_resource = 1;
        h.synapses[j] = s;
    }
}

// native.cpp:97-98: the checks and messages of the reference's ns_step.
void check_step(const HostBrain &b, const float *input, int ticks, float gain) {
    if (ticks < 1 || ticks > 10000 || !std::isfinite(gain) || gain < 0)
        throw std::runtime_error("Invalid neural integration interval");
    for (int c = 0; c < b.channels; ++c)
        if (!std::isfinite(input[c]) || input[c] < 0 || input[c] > 1)
            throw std::runtime_error("Invalid encoded observation");
}

// Lay a one-brain pool out in `pool` (its arena allocates, or carves a block
// given to it beforehand), with the call's input and output rows after it.
void build_one_brain(DevicePool &pool, const HostBrain &b, float *&d_in, float *&d_out) {
    std::vector<BrainSpec> specs{BrainSpec{&b, 0, 0, 0, false}};
    pool.build(specs, int(b.motors.size()), default_tile({&b}), nullptr);
    d_in = pool.arena.alloc<float>(b.channels);
    d_out = pool.arena.alloc<float>(int64_t(b.motors.size()));
}

// One ns_step call on a built one-brain pool, synchronous on the null stream:
// the handle's state in, `ticks` ticks (resident launch, gym homeostasis),
// outputs and state out.  A nonfinite voltage completes every tick, as the
// device batch does (gym_adapter/gpu.cu), keeps that state and reports -1.
void step_one_brain(DevicePool &pool, GenomeHandle &h, float *d_in, float *d_out,
                    const float *input, int ticks, float gain, float *output) {
    const HostBrain &b = h.brain();
    pool.upload_state(h.neurons, h.synapses, h.hist, std::vector<int64_t>{h.tick});
    NS_HIP(hipMemcpy(d_in, input, sizeof(float) * b.channels, hipMemcpyHostToDevice));
    const uint8_t run = uint8_t(1 | (pool.fallback[0] ? 2 : 0));
    NS_HIP(hipMemcpy(pool.view.brun, &run, 1, hipMemcpyHostToDevice));
    run_pool(pool, NS_FORM_RESIDENT, ticks, d_in, gain, d_out, int64_t(b.motors.size()), 1, nullptr);
    NS_HIP(hipGetLastError());
    NS_HIP(hipStreamSynchronize(nullptr));
    NS_HIP(hipMemcpy(output, d_out, sizeof(float) * b.motors.size(), hipMemcpyDeviceToHost));
    int32_t status = 0;
    NS_HIP(hipMemcpy(&status, pool.view.bstatus, sizeof status, hipMemcpyDeviceToHost));
    std::vector<int64_t> ticks_out;
    pool.download_state(h.neurons, h.synapses, h.hist, ticks_out);
    h.tick = ticks_out[0];
    if (status) throw std::runtime_error("Nonfinite neural state");
}

}  // namespace nsa

using namespace nsa;

extern "C" {

void ns_reset(void *p, uint32_t seed) {
    reset_handle(*static_cast<GenomeHandle *>(p), seed);
}

int ns_step(void *p, const float *input, int ticks, float gain, float *output) {
    try {
        auto &h = *static_cast<GenomeHandle *>(p);
        check_step(h.brain(), input, ticks, gain);
        if (!h.device) {
            auto dev = std::make_unique<HandleDevice>();
            build_one_brain(dev->pool, h.brain(), dev->d_in, dev->d_out);
            h.device = std::move(dev);
        }
        step_one_brain(h.device->pool, h, h.device->d_in, h.device->d_out, input, ticks, gain, output);
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

int ns_export(void *p, ns_neuron_state *neurons, ns_synapse_state *synapses, uint32_t *spike_bits,
              int64_t *tick) {
    try {
        const auto &h = *static_cast<GenomeHandle *>(p);
        if (neurons) std::memcpy(neurons, h.neurons.data(), h.neurons.size() * sizeof(ns_neuron_state));
        if (synapses && !h.synapses.empty())
            std::memcpy(synapses, h.synapses.data(), h.synapses.size() * sizeof(ns_synapse_state));
        if (spike_bits) std::memcpy(spike_bits, h.hist.data(), h.hist.size() * sizeof(uint32_t));
        if (tick) *tick = h.tick;
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

}  // extern "C"
