// Accessors over the reference's own CPU path — TEST INFRASTRUCTURE ONLY.
//
// Compiled by oracle/Makefile together with the reference sources where they
// lie (gym_adapter/native.cpp, which includes src/physiology.cuh and
// gym_adapter/core.cuh) into oracle/_ref/libns_ref.so.  Nothing here restates
// the reference: it only exposes the reference's functions and its Brain
// internals (native.cpp:23-37) to the parity tests through a C ABI.
#include "native.cpp"

#include <cstring>

extern "C" {

// Full state of a native.cpp Brain handle.  history9 is count x 9 bytes.
void ref_brain_export(void *p, NeuronState *ns, SynapseState *ss, unsigned char *history9,
                      int64_t *tick) {
    auto &b = *static_cast<Brain *>(p);
    if (ns) std::memcpy(ns, b.neurons.data(), b.neurons.size() * sizeof(NeuronState));
    if (ss && !b.synapses.empty())
        std::memcpy(ss, b.synapses.data(), b.synapses.size() * sizeof(SynapseState));
    if (history9) std::memcpy(history9, b.history.data(), b.history.size() * 9);
    if (tick) *tick = b.tick;
}

void ref_brain_genes(void *p, NeuronGene *ng, SynapseGene *sg) {
    auto &b = *static_cast<Brain *>(p);
    if (ng) std::memcpy(ng, b.genes.data(), b.genes.size() * sizeof(NeuronGene));
    if (sg && !b.connections.empty())
        std::memcpy(sg, b.connections.data(), b.connections.size() * sizeof(SynapseGene));
}

float ref_fn_dendritic_transfer(float c) { return dendritic_transfer(c); }
void ref_fn_profile(int type, NeuronGene *out) { *out = initial_profile((unsigned char)type); }
void ref_fn_initial_neuron(const NeuronGene *g, uint32_t seed, NeuronState *out) {
    *out = initial_neuron_state(*g, seed);
}
void ref_fn_initial_synapse(const SynapseGene *g, uint32_t seed, SynapseState *out) {
    *out = initial_synapse_state(*g, seed);
}
float ref_fn_receptor_current(const NeuronGene *g, NeuronState *s, float level) {
    return receptor_current(*g, *s, level);
}
void ref_fn_neuron_advance(const NeuronGene *g, NeuronState *s, float current, const float *exc,
                           float inh) {
    neuron_advance(*g, *s, current, exc, inh);
}
float ref_fn_synapse_release(const SynapseGene *g, SynapseState *s, int64_t tick) {
    return synapse_release(*g, *s, tick);
}
void ref_fn_plasticity(const SynapseGene *g, SynapseState *s, int arrived, int post_spike, int pair,
                       float depol, int mod) {
    synapse_plasticity_step(*g, *s, arrived, post_spike, pair, depol, true, (unsigned char)mod);
}
void ref_fn_homeostasis(const NeuronGene *g, NeuronState *s, float elapsed, float hz) {
    homeostatic_advance(*g, *s, elapsed, hz);
}
void ref_fn_plasticity_quiet(const SynapseGene *g, SynapseState *s, int ticks) {
    synapse_plasticity_quiet(*g, *s, ticks);
}

}  // extern "C"
