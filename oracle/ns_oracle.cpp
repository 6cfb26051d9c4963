// CPU oracle for the per-tick brain update — TEST INFRASTRUCTURE ONLY.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load this library, and only as the checker or the timed CPU baseline.  The
// product (libns_amd.so) never links or calls it.
//
// It restates, in plain host C++ with the host libm (expf/logf), the
// arithmetic of the reference's src/physiology.cuh and the tick order of
// gym_adapter/core.cuh (batched episodes) and src/main.cu (world neural
// step).  Each function cites the reference lines it follows.  Parity of this
// restatement is pinned in tests/test_oracle_*.py against the reference's own
// sources host-compiled into oracle/_ref (when /root/reference is present),
// against committed golden vectors under tests/golden/, and against closed
// forms.
//
// Build: oracle/Makefile (g++ -O2 -std=c++17, no -march, no fast-math, so
// float arithmetic is plain SSE IEEE single precision without contraction).
#include "../include/ns_amd.h"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>
#include <limits>

namespace {

using NG = ns_neuron_gene;
using NS = ns_neuron_state;
using SG = ns_synapse_gene;
using SS = ns_synapse_state;

thread_local std::string g_error;

constexpr int kBranches = 4;                      // physiology.cuh:17
constexpr unsigned kModNE = 1, kMod5HT = 2, kModDA = 4;  // physiology.cuh:23-25

// physiology.cuh:78-80
inline float clamp_range(float x, float lo, float hi) { return fminf(hi, fmaxf(lo, x)); }

// ---------------------------------------------------------------------------
// Neuron profiles: physiology.cuh:82-176, prinz_cell.cuh:24-28, lp_cell.cuh:24-27
// ---------------------------------------------------------------------------
NG blank_gene() {
    NG g;
    std::memset(&g, 0, sizeof g);
    g.sensory_channel = -1;          // physiology.cuh:41
    g.receptor_decay_ms = 3;         // physiology.cuh:43
    g.target_affinity = 1;           // physiology.cuh:44
    return g;
}

NG naud_profile(int type) {          // physiology.cuh:82-89 (recorded_profile)
    NG g = blank_gene();
    if (type == 1) {
        const float p[11] = {59, 2.9f, -62, -42, 3.0f, 1.8f, 16, 61, -54, 0, 0};
        std::memcpy(&g.capacitance_pf, p, sizeof p);
        g.type = 1; g.active = 1; g.modulator = 0; g.inhibitory_transmitter = 1;
        return g;
    }
    const float p[11] = {104, 4.3f, -65, -52, 0.8f, -0.8f, 88, 65, -53, 0, 0};
    std::memcpy(&g.capacitance_pf, p, sizeof p);
    g.type = 0; g.active = 1; g.modulator = 0;
    return g;
}

NG profile_for(int type) {           // physiology.cuh:169-176 (initial_profile)
    NG g;
    switch (type) {
    case 0:                          // physiology.cuh:94-100
        g = naud_profile(0);
        g.reset_mv = -68.0f; g.afterdepol_spike_pa = 120.0f; g.afterdepol_tau_ms = 45.0f;
        return g;
    case 1:
        return naud_profile(1);
    case 2:                          // physiology.cuh:102-107
        g = naud_profile(1); g.type = 2;
        g.lp_scale[0] = g.lp_scale[1] = g.lp_scale[2] = g.lp_scale[3] = 1;
        return g;
    case 3: {                        // physiology.cuh:109-114, prinz_cell.cuh:24-28
        g = naud_profile(1); g.type = 3;
        const float pm[8] = {400 * 628.0f, 2.5f * 628, 6 * 628.0f, 50 * 628.0f,
                             10 * 628.0f, 100 * 628.0f, .01f * 628, 0};
        std::memcpy(g.prinz_g, pm, sizeof pm);
        return g;
    }
    case 4:                          // physiology.cuh:139-150
        g = naud_profile(0); g.type = 4;
        g.capacitance_pf = 8.7f; g.leak_ns = 8.7f / 12.0f; g.rest_mv = -60.0f;
        g.threshold_mv = -40.0f; g.reset_mv = -60.0f; g.adapt_ns = 0; g.spike_adapt_pa = 0;
        return g;
    case 5:                          // physiology.cuh:119-134
        g = naud_profile(0); g.type = 5;
        g.capacitance_pf = 10.0f; g.leak_ns = 0.5f; g.rest_mv = -60.0f;
        g.threshold_mv = -40.0f; g.reset_mv = -60.0f; g.adapt_ns = 0; g.spike_adapt_pa = 0;
        g.receptor_baseline_hz = 7.0f; g.receptor_gain_hz = 43.0f;
        g.receptor_event_ns = 2.0f; g.receptor_decay_ms = 3.0f;
        return g;
    default:                         // physiology.cuh:155-163
        g = naud_profile(0); g.type = 6;
        g.capacitance_pf = 1200.0f; g.leak_ns = 1000.0f / 65.6f;
        g.rest_mv = -45.1f; g.threshold_mv = -51.8f;
        return g;
    }
}

inline bool excitatory_cell(unsigned type) { return type == 0 || type == 4 || type == 5; }  // :165-167

inline uint32_t seed_mix(uint32_t seed) {        // physiology.cuh:183-186, 395-398
    uint32_t x = seed ? seed : 1;
    x ^= x >> 16; x *= 0x7feb352du;
    x ^= x >> 15; x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// physiology.cuh:179-202, lp_cell.cuh:24-27, prinz_cell.cuh:34-40
NS fresh_neuron(const NG &g, uint32_t seed) {
    NS s;
    std::memset(&s, 0, sizeof s);
    s.voltage_mv = g.rest_mv;
    uint32_t x = seed_mix(seed);
    s.receptor_rng = x ? x : 1;
    if (g.type == 2) {
        const float lp[16] = {-56, 0, .2f, 0, 1, 0, 0, 1, 0, 0, 1, 1, 0, .04f, .01f, -56};
        std::memcpy(s.cell, lp, sizeof lp);
        s.voltage_mv = s.cell[15] - 9.342518882f;
    }
    if (g.type == 3) {
        std::memset(s.cell, 0, sizeof s.cell);
        s.cell[0] = -50.0f; s.cell[1] = .05f;
        for (int i = 0; i < 4; ++i) s.cell[9 + i] = 1.0f;
        s.voltage_mv = s.cell[0];
    }
    if (g.type == 6) {
        for (int i = 0; i < 8; ++i) s.cell[i] = 0;
        s.cell[0] = g.rest_mv;
        s.cell[1] = g.threshold_mv;
    }
    return s;
}

inline int sites_of(const SG &g) {               // physiology.cuh:376-378
    return g.release_sites ? (g.release_sites > 64 ? 64 : g.release_sites) : 14;
}

SS fresh_synapse(const SG &g, uint32_t seed) {    // physiology.cuh:393-404
    uint32_t x = seed_mix(seed);
    if (!x) x = 1;
    SS s;
    std::memset(&s, 0, sizeof s);
    s.weight_ns = g.initial_weight_ns;
    s.da_age_ms = -1; s.spine_volume = 1; s.structural_age_ms = -1;
    s.pre_age_ms = 1000000; s.post_age_ms = 1000000;
    s.last_arrival_tick = -1000000;
    int n = sites_of(g);
    s.ready_mask = n == 64 ? ~uint64_t(0) : (uint64_t(1) << n) - 1;
    s.rng_state = x;
    s.arrived = 0;
    s.slow_resource = 1;
    return s;
}

// ---------------------------------------------------------------------------
// Membrane dynamics
// ---------------------------------------------------------------------------

// physiology.cuh:206-219
float receptor_drive(const NG &g, NS &s, float level) {
    if (g.sensory_channel < 0 || g.receptor_event_ns <= 0) return 0;
    float rate = clamp_range(g.receptor_baseline_hz + g.receptor_gain_hz * level, 0, 1000);
    uint32_t r = s.receptor_rng;
    r ^= r << 13; r ^= r >> 17; r ^= r << 5;
    s.receptor_rng = r;
    float u = (r >> 8) * (1.0f / 16777216.0f);
    s.receptor_ns *= expf(-1.0f / g.receptor_decay_ms);
    if (u < rate * (1.0f / 1000.0f)) s.receptor_ns += g.receptor_event_ns;
    return s.receptor_ns * (0.0f - s.voltage_mv);
}

// physiology.cuh:226-230
float branch_io(float c) {
    float regen = 5.0f / (1.0f + expf(-(c - 4.0f) / 0.4f));
    float floor_term = 5.0f / (1.0f + expf(10.0f));
    return fminf(11.0f, c + regen - floor_term);
}

// physiology.cuh:243-253
void homeostasis(const NG &g, NS &s, float elapsed_ms, float observed_hz) {
    if (g.type != 0) return;
    float ratio = clamp_range(observed_hz / 0.4f, 0.0f, 2.65f);
    float change = ratio < 1.0f ? logf(1.92f) * (1.0f - ratio)
                                : logf(0.70f) * ((ratio - 1.0f) / 1.65f);
    s.log_ampa_scale = clamp_range(s.log_ampa_scale + change * (elapsed_ms / 172800000.0f),
                                   -3.0f, 3.0f);
}

// physiology.cuh:258-287 (LFS cell in cell[0..7]: voltage, threshold, activated[3], open[3])
void lfs_tick(const NG &g, NS &s, float current_pa, float exc_ns, float inh_ns) {
    const float w[3] = {250.0f, 10.0f, 4.0f};
    const float open_ms[3] = {10.0f, 40.0f, 750.0f};
    const float close_ms[3] = {50.0f, 750.0f, 2000.0f};
    float *c = s.cell;
    c[1] = g.threshold_mv + (c[1] - g.threshold_mv) * expf(-1.0f / 10.0f);
    float under = 0;
    for (int j = 0; j < 3; ++j) {
        c[5 + j] += c[2 + j] / open_ms[j] - c[5 + j] / close_ms[j];
        c[2 + j] *= expf(-1.0f / open_ms[j]);
        under += w[j] * c[5 + j];
    }
    float v = c[0];
    float i = g.leak_ns * (g.rest_mv - v) + exc_ns * (10.0f - v) + inh_ns * (-60.0f - v) +
              under * (-80.0f - v) + current_pa;
    c[0] = v + i / g.capacitance_pf;
    s.spike = c[0] >= c[1];
    if (s.spike) {
        c[1] = 200.0f;
        for (int j = 0; j < 3; ++j) c[2 + j] += 1.0f;
        ++s.total_spikes;
    }
    s.voltage_mv = c[0];
}

// lp_cell.cuh:29-31, 33-37
inline float lp_sig(float v, float mid, float slope) { return 1.0f / (1.0f + expf((v - mid) / slope)); }
inline float lp_ratio(float num, float den, float lim) { return fabsf(den) < 1e-5f ? lim : num / den; }

// lp_cell.cuh:41-96
void lp_rhs(const float *x, float i_na, const float *sc, float *d) {
    const float va = x[0], vs = x[15];
    const float na = x[1] * x[1] * x[1] * x[2] * 7.15f * (va - 50.0f);
    const float kd = x[8] * x[8] * x[8] * x[8] * 1.43f * (va + 72.0f);
    const float mk = .2613041227f * x[14] * (va + 80.0f);
    const float aleak = sc[3] * .001004288227f * (va + 50.0f);
    const float coup = .004017089073f * (vs - va);
    d[0] = (-na - kd - mk - aleak + coup) / .1148312178f;
    const float ev = expf(vs / 11.4866f);
    const float ghk = fabsf(vs) < .001f ? 11.4866f * (x[13] - 15123.5f)
                                       : vs * (x[13] * ev - 15123.5f) / (ev - 1.0f);
    const float cal = sc[0] * (x[3] * x[4] * .1247796343f + x[5] * .07142141077f) * .00155982f * ghk * .9f;
    const float cak = sc[1] * x[6] * x[7] * 1.663612026f * (vs + 72.0f) * .9f;
    const float amix = lp_sig(vs, 55.0f, -15.0f);
    const float ak = x[9] * x[9] * x[9] * .8008791415f * (vs + 72.0f) *
                     (amix * x[10] + (1.0f - amix) * x[11]) * .9f;
    const float ih = sc[2] * x[12] * .0126874385f * (vs + 20.0f) * .9f;
    const float sleak = sc[3] * .0006901584324f * (vs + 50.0f);
    d[15] = (-cal - cak - ak - ih - sleak - coup + (i_na - 2.0f) * .01222505828f) / .0543909161f;
    float a = lp_ratio(-16.64f - .32f * va, expf(-13.0f - va / 4.0f) - 1.0f, 1.28f);
    float b = lp_ratio(.28f * va + 7.0f, expf(va / 5.0f + 5.0f) - 1.0f, 1.4f);
    d[1] = a * (1.0f - x[1]) - b * x[1];
    a = .128f * expf(-2.66666666667f - va / 18.0f);
    b = 4.0f / (expf(-5.0f - va / 5.0f) + 1.0f);
    d[2] = a * (1.0f - x[2]) - b * x[2];
    a = lp_ratio(-1.6f - .032f * va, expf(-10.0f - va / 5.0f) - 1.0f, .16f);
    b = .5f * expf(-1.375f - va / 40.0f);
    d[8] = a * (1.0f - x[8]) - b * x[8];
    d[3] = (lp_sig(vs, 15.0f, -9.8f) - x[3]) * .045f;
    d[4] = (lp_sig(vs, -40.0f, 3.2f) - x[4]) * .02f * lp_sig(vs, -15.0f, -10.0f);
    d[5] = (lp_sig(vs, 29.1336f, -4.43083f) - x[5]) * .0608644f;
    const float ca = x[13];
    const float tgt = lp_sig(vs, -.6f * ca, -23.0f) * lp_sig(vs, -16.0f - .6f * ca, -5.0f) * ca / (2.5f + ca);
    d[6] = (tgt - x[6]) * .6f;
    d[7] = (.7f / (.6f + ca) - x[7]) * .035f;
    d[9] = (lp_sig(vs, -12.0f, -26.0f) - x[9]) * .14f;
    const float ainf = lp_sig(vs, -62.0f, 6.0f);
    d[10] = (ainf - x[10]) * .05f;
    d[11] = (ainf - x[11]) * .0036f * lp_sig(vs, -40.0f, -12.0f);
    d[12] = (lp_sig(vs, -70.0f, 8.0f) - x[12]) * .00033f / lp_sig(vs, -110.0f, -21.6f);
    d[13] = -.05049639866f * cal - .01729930764f * (ca - .02f);
    d[14] = (lp_sig(va, -26.9916223f, -5.957441281f) - x[14]) * .0001386796618f *
            lp_sig(va, -60.57875745f, -13.25674673f);
}

// lp_cell.cuh:98-113
void lp_tick(float *x, float i_na, const float *sc) {
    const float dt = .05f;
    for (int step = 0; step < 20; ++step) {
        float k1[16], k2[16], k3[16], k4[16], t[16];
        lp_rhs(x, i_na, sc, k1);
        for (int j = 0; j < 16; ++j) t[j] = x[j] + .5f * dt * k1[j];
        lp_rhs(t, i_na, sc, k2);
        for (int j = 0; j < 16; ++j) t[j] = x[j] + .5f * dt * k2[j];
        lp_rhs(t, i_na, sc, k3);
        for (int j = 0; j < 16; ++j) t[j] = x[j] + dt * k3[j];
        lp_rhs(t, i_na, sc, k4);
        for (int j = 0; j < 16; ++j) x[j] += dt * (k1[j] + 2 * k2[j] + 2 * k3[j] + k4[j]) / 6.0f;
    }
}

// prinz_cell.cuh:41-46
inline float pz_gate(float v, float a, float b) { return 1.0f / (1.0f + expf((v + a) / b)); }
inline float pz_gate2(float v, float a, float b, float c, float d) {
    return 1.0f / (expf((v + a) / b) + expf((v + c) / d));
}

// prinz_cell.cuh:44-108.  Cell layout in cell[]: v, ca, m[7], h[4].
void pz_substep(float *c, const float *gg, float ext_pa, float clamp_ns, float clamp_rev) {
    const float dt = .025f;
    float *m = c + 2, *h = c + 9;
    const float v = c[0], ca = fmaxf(c[1], 0.000001f);
    const float eca = (8.31451e3f * 283.0f / (2.0f * 96485.3415f)) * logf(3000.0f / ca);
    const float cond[8] = {gg[0] * m[0] * m[0] * m[0] * h[0], gg[1] * m[1] * m[1] * m[1] * h[1],
                           gg[2] * m[2] * m[2] * m[2] * h[2], gg[3] * m[3] * m[3] * m[3] * h[3],
                           gg[4] * m[4] * m[4] * m[4] * m[4], gg[5] * m[5] * m[5] * m[5] * m[5],
                           gg[6] * m[6], gg[7]};
    const float rev[8] = {50, eca, eca, -80, -80, -80, -20, -50};
    float gsum = clamp_ns;
    float wsum = ext_pa + clamp_ns * clamp_rev;
    for (int i = 0; i < 8; ++i) { gsum += cond[i]; wsum += cond[i] * rev[i]; }
    if (fabsf(gsum) < .000001f)
        c[0] = v + dt * wsum / 628.3f;
    else {
        const float target = wsum / gsum;
        c[0] = target + (v - target) * expf(-dt * gsum / 628.3f);
    }
    const float ica = (cond[1] + cond[2]) * (v - eca) / 1000000.0f;
    const float ca_t = .05f - 14961.0f * ica;
    c[1] = fmaxf(.000001f, ca_t + (ca - ca_t) * expf(-dt / 200.0f));
    const float mi[7] = {pz_gate(v, 25.5f, -5.29f), pz_gate(v, 27.1f, -7.2f), pz_gate(v, 33, -8.1f),
                         pz_gate(v, 27.2f, -8.7f), (ca / (ca + 3)) * pz_gate(v, 28.3f, -12.6f),
                         pz_gate(v, 12.3f, -11.8f), pz_gate(v, 75, 5.5f)};
    const float mt[7] = {2.64f - 2.52f * pz_gate(v, 120, -25),
                         43.4f - 42.6f * pz_gate(v, 68.1f, -20.5f),
                         2.8f + 14 * pz_gate2(v, 27, 10, 70, -13),
                         23.2f - 20.8f * pz_gate(v, 32.9f, -15.2f),
                         180.6f - 150.2f * pz_gate(v, 46, -22.7f),
                         14.4f - 12.8f * pz_gate(v, 28.3f, -19.2f),
                         2.0f / (expf(-14.59f - .086f * v) + expf(-1.87f + .0701f * v))};
    for (int i = 0; i < 7; ++i) m[i] = mi[i] + (m[i] - mi[i]) * expf(-dt / mt[i]);
    const float hi[4] = {pz_gate(v, 48.9f, 5.18f), pz_gate(v, 32.1f, 5.5f), pz_gate(v, 60, 6.2f),
                         pz_gate(v, 56.9f, 4.9f)};
    const float ht[4] = {(1.34f / (1 + expf((v + 62.9f) / -10))) * (1.5f + 1 / (1 + expf((v + 34.9f) / 3.6f))),
                         210 - 179.6f * pz_gate(v, 55, -16.9f),
                         120 + 300 * pz_gate2(v, 55, 9, 65, -16),
                         77.2f - 58.4f * pz_gate(v, 38.9f, -26.5f)};
    for (int i = 0; i < 4; ++i) h[i] = hi[i] + (h[i] - hi[i]) * expf(-dt / ht[i]);
}

// prinz_cell.cuh:110-121
bool pz_tick(float *c, const float *gg, float ext_pa, float clamp_ns, float clamp_rev) {
    bool fired = false;
    for (int i = 0; i < 40; ++i) {
        const float before = c[0];
        pz_substep(c, gg, ext_pa, clamp_ns, clamp_rev);
        fired |= before < 0 && c[0] >= 0;
    }
    return fired;
}

// physiology.cuh:290-374 (clamp terms fixed at their defaults 0 nS, -68 mV)
void membrane_tick(const NG &g, NS &s, float current_pa, const float *exc, float inh) {
    const float clamp_ns = 0.0f, clamp_rev = -68.0f;
    for (int b = 0; b < kBranches; ++b) s.dendrite_exc_ns[b] += exc[b];
    float sub = fmaxf(0.0f, fminf(s.voltage_mv, -50.0f) - g.rest_mv);
    s.recent_depol_mv += 0.1f * (sub - s.recent_depol_mv);
    s.inh_ns += inh;
    s.spike = 0;
    const float dec1e = expf(-1.0f / 3.0f), dec1i = expf(-1.0f / 10.0f);
    if (g.type == 6) {
        float e = 0;
        for (int b = 0; b < kBranches; ++b) e += s.dendrite_exc_ns[b];
        lfs_tick(g, s, current_pa, e, s.inh_ns);
        for (int b = 0; b < kBranches; ++b) s.dendrite_exc_ns[b] *= dec1e;
        s.inh_ns *= dec1i;
        return;
    }
    if (g.type == 2) {
        float before = s.cell[0];
        float e = 0;
        for (int b = 0; b < kBranches; ++b) e += s.dendrite_exc_ns[b];
        float ext = current_pa + e * (0.0f - s.voltage_mv) + s.inh_ns * (-75.0f - s.voltage_mv) -
                    clamp_ns * (s.voltage_mv - clamp_rev);
        lp_tick(s.cell, ext / 1000.0f, g.lp_scale);
        s.voltage_mv = s.cell[15] - 9.342518882f;
        s.spike = before < 0 && s.cell[0] >= 0;
        s.total_spikes += s.spike;
        for (int b = 0; b < kBranches; ++b) s.dendrite_exc_ns[b] *= dec1e;
        s.inh_ns *= dec1i;
        return;
    }
    if (g.type == 3) {
        float e = 0;
        for (int b = 0; b < kBranches; ++b) e += s.dendrite_exc_ns[b];
        const float ext = current_pa + e * (0.0f - s.voltage_mv) + s.inh_ns * (-75.0f - s.voltage_mv);
        s.spike = pz_tick(s.cell, g.prinz_g, ext, g.prinz_ambient_leak_ns + clamp_ns, clamp_rev);
        s.total_spikes += s.spike;
        s.voltage_mv = s.cell[0];
        for (int b = 0; b < kBranches; ++b) s.dendrite_exc_ns[b] *= dec1e;
        s.inh_ns *= dec1i;
        return;
    }
    const float h = 1.0f / 4;
    const float dq_e = expf(-h / 3.0f), dq_i = expf(-h / 10.0f);
    for (int step = 0; step < 4; ++step) {
        float v = s.voltage_mv;
        s.adaptation_pa += h * (g.adapt_ns * (v - g.rest_mv) - s.adaptation_pa) / g.adapt_tau_ms;
        float x = clamp_range((v - g.threshold_mv) / g.slope_mv, -20.0f, 10.0f);
        float soma = 0;
        for (int b = 0; b < kBranches; ++b)
            soma += g.type == 0 ? branch_io(s.dendrite_exc_ns[b]) : s.dendrite_exc_ns[b];
        float dv = g.leak_ns * (g.rest_mv - v) + g.leak_ns * g.slope_mv * expf(x) +
                   soma * (0.0f - v) + s.inh_ns * (-75.0f - v) + current_pa + s.afterdepol_pa -
                   s.adaptation_pa - clamp_ns * (v - clamp_rev);
        s.voltage_mv = v + h * dv / g.capacitance_pf;
        if (s.voltage_mv >= 0.0f) {
            s.spike = 1;
            ++s.total_spikes;
            s.voltage_mv = g.reset_mv;
            s.adaptation_pa += g.spike_adapt_pa;
            s.afterdepol_pa += g.afterdepol_spike_pa;
        }
        if (g.afterdepol_tau_ms > 0) s.afterdepol_pa *= expf(-h / g.afterdepol_tau_ms);
        for (int b = 0; b < kBranches; ++b) s.dendrite_exc_ns[b] *= dq_e;
        s.inh_ns *= dq_i;
    }
}

// ---------------------------------------------------------------------------
// Synapses
// ---------------------------------------------------------------------------
inline float uniform24(SS &s) {                   // physiology.cuh:380-391
    uint32_t x = s.rng_state;
    x ^= x << 13; x ^= x >> 17; x ^= x << 5;
    s.rng_state = x;
    return (x >> 8) * (1.0f / 16777216.0f);
}

// physiology.cuh:425-454
float release(const SG &g, SS &s, int64_t tick) {
    float p_recover = 0;
    if (s.last_arrival_tick > -1000000) {
        float dt = (tick - s.last_arrival_tick) * 1.0f;
        p_recover = 1.0f - expf(-dt / g.recovery_ms);
        if (g.facilitation_ms > 0) s.running_utilization *= expf(-dt / g.facilitation_ms);
        if (g.slow_recovery_ms > 0)
            s.slow_resource = 1.0f - (1.0f - s.slow_resource) * expf(-dt / g.slow_recovery_ms);
    }
    s.running_utilization = g.facilitation_ms > 0
        ? s.running_utilization + g.utilization * (1.0f - s.running_utilization)
        : g.utilization;
    int n = sites_of(g), out = 0;
    for (int k = 0; k < n; ++k) {
        uint64_t bit = uint64_t(1) << k;
        if (!(s.ready_mask & bit) && uniform24(s) < p_recover) s.ready_mask |= bit;
        if ((s.ready_mask & bit) && uniform24(s) < s.running_utilization) {
            s.ready_mask &= ~bit;
            ++out;
        }
    }
    s.last_arrival_tick = tick;
    float y = s.weight_ns * out / n * s.slow_resource;
    if (g.slow_recovery_ms > 0) s.slow_resource *= 1.0f - g.slow_depletion;
    return y;
}

float da_envelope(float age) {                    // physiology.cuh:408-412
    if (age < 0 || age > 20000.0f) return 0;
    float x = age / 2000.0f;
    return x * x * x * x * expf(4.0f * (1.0f - x));
}

float spine_window(float age) {                   // physiology.cuh:416-420
    if (age < 0 || age > 5000.0f) return 0;
    float x = age / 1000.0f;
    return x * x * x * expf(3.0f * (1.0f - x));
}

// physiology.cuh:459-555 with nmda/l-type/cp-ampa/pka all available
void plasticity(const SG &g, SS &s, bool arrived, bool post_spike, bool pair, float depol,
                unsigned mod) {
    if (!g.plasticity || !pair) return;
    const float w_max = 2.0f * g.initial_weight_ns;
    if (g.plasticity == 5) {
        if (s.structural_age_ms >= 0) {
            s.structural_age_ms += 1.0f;
            if (s.structural_age_ms > 5000.0f) s.structural_age_ms = -1;
        }
        s.pre_age_ms += 1.0f;
        if (arrived) s.pre_age_ms = 0;
        if (post_spike && s.pre_age_ms <= 15.0f && s.structural_age_ms < 0) s.structural_age_ms = 0;
        if ((mod & kModDA) && (g.mod_receptors & kModDA))
            s.spine_volume = clamp_range(s.spine_volume + 0.025f * spine_window(s.structural_age_ms),
                                         1.0f, 2.0f);
        return;
    }
    if (g.plasticity == 4) {
        if (s.da_age_ms >= 0) {
            s.da_age_ms += 1.0f;
            if (s.da_age_ms > 20000.0f) s.da_age_ms = -1;
        }
        s.pre_age_ms += 1.0f;
        if (arrived) s.pre_age_ms = 0;
        if (post_spike && s.pre_age_ms <= 15.0f && (s.da_age_ms < 0 || s.da_age_ms > 30.0f)) {
            s.da_age_ms = 0;
            s.weight_ns -= 0.005383f * g.initial_weight_ns;
        }
        if ((mod & kModDA) && (g.mod_receptors & kModDA)) {
            s.weight_ns += 0.00453f * da_envelope(s.da_age_ms) * g.initial_weight_ns;
            s.da_age_ms = -1;
        }
        s.weight_ns = clamp_range(s.weight_ns, 0.0f, w_max);
        return;
    }
    if (g.plasticity == 3) {
        s.eligible_ltp *= expf(-1.0f / 7200.0f);
        s.eligible_ltd *= expf(-1.0f / 3600.0f);
        s.pre_age_ms += 1.0f;
        s.post_age_ms += 1.0f;
        if (arrived) {
            if (s.post_age_ms <= 15.0f) s.eligible_ltd += 1.0f;
            s.pre_age_ms = 0;
        }
        if (post_spike) {
            if (s.pre_age_ms <= 15.0f) s.eligible_ltp += 1.0f;
            s.post_age_ms = 0;
        }
        if ((mod & kModNE) && (g.mod_receptors & kModNE)) {
            s.weight_ns += 0.0048f * s.eligible_ltp * g.initial_weight_ns;
            s.eligible_ltp = 0;
        }
        if ((mod & kMod5HT) && (g.mod_receptors & kMod5HT)) {
            s.weight_ns -= 0.0075f * s.eligible_ltd * g.initial_weight_ns;
            s.eligible_ltd = 0;
        }
        s.weight_ns = clamp_range(s.weight_ns, 0.0f, w_max);
        return;
    }
    s.pre_trace *= expf(-1.0f / 17.0f);
    s.post_trace *= expf(-1.0f / 34.0f);
    s.pre_age_ms += 1.0f;
    s.post_age_ms += 1.0f;
    if (arrived) {
        if (s.post_age_ms <= (g.plasticity == 2 ? 50.0f : 20.0f))
            s.weight_ns *= 1.0f - 0.004f * s.post_trace;
        s.pre_trace += 1.0f;
        s.pre_age_ms = 0;
    }
    if (post_spike) {
        if (s.pre_age_ms <= 20.0f && (g.plasticity == 1 || depol >= 2.3f))
            s.weight_ns += 0.016f * s.pre_trace * (2.0f * g.initial_weight_ns - s.weight_ns);
        s.post_trace += 1.0f;
        s.post_age_ms = 0;
    }
    s.weight_ns = clamp_range(s.weight_ns, 0.0f, w_max);
}

// ---------------------------------------------------------------------------
// Brain harness.  Gym order: gym_adapter/core.cuh:25-75.  World rules:
// src/main.cu:659-801 (delivery, mod-event stamps), 710-746 (neuron_step),
// 1287-1300 (homeostasis_step).
// ---------------------------------------------------------------------------
constexpr int kWorldSenses = 19, kWorldMotors = 5;   // main.cu S, M

struct Brain {
    std::vector<NG> genes;
    std::vector<SG> edges;
    std::vector<NS> neurons;
    std::vector<SS> synapses;
    std::vector<int> motors, sensory;
    std::vector<uint8_t> receptor;        // receptor path for this neuron
    std::vector<int> ports;               // world: neuron_port (motor j = port kWorldSenses + j)
    std::vector<uint32_t> orig_n, orig_s;
    std::vector<uint8_t> history;         // count x ring
    std::vector<float> exc, inh;
    std::vector<uint8_t> mod;
    std::vector<uint64_t> stamps;         // count x 6, main.cu:118-120
    int ring = 9, channels = 1;
    int world = 0;                        // main.cu rules
    int frozen = 0;                       // main.cu:142-146
    int64_t tick = 0;
};

bool valid_edge(const Brain &b, const SG &g) {
    return g.active && b.genes[g.pre].active && b.genes[g.post].active;
}

// One tick; returns false on a nonfinite voltage (core.cuh:55).
bool brain_tick(Brain &b, const float *input, float gain, uint8_t *raster) {
    const int n = int(b.genes.size()), m = int(b.edges.size());
    std::fill(b.exc.begin(), b.exc.end(), 0.0f);             // core.cuh:35
    std::fill(b.inh.begin(), b.inh.end(), 0.0f);
    std::fill(b.mod.begin(), b.mod.end(), 0);
    bool ok = true;
    for (int j = 0; j < m; ++j) {                            // core.cuh:37-47
        const SG &g = b.edges[j];
        SS &s = b.synapses[j];
        s.arrived = 0;
        if (!valid_edge(b, g)) continue;
        s.arrived = b.tick >= g.delay && b.history[size_t(g.pre) * b.ring + (b.tick - g.delay) % b.ring];
        if (!s.arrived) continue;
        const NG &src = b.genes[g.pre];
        float r = !src.inhibitory_transmitter || src.modulator ? release(g, s, b.tick) : s.weight_ns;
        if (src.modulator) {
            if (r > 0) {
                b.mod[g.post] |= 1u << (src.modulator - 1);
                uint64_t &st = b.stamps[size_t(g.post) * 6 + (src.modulator - 1) * 2 + (b.tick & 1)];
                st = std::max<uint64_t>(st, uint64_t(b.tick + 1));   // main.cu:669-673
            }
        } else if (src.inhibitory_transmitter) {
            b.inh[g.post] += r;
        } else {
            b.exc[size_t(g.post) * 4 + g.branch] += r * expf(b.neurons[g.post].log_ampa_scale);
        }
    }
    for (int i = 0; i < n; ++i) {                            // core.cuh:49-58, main.cu:710-746
        const NG &g = b.genes[i];
        NS &s = b.neurons[i];
        if (g.active) {
            float current = 0;
            int ch = b.sensory[i];
            if (ch >= 0) {
                float level = input[ch];
                // gym: core.cuh:53 (input_gain * level); world: main.cu:736
                // (level * 230 pA whatever the caller's gain)
                current = b.receptor[i] ? receptor_drive(g, s, level)
                                        : b.world ? level * 230.0f : gain * level;
            }
            membrane_tick(g, s, current, &b.exc[size_t(i) * 4], b.inh[i]);
            if (!std::isfinite(s.voltage_mv)) ok = false;
        } else if (!b.world) {
            s.spike = 0;
        }
        if (g.active || !b.world) b.history[size_t(i) * b.ring + b.tick % b.ring] = s.spike;
        if (raster) raster[i] = s.spike;
    }
    for (int j = 0; j < m; ++j) {                            // core.cuh:60-66, main.cu:748-767
        const SG &g = b.edges[j];
        SS &s = b.synapses[j];
        if (!valid_edge(b, g)) continue;
        if (b.world && b.frozen) continue;
        const NG &a = b.genes[g.pre], &z = b.genes[g.post];
        bool pair = !a.inhibitory_transmitter && excitatory_cell(a.type) && !a.modulator &&
                    !z.inhibitory_transmitter && excitatory_cell(z.type);
        plasticity(g, s, s.arrived, b.neurons[g.post].spike, pair, b.neurons[g.post].recent_depol_mv,
                   b.mod[g.post]);
    }
    return ok;
}

Brain *make_brain(int count, int edges, int channels, int outputs, const float *ng, const float *sg,
                  const int *motors, int max_delay) {
    // gym_adapter/native.cpp:46-82 with the delay bound raised to max_delay
    if (count < 1 || edges < 0 || channels < 1 || channels > 65535 || outputs < 1)
        throw std::runtime_error("Invalid brain dimensions");
    if (max_delay < 8 || max_delay > NS_MAX_DELAY) throw std::runtime_error("Invalid delay bound");
    auto b = new Brain;
    try {
        b->channels = channels;
        b->ring = max_delay + 1;
        for (int i = 0; i < count; ++i) {
            const float *p = ng + size_t(i) * 10;
            if (p[0] < 0 || p[0] > 6 || p[1] < -1 || p[1] >= channels || p[2] < 0 || p[2] > 1 ||
                p[3] < 0 || p[3] > 1 || p[4] < 0 || p[4] > 3)
                throw std::runtime_error("Invalid neuron gene");
            NG g = profile_for(int(uint8_t(p[0])));
            b->sensory.push_back(int(p[1]));
            g.sensory_channel = p[1] < 0 ? -1 : 0;
            g.active = uint8_t(p[2]);
            g.inhibitory_transmitter = uint8_t(p[3]);
            g.modulator = uint8_t(p[4]);
            if (p[5] <= 0 || p[6] <= 0 || p[7] < -80 || p[7] > -20 || p[8] < 0 || p[9] < 0)
                throw std::runtime_error("Invalid membrane parameters");
            g.capacitance_pf *= p[5];
            g.leak_ns *= p[6];
            g.threshold_mv = p[7];
            g.receptor_baseline_hz = 0;
            g.receptor_gain_hz = p[8];
            g.receptor_event_ns = p[9];
            b->genes.push_back(g);
            b->receptor.push_back(g.type == 5);
        }
        for (int i = 0; i < edges; ++i) {
            const float *p = sg + size_t(i) * 11;
            if (p[0] < 0 || p[0] >= count || p[1] < 0 || p[1] >= count || p[2] < 0 || p[3] < 0 ||
                p[3] > 1 || p[4] < 1 || p[4] > max_delay || p[5] < 0 || p[5] >= 4 || p[6] <= 0 ||
                p[6] > 1 || p[7] <= 0 || p[8] < 0 || p[9] < 0 || p[9] > 5 || p[10] < 0 || p[10] > 7)
                throw std::runtime_error("Invalid synapse gene");
            SG g;
            std::memset(&g, 0, sizeof g);
            g.pre = int(p[0]); g.post = int(p[1]); g.initial_weight_ns = p[2];
            g.active = uint8_t(p[3]); g.delay = uint8_t(p[4]); g.branch = uint8_t(p[5]);
            g.utilization = p[6]; g.recovery_ms = p[7]; g.facilitation_ms = p[8];
            g.plasticity = uint8_t(p[9]); g.mod_receptors = uint8_t(p[10]); g.release_sites = 14;
            b->edges.push_back(g);
        }
        for (int i = 0; i < outputs; ++i) {
            if (motors[i] < 0 || motors[i] >= count) throw std::runtime_error("Invalid motor index");
            b->motors.push_back(motors[i]);
        }
    } catch (...) {
        delete b;
        throw;
    }
    return b;
}

void size_scratch(Brain &b) {
    size_t n = b.genes.size();
    b.neurons.resize(n);
    b.synapses.resize(b.edges.size());
    b.history.assign(n * b.ring, 0);
    b.exc.assign(n * 4, 0.0f);
    b.inh.assign(n, 0.0f);
    b.mod.assign(n, 0);
    b.stamps.assign(n * 6, 0);
}

void reset_brain(Brain &b, uint32_t seed) {          // core.cuh:25-29, native.cpp:89-94
    b.tick = 0;
    for (size_t i = 0; i < b.genes.size(); ++i) {
        uint32_t o = b.orig_n.empty() ? uint32_t(i) : b.orig_n[i];
        b.neurons[i] = fresh_neuron(b.genes[i], seed + o * 7919u);
    }
    std::fill(b.history.begin(), b.history.end(), 0);
    std::fill(b.stamps.begin(), b.stamps.end(), 0);
    for (size_t i = 0; i < b.edges.size(); ++i) {
        uint32_t o = b.orig_s.empty() ? uint32_t(i) : b.orig_s[i];
        b.synapses[i] = fresh_synapse(b.edges[i], seed + o * 104729u);
    }
}

// core.cuh:30-75.  Returns false on a nonfinite voltage (the state is then
// left as the reference's GPU batch leaves it: every tick completes).
bool step_brain(Brain &b, const float *input, int ticks, float gain, float *output, uint8_t *raster) {
    bool ok = true;
    const int n = int(b.genes.size());
    for (size_t j = 0; j < b.motors.size(); ++j) output[j] = 0;
    for (int t = 0; t < ticks; ++t, ++b.tick) {
        ok &= brain_tick(b, input, gain, raster ? raster + size_t(t) * n : nullptr);
        if (b.ports.empty()) {
            for (size_t j = 0; j < b.motors.size(); ++j) output[j] += b.neurons[b.motors[j]].spike;
        } else {
            // main.cu:742-743: every active neuron on a motor port counts its spike
            for (int i = 0; i < n; ++i) {
                const int m = b.ports[i] - kWorldSenses;
                if (b.genes[i].active && b.neurons[i].spike && m >= 0 && m < kWorldMotors)
                    output[m] += 1;
            }
        }
    }
    if (!b.world) {
        for (int j = 0; j < n; ++j)                          // core.cuh:70-72
            if (b.genes[j].active)
                homeostasis(b.genes[j], b.neurons[j], float(ticks),
                            b.neurons[j].total_spikes * 1000.0f / float(ticks > b.tick ? ticks : b.tick));
    }
    return ok;
}

}  // namespace

// ===========================================================================
// C API (ctypes)
// ===========================================================================
extern "C" {

const char *ora_error() { return g_error.c_str(); }

void *ora_create(int count, int edges, int channels, int outputs, const float *ng, const float *sg,
                 const int *motors, int max_delay) {
    try {
        Brain *b = make_brain(count, edges, channels, outputs, ng, sg, motors, max_delay);
        size_scratch(*b);
        return b;
    } catch (const std::exception &e) {
        g_error = e.what();
        return nullptr;
    }
}

// Brain from full gene records (world-pool style).  sensory[i] = input index
// or -1; receptor[i] = 1 selects receptor_current for that neuron.
void *ora_create_records(int count, int edges, int channels, const ns_neuron_gene *genes,
                         const ns_synapse_gene *syn, const int *sensory, const uint8_t *receptor,
                         int outputs, const int *motors, int max_delay, int world) {
    auto b = new Brain;
    b->channels = channels;
    b->ring = max_delay + 1;
    b->world = world;
    b->genes.assign(genes, genes + count);
    b->edges.assign(syn, syn + edges);
    b->sensory.assign(sensory, sensory + count);
    b->receptor.assign(receptor, receptor + count);
    b->motors.assign(motors, motors + outputs);
    size_scratch(*b);
    return b;
}

// One animal's brain under the world rules from its checkpoint genes (pre/post
// animal-local) and neuron ports, with the routing main.cu derives from them:
// a sensory channel in [0, 19) reads that sense level (main.cu:729-737),
// through receptor_current for type 5 below channel 6 and as 230 pA per level
// otherwise; motor j counts the spikes of neurons on port 19 + j
// (main.cu:742-743).  ora_step's input is the animal's 19 sense levels and its
// output the 5 motor counts.
void *ora_create_world(int count, int edges, const ns_neuron_gene *genes, const ns_synapse_gene *syn,
                       const int *ports, int max_delay) {
    auto b = new Brain;
    b->channels = kWorldSenses;
    b->ring = max_delay + 1;
    b->world = 1;
    b->genes.assign(genes, genes + count);
    b->edges.assign(syn, syn + edges);
    b->ports.assign(ports, ports + count);
    for (int i = 0; i < count; ++i) {
        const int ch = genes[i].sensory_channel;
        const bool sensed = ch >= 0 && ch < kWorldSenses;
        b->sensory.push_back(sensed ? ch : -1);
        b->receptor.push_back(sensed && genes[i].type == 5 && ch < 6);
    }
    b->motors.assign(kWorldMotors, 0);
    size_scratch(*b);
    return b;
}

void ora_destroy(void *p) { delete static_cast<Brain *>(p); }

// The brain's gene records (neurons, then synapses with brain-local pre/post).
void ora_genes(void *p, ns_neuron_gene *genes, ns_synapse_gene *syn) {
    auto &b = *static_cast<Brain *>(p);
    if (genes) std::memcpy(genes, b.genes.data(), b.genes.size() * sizeof(NG));
    if (syn && !b.edges.empty()) std::memcpy(syn, b.edges.data(), b.edges.size() * sizeof(SG));
}

void ora_original_indices(void *p, const uint32_t *n, const uint32_t *s) {
    auto &b = *static_cast<Brain *>(p);
    b.orig_n.assign(n, n + b.genes.size());
    if (!b.edges.empty()) b.orig_s.assign(s, s + b.edges.size());
}

void ora_reset(void *p, uint32_t seed) { reset_brain(*static_cast<Brain *>(p), seed); }

int ora_step(void *p, const float *input, int ticks, float gain, float *output, uint8_t *raster) {
    return step_brain(*static_cast<Brain *>(p), input, ticks, gain, output, raster) ? 0 : -1;
}

void ora_set_tick(void *p, int64_t tick) { static_cast<Brain *>(p)->tick = tick; }
void ora_set_frozen(void *p, int frozen) { static_cast<Brain *>(p)->frozen = frozen; }

// homeostasis_step (main.cu:1287-1300) for one animal of age `age` steps.
void ora_world_homeostasis(void *p, int age) {
    auto &b = *static_cast<Brain *>(p);
    if (b.frozen) return;
    for (size_t i = 0; i < b.genes.size(); ++i) {
        if (!b.genes[i].active) continue;
        float elapsed = fmaxf(16.0f, age * 16.0f);
        float hz = b.neurons[i].total_spikes * 1000.0f / elapsed;
        homeostasis(b.genes[i], b.neurons[i], 16.0f, hz);
    }
}

// history9: count x 9 bytes in core.cuh's slot = tick % 9 layout, rebuilt
// from the ring.  ring_out: count x ring raw bytes.  stamps: count x 6.
void ora_export(void *p, ns_neuron_state *ns, ns_synapse_state *ss, uint8_t *history9,
                uint8_t *ring_out, int64_t *tick, uint64_t *stamps) {
    auto &b = *static_cast<Brain *>(p);
    const size_t n = b.genes.size();
    if (ns) std::memcpy(ns, b.neurons.data(), n * sizeof(NS));
    if (ss && !b.edges.empty()) std::memcpy(ss, b.synapses.data(), b.edges.size() * sizeof(SS));
    if (history9) {
        for (size_t i = 0; i < n; ++i)
            for (int k = 0; k < 9; ++k) {
                // slot k holds the most recent completed tick tau <= tick-1 with tau % 9 == k
                int64_t last = b.tick - 1;
                int64_t back = ((last - k) % 9 + 9) % 9;
                int64_t tau = last - back;
                history9[i * 9 + k] = tau < 0 ? 0 : b.history[i * b.ring + tau % b.ring];
            }
    }
    if (ring_out) std::memcpy(ring_out, b.history.data(), b.history.size());
    if (tick) *tick = b.tick;
    if (stamps) std::memcpy(stamps, b.stamps.data(), n * 6 * sizeof(uint64_t));
}

void ora_import(void *p, const ns_neuron_state *ns, const ns_synapse_state *ss, const uint8_t *ring,
                int64_t tick) {
    auto &b = *static_cast<Brain *>(p);
    if (ns) std::memcpy(b.neurons.data(), ns, b.genes.size() * sizeof(NS));
    if (ss && !b.edges.empty()) std::memcpy(b.synapses.data(), ss, b.edges.size() * sizeof(SS));
    if (ring) std::memcpy(b.history.data(), ring, b.history.size());
    b.tick = tick;
}

int ora_ring(void *p) { return static_cast<Brain *>(p)->ring; }

// Threaded batch: brains are independent within a call, so they shard over
// host threads without synchronisation (the CPU baseline).
int ora_batch_step(void **brains, int count, const float *inputs, int channels, const uint8_t *mask,
                   int ticks, float gain, float *outputs, int outputs_per_brain, int *status,
                   int threads) {
    if (threads < 1) threads = 1;
    auto work = [&](int lo, int hi) {
        for (int i = lo; i < hi; ++i) {
            float *out = outputs + size_t(i) * outputs_per_brain;
            if (mask && !mask[i]) {
                for (int j = 0; j < outputs_per_brain; ++j) out[j] = 0;
                if (status) status[i] = 0;
                continue;
            }
            bool ok = step_brain(*static_cast<Brain *>(brains[i]), inputs + size_t(i) * channels, ticks,
                                 gain, out, nullptr);
            if (status) status[i] = ok ? 0 : -1;
        }
    };
    if (threads == 1) {
        work(0, count);
        return 0;
    }
    // Interleaved static partition by cumulative synapse+neuron cost.
    std::vector<double> cost(count);
    double total = 0;
    for (int i = 0; i < count; ++i) {
        auto &b = *static_cast<Brain *>(brains[i]);
        cost[i] = double(b.genes.size()) * 8 + double(b.edges.size());
        total += cost[i];
    }
    std::vector<int> cut(threads + 1, count);
    cut[0] = 0;
    double acc = 0;
    int t = 1;
    for (int i = 0; i < count && t < threads; ++i) {
        acc += cost[i];
        if (acc >= total * t / threads) cut[t++] = i + 1;
    }
    std::vector<std::thread> pool;
    for (int k = 0; k < threads; ++k)
        if (cut[k] < cut[k + 1]) pool.emplace_back(work, cut[k], cut[k + 1]);
    for (auto &th : pool) th.join();
    return 0;
}

// ---- the world step around the neural ticks (main.cu:612-657, 1138-1300) --
// A restatement of the C2 world: toroidal (or the reference's clamped) grid,
// food kinds at a density each, regrowth; the reference's atomics replaced
// by ascending animal order: the lowest-index eater takes a cell's item,
// cells list their living animals in ascending index, attackers bite in
// index order within their cell.  Sines and cosines: Cody-Waite reduction by
// pi/2 (fdlibm's three-part constant) and Taylor polynomials in double,
// rounded once to float.
}  // extern "C"
namespace {

void world_sincos(float angle, float *s, float *c) {
    const double x = (double)angle;
    const double k = std::rint(x * 0.63661977236758134308);
    double r = x - k * 1.57079632673412561417e+00;
    r = r - k * 6.07710050650619224932e-11;
    r = r - k * 2.02226624879595063154e-21;
    const double r2 = r * r;
    const double sr = r + r * r2 * (-1.0 / 6 + r2 * (1.0 / 120 + r2 * (-1.0 / 5040 + r2 * (1.0 / 362880 +
                      r2 * (-1.0 / 39916800 + r2 * (1.0 / 6227020800.0))))));
    const double cr = 1.0 + r2 * (-0.5 + r2 * (1.0 / 24 + r2 * (-1.0 / 720 + r2 * (1.0 / 40320 +
                      r2 * (-1.0 / 3628800 + r2 * (1.0 / 479001600.0))))));
    const long long q = (long long)k & 3;
    *s = (float)(q == 0 ? sr : q == 1 ? cr : q == 2 ? -sr : -cr);
    *c = (float)(q == 0 ? cr : q == 1 ? -sr : q == 2 ? -cr : sr);
}

uint64_t eco_random64(uint64_t &s) {                      // main.cu:331-334
    s ^= s >> 12; s ^= s << 25; s ^= s >> 27;
    return s * 2685821657736338717ULL;
}
float eco_uniform(uint64_t &s) { return (float)((eco_random64(s) >> 40) * (1.0 / 16777216.0)); }
uint64_t eco_seed_for(uint64_t seed, uint64_t i) {        // main.cu:341-346
    uint64_t x = seed + 0x9e3779b97f4a7c15ULL * (i + 1);
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return (x ^ (x >> 31)) | 1ULL;
}
float eco_clamp(float x, float lo, float hi) { return fminf(hi, fmaxf(lo, x)); }

struct WorldAnimal {                                      // ns_world_animal
    float x, y, heading, energy, health, fullness, pain, taste[4], muscle[3];
    int32_t age, alive;
};

struct Ecology {
    int P = 0, W = 0, H = 0, torus = 1, instant = 0;
    float regrowth = 0;
    uint64_t seed = 0;
    std::vector<WorldAnimal> an;
    std::vector<int> active, food, claim;
    std::vector<uint8_t> kind, eat, bite;
    std::vector<float> senses, motors;
    std::vector<std::vector<int>> lists;                  // per cell, ascending

    int cell(float x, float y) const {                    // main.cu:347-351 / the torus
        if (torus) {
            int ix = int(floorf(x)) % W, iy = int(floorf(y)) % H;
            if (ix < 0) ix += W;
            if (iy < 0) iy += H;
            return iy * W + ix;
        }
        return int(floorf(eco_clamp(y, 0.0f, float(H) - 0.001f))) * W +
               int(floorf(eco_clamp(x, 0.0f, float(W) - 0.001f)));
    }
    float place(float v, int extent) const {
        if (!torus) return eco_clamp(v, 0.0f, float(extent) - 0.001f);
        if (v < 0.0f) v += float(extent);
        if (v >= float(extent)) v -= float(extent);
        return v;
    }
    void index() {                                        // main.cu:612-617
        for (auto &l : lists) l.clear();
        for (int a = 0; a < P; ++a)
            if (an[a].alive == 1) lists[cell(an[a].x, an[a].y)].push_back(a);
    }
    void sense(int a) {                                   // main.cu:619-657
        const WorldAnimal &b = an[a];
        float *out = &senses[size_t(a) * 19];
        float rs[3], rc[3];
        for (int d = 0; d < 3; ++d) world_sincos(b.heading + float(d - 1) * 0.7f, &rs[d], &rc[d]);
        for (int s = 0; s < 6; ++s) {
            const int d = s % 3;
            float level = 0;
            for (int range = 1; range <= 4; ++range) {
                const int c = cell(b.x + rc[d] * float(range), b.y + rs[d] * float(range));
                if (kind[c] == s / 3 && food[c] > 0) level += 0.25f;
            }
            out[s] = eco_clamp(level, 0, 1);
        }
        out[6] = eco_clamp((16.0f - b.energy) / 16.0f, 0, 1);
        out[7] = eco_clamp(b.fullness / 8.0f, 0, 1);
        out[8] = eco_clamp(b.pain / 3.0f, 0, 1);
        for (int k = 0; k < 3; ++k) out[9 + k] = eco_clamp(b.taste[k], 0, 1);
        for (int d = 0; d < 3; ++d) {
            float level = 0;
            for (int range = 1; range <= 4; ++range) {
                const int c = cell(b.x + rc[d] * float(range), b.y + rs[d] * float(range));
                for (int other : lists[c])
                    if (other != a && an[other].alive == 1) level += 0.25f;
            }
            out[12 + d] = eco_clamp(level, 0, 1);
        }
        out[15] = eco_clamp(b.taste[3], 0, 1);
        const int here = cell(b.x, b.y);
        for (int k = 0; k < 3; ++k) out[16 + k] = food[here] > 0 && kind[here] == k ? 1.0f : 0.0f;
    }
};

}  // namespace
extern "C" {

void *ora_ecology_create(int animals, int width, int height, int torus, float food_density,
                         float regrowth, uint64_t seed, int instant, const int *active_neurons) {
    auto *e = new Ecology;
    e->P = animals; e->W = width; e->H = height; e->torus = torus; e->instant = instant;
    e->regrowth = regrowth; e->seed = seed;
    e->an.assign(animals, WorldAnimal{});
    e->active.assign(active_neurons, active_neurons + animals);
    for (int a = 0; a < animals; ++a) {                   // main.cu:530-545
        uint64_t rng = eco_seed_for(seed ^ 0x5eedf00du, uint64_t(a));
        WorldAnimal &b = e->an[a];
        b.x = eco_uniform(rng) * (torus ? float(width) : float(width - 1));
        b.y = eco_uniform(rng) * (torus ? float(height) : float(height - 1));
        b.heading = eco_uniform(rng) * 2.0f * 3.14159265358979323846f;
        b.energy = 14.0f;
        b.health = 10.0f;
        b.alive = 1;
    }
    const int cells = width * height;
    e->food.assign(cells, 0);
    e->kind.assign(cells, 0);
    for (int c = 0; c < cells; ++c) {                     // main.cu:458-462
        uint64_t rng = eco_seed_for(seed, uint64_t(c));
        e->kind[c] = uint8_t(int(eco_uniform(rng) * 3));
        e->food[c] = eco_uniform(rng) < 3.0f * food_density ? 1 : 0;
    }
    e->claim.assign(cells, 0);
    e->lists.assign(cells, {});
    e->eat.assign(animals, 0);
    e->bite.assign(animals, 0);
    e->senses.assign(size_t(animals) * 19, 0.0f);
    e->motors.assign(size_t(animals) * 5, 0.0f);
    e->index();
    for (int a = 0; a < animals; ++a)
        if (e->an[a].alive == 1) e->sense(a);
    return e;
}

void ora_ecology_destroy(void *p) { delete static_cast<Ecology *>(p); }

// One world step with brains[a] the animal's world-rules oracle brain.
void ora_ecology_step(void *p, void **brains, int64_t step) {
    Ecology &e = *static_cast<Ecology *>(p);
    const int P = e.P;
    for (int a = 0; a < P; ++a) {                         // the 16 neural ticks
        float *m = &e.motors[size_t(a) * 5];
        for (int k = 0; k < 5; ++k) m[k] = 0;
        if (e.an[a].alive == 1)
            step_brain(*static_cast<Brain *>(brains[a]), &e.senses[size_t(a) * 19], 16, 230.0f, m, nullptr);
    }
    std::fill(e.claim.begin(), e.claim.end(), std::numeric_limits<int>::max());
    std::vector<int> food_cell(P, 0);
    for (int a = 0; a < P; ++a) {                         // body_step, main.cu:1146-1213
        WorldAnimal &b = e.an[a];
        if (b.alive != 1) continue;
        const float old_x = b.x, old_y = b.y;
        const float *mo = &e.motors[size_t(a) * 5];
        float force[3];
        for (int m = 0; m < 3; ++m)
            force[m] = e.instant ? mo[m] / 16.0f
                                 : (b.muscle[m] = fminf(1.0f, b.muscle[m] * 0.92311635f + mo[m] * 0.30f));
        const float thrust = force[0], left = force[1], right = force[2];
        e.eat[a] = mo[3] > 0;
        e.bite[a] = mo[4] > 0;
        b.heading += (right - left) * 1.2f;
        float s, c;
        world_sincos(b.heading, &s, &c);
        b.x = e.place(old_x + c * thrust * 1.5f, e.W);
        b.y = e.place(old_y + s * thrust * 1.5f, e.H);
        b.energy -= 0.003f + float(e.active[a]) * 0.00002f + thrust * 0.02f + (e.bite[a] ? 0.02f : 0.0f);
        b.fullness = fmaxf(0, b.fullness - 0.05f);
        b.pain *= 0.8f;
        for (int k = 0; k < 4; ++k) b.taste[k] *= 0.7f;
        food_cell[a] = e.cell(old_x, old_y);
        if (e.eat[a]) e.claim[food_cell[a]] = std::min(e.claim[food_cell[a]], a);
    }
    for (int a = 0; a < P; ++a) {                         // feed, starvation, age, death
        WorldAnimal &b = e.an[a];
        if (b.alive != 1) continue;
        const int fc = food_cell[a];
        if (e.eat[a] && e.claim[fc] == a && e.food[fc] > 0) {
            e.food[fc] -= 1;
            const int kind = e.kind[fc];
            b.taste[kind] = 1;
            b.fullness += 1;
            if (kind == 0) b.energy = fminf(30, b.energy + 4);
            if (kind == 1) { b.energy = fminf(30, b.energy + 2); b.health -= 1.4f; b.pain += 1.4f; }
            if (b.fullness > 8) { b.health -= 0.8f; b.pain += 0.8f; }
        }
        if (b.energy <= 0) { b.energy = 0; b.health -= 0.15f; b.pain += 0.15f; }
        ++b.age;
        if (b.health <= 0 || b.age >= 10000) b.alive = 0;
    }
    e.index();
    for (const auto &l : e.lists)                         // bite_animals, main.cu:1251-1275
        for (int a : l) {
            if (e.an[a].alive != 1 || !e.bite[a]) continue;
            for (int prey : l) {
                if (prey == a || e.an[prey].alive != 1) continue;
                const float previous = e.an[prey].health;
                e.an[prey].health = previous - 2.0f;
                e.an[prey].pain += 2.0f;
                if (previous <= 2.0f) {
                    e.an[prey].alive = -2;
                    e.an[a].energy += fmaxf(0.0f, e.an[prey].energy) * 0.45f;
                    e.an[a].taste[3] = 1.0f;
                    e.an[a].fullness += 1.0f;
                }
                break;
            }
        }
    for (int a = 0; a < P; ++a) {                         // finish_predation; homeostasis_step
        if (e.an[a].alive == -2) e.an[a].alive = 0;
        if (e.an[a].alive == 1) ora_world_homeostasis(brains[a], e.an[a].age);
    }
    for (int c = 0; c < e.W * e.H; ++c) {                 // regrow, main.cu:1138-1144
        if (e.food[c] != 0) continue;
        uint64_t rng = eco_seed_for(e.seed + uint64_t(step), uint64_t(c));
        if (eco_uniform(rng) < e.regrowth) e.food[c] = 1;
    }
    e.index();
    for (int a = 0; a < P; ++a)
        if (e.an[a].alive == 1) e.sense(a);
}

void ora_ecology_export(void *p, void *animals, int32_t *food, uint8_t *kind, float *senses, float *motors) {
    const Ecology &e = *static_cast<Ecology *>(p);
    if (animals) std::memcpy(animals, e.an.data(), e.an.size() * sizeof(WorldAnimal));
    if (food) std::memcpy(food, e.food.data(), e.food.size() * sizeof(int32_t));
    if (kind) std::memcpy(kind, e.kind.data(), e.kind.size());
    if (senses) std::memcpy(senses, e.senses.data(), e.senses.size() * sizeof(float));
    if (motors) std::memcpy(motors, e.motors.data(), e.motors.size() * sizeof(float));
}

float ora_world_sin(float angle) { float s, c; world_sincos(angle, &s, &c); return s; }
float ora_world_cos(float angle) { float s, c; world_sincos(angle, &s, &c); return c; }

// ---- function-level probes (pinned against oracle/_ref's ref_fn_*) --------
void ora_fn_profile(int type, ns_neuron_gene *out) { *out = profile_for(type); }
void ora_fn_initial_neuron(const ns_neuron_gene *g, uint32_t seed, ns_neuron_state *out) {
    *out = fresh_neuron(*g, seed);
}
void ora_fn_initial_synapse(const ns_synapse_gene *g, uint32_t seed, ns_synapse_state *out) {
    *out = fresh_synapse(*g, seed);
}
float ora_fn_receptor_current(const ns_neuron_gene *g, ns_neuron_state *s, float level) {
    return receptor_drive(*g, *s, level);
}
void ora_fn_neuron_advance(const ns_neuron_gene *g, ns_neuron_state *s, float current, const float *exc,
                           float inh) {
    membrane_tick(*g, *s, current, exc, inh);
}
float ora_fn_synapse_release(const ns_synapse_gene *g, ns_synapse_state *s, int64_t tick) {
    return release(*g, *s, tick);
}
void ora_fn_plasticity(const ns_synapse_gene *g, ns_synapse_state *s, int arrived, int post_spike,
                       int pair, float depol, int mod) {
    plasticity(*g, *s, arrived, post_spike, pair, depol, unsigned(mod));
}
void ora_fn_homeostasis(const ns_neuron_gene *g, ns_neuron_state *s, float elapsed, float hz) {
    homeostasis(*g, *s, elapsed, hz);
}
float ora_fn_expf(float x) { return expf(x); }
float ora_fn_branch_io(float c) { return branch_io(c); }
// Number of float bit patterns b in [lo, hi), stepping by `step`, for which
// branch_io(b) is not bitwise +0.
uint64_t ora_branch_io_nonzero(uint32_t lo, uint32_t hi, uint32_t step) {
    uint64_t bad = 0;
    for (uint64_t b = lo; b < hi; b += step) {
        uint32_t u = uint32_t(b), r;
        float c;
        std::memcpy(&c, &u, 4);
        const float y = branch_io(c);
        std::memcpy(&r, &y, 4);
        bad += r != 0;
    }
    return bad;
}
float ora_fn_logf(float x) { return logf(x); }

// ---------------------------------------------------------------------------
// Serving tasks on the CPU (the serving loop's linear and maze tasks, for the
// serving CPU baseline and the tests).  Whole requests run here, one oracle
// brain each, so a thread pool runs them without the caller's interpreter.
// ---------------------------------------------------------------------------
static uint64_t task_splitmix(uint64_t &z) {
    z += 0x9e3779b97f4a7c15ull;
    uint64_t r = z;
    r = (r ^ (r >> 30)) * 0xbf58476d1ce4e5b9ull;
    r = (r ^ (r >> 27)) * 0x94d049bb133111ebull;
    return r ^ (r >> 31);
}

// Linear request: x' = A x + B u + sigma * (sum of 12 uniforms - 6), u the two
// motor counts per tick; return = -sum |x|^2.  Returns the steps run.
int ora_linear_episode(void *p, uint32_t task_seed, int steps, float sigma, int ticks, float gain,
                       float *ret, float *x_out, int32_t *motors) {
    auto &b = *static_cast<Brain *>(p);
    uint64_t z = 0x9e3779b97f4a7c15ull * (uint64_t(task_seed) + 1);
    auto uni53 = [&]() { return float(double(task_splitmix(z) >> 11) * (1.0 / 9007199254740992.0)); };
    float x[4], A[16], B[8];
    for (int i = 0; i < 4; ++i) x[i] = 2.0f * uni53() - 1.0f;
    for (int i = 0; i < 4; ++i) {
        float row[4], norm = 0.0f;
        for (int j = 0; j < 4; ++j) { row[j] = 2.0f * uni53() - 1.0f; norm += std::fabs(row[j]); }
        for (int j = 0; j < 4; ++j) A[i * 4 + j] = 0.9f * row[j] / norm;
    }
    for (int k = 0; k < 8; ++k) B[k] = uni53() - 0.5f;
    uint32_t s = task_seed * 2654435761u + 12345u;
    if (!s) s = 1;
    auto uni24 = [&]() {
        s ^= s << 13; s ^= s >> 17; s ^= s << 5;
        return (s >> 8) * (1.0f / 16777216.0f);
    };
    float r = 0.0f, m0 = 0.0f, m1 = 0.0f, in[8], out[2];
    for (int k = 0; k < steps; ++k) {
        for (int d = 0; d < 4; ++d) {
            in[2 * d] = std::fmin(1.0f, std::fmax(0.0f, x[d]));
            in[2 * d + 1] = std::fmin(1.0f, std::fmax(0.0f, -x[d]));
        }
        step_brain(b, in, ticks, gain, out, nullptr);
        const float u0 = out[0] / float(ticks), u1 = out[1] / float(ticks);
        float nx[4];
        for (int i = 0; i < 4; ++i) {
            float acc = 0.0f;
            for (int j = 0; j < 4; ++j) acc += A[i * 4 + j] * x[j];
            acc += B[i * 2] * u0;
            acc += B[i * 2 + 1] * u1;
            float n = -6.0f;
            for (int q = 0; q < 12; ++q) n += uni24();
            nx[i] = acc + sigma * n;
        }
        float cost = 0.0f;
        for (int i = 0; i < 4; ++i) { x[i] = nx[i]; cost += nx[i] * nx[i]; }
        r -= cost;
        m0 += out[0];
        m1 += out[1];
    }
    *ret = r;
    for (int i = 0; i < 4; ++i) x_out[i] = x[i];
    motors[0] = int32_t(m0);
    motors[1] = int32_t(m1);
    return steps;
}

// Depth-first 31 x 31 maze `index` of a request seeded `seed` (rooms at odd
// coordinates, carve from (1, 1), neighbours north/east/south/west, one drawn
// as SplitMix64 mod count; goal room drawn from the other 224).
static void task_maze(uint32_t seed, uint32_t index, uint32_t rows[31], int &gx, int &gy) {
    uint64_t z = (0x9e3779b97f4a7c15ull * (uint64_t(seed) + 1)) ^ (uint64_t(index) * 0xd1b54a32d192ed03ull);
    for (int y = 0; y < 31; ++y) rows[y] = 0;
    std::vector<uint8_t> seen(225, 0);
    std::vector<int> stack{0};
    rows[1] |= 1u << 1;
    seen[0] = 1;
    while (!stack.empty()) {
        const int r = stack.back(), rx = r % 15, ry = r / 15;
        int cand[4], nc = 0;
        if (ry > 0 && !seen[r - 15]) cand[nc++] = r - 15;
        if (rx < 14 && !seen[r + 1]) cand[nc++] = r + 1;
        if (ry < 14 && !seen[r + 15]) cand[nc++] = r + 15;
        if (rx > 0 && !seen[r - 1]) cand[nc++] = r - 1;
        if (!nc) { stack.pop_back(); continue; }
        const int q = cand[task_splitmix(z) % uint64_t(nc)];
        const int ax = 1 + 2 * rx, ay = 1 + 2 * ry, bx = 1 + 2 * (q % 15), by = 1 + 2 * (q / 15);
        rows[(ay + by) / 2] |= 1u << ((ax + bx) / 2);
        rows[by] |= 1u << bx;
        seen[q] = 1;
        stack.push_back(q);
    }
    const int g = 1 + int(task_splitmix(z) % 224);
    gx = 1 + 2 * (g % 15);
    gy = 1 + 2 * (g / 15);
}

void ora_maze_layout(uint32_t seed, uint32_t index, uint32_t *rows, uint32_t *goal) {
    int gx, gy;
    task_maze(seed, index, rows, gx, gy);
    goal[0] = uint32_t(gx);
    goal[1] = uint32_t(gy);
}

// Maze request: `mazes` mazes in a row, the brain kept; observation walls
// N/E/S/W, goal E/W/S/N, reward; most motor spikes moves N/E/S/W (ties to
// the lower motor, none = stay); the goal pays 1 - 0.9 steps / horizon.
// state4 = {solved, steps, mazes played, 0}.  Returns the steps run.
int ora_maze_request(void *p, uint32_t task_seed, int mazes, int horizon, int ticks, float gain,
                     float *ret, float *state4, int32_t *motors) {
    auto &b = *static_cast<Brain *>(p);
    uint32_t rows[31];
    int gx, gy, x = 1, y = 1, steps = 0, total = 0, solved = 0, index = 0;
    bool pending = false;
    task_maze(task_seed, 0, rows, gx, gy);
    auto open = [&](int cx, int cy) { return cx >= 0 && cx < 31 && cy >= 0 && cy < 31 && ((rows[cy] >> cx) & 1u); };
    float r = 0.0f, mot[4] = {0, 0, 0, 0}, in[9], out[4];
    while (index < mazes) {
        in[0] = open(x, y - 1) ? 0.0f : 1.0f;
        in[1] = open(x + 1, y) ? 0.0f : 1.0f;
        in[2] = open(x, y + 1) ? 0.0f : 1.0f;
        in[3] = open(x - 1, y) ? 0.0f : 1.0f;
        in[4] = gx > x ? 1.0f : 0.0f;
        in[5] = gx < x ? 1.0f : 0.0f;
        in[6] = gy > y ? 1.0f : 0.0f;
        in[7] = gy < y ? 1.0f : 0.0f;
        in[8] = pending ? 1.0f : 0.0f;
        step_brain(b, in, ticks, gain, out, nullptr);
        pending = false;
        int best = -1;
        float top = 0.0f;
        for (int m = 0; m < 4; ++m) {
            if (out[m] > top) { top = out[m]; best = m; }
            mot[m] += out[m];
        }
        if (best >= 0) {
            const int nx = x + (best == 1) - (best == 3), ny = y + (best == 2) - (best == 0);
            if (open(nx, ny)) { x = nx; y = ny; }
        }
        ++steps;
        ++total;
        bool next = false;
        if (x == gx && y == gy) {
            r += float(1.0 - 0.9 * (double(steps) / double(horizon)));
            ++solved;
            pending = true;
            next = true;
        } else if (steps >= horizon) {
            next = true;
        }
        if (!next) continue;
        if (++index >= mazes) break;
        task_maze(task_seed, uint32_t(index), rows, gx, gy);
        x = y = 1;
        steps = 0;
    }
    *ret = r;
    state4[0] = float(solved);
    state4[1] = float(total);
    state4[2] = float(index);
    state4[3] = 0.0f;
    motors[0] = int32_t(mot[0]);
    motors[1] = int32_t(mot[1]);
    return total;
}

// World structs of the checkpoint container, field lists as src/main.cu:66-104
// and 1845-1849 declare them (S = 19, M = 5), laid out by this compiler:
// out = {sizeof Animal, offsetof alive, neuron_count, synapse_count,
// sizeof BehaviorTotals, LifeRecord, BiteEvent, PlasticityCohort, Behavior,
// CheckpointHeader (checkpoint_impl.cuh:6-19)}.
void ora_world_layout(uint64_t *out) {
    struct Behavior {
        int opportunities[3][2][3], accepted[3][2][3];
        int empty_steps[3][2], empty_eat[3][2];
        float approach_progress[3][3], approach_travel[3][3];
        int first_b_age, offspring, death_cause;
    };
    struct Animal {
        float x, y, heading, energy, health, fullness, pain;
        float taste[4], muscle[3];
        unsigned long long rng;
        int id, parent_id, generation, age, alive;
        int neuron_head, synapse_head, neuron_count, synapse_count;
        int ports[19 + 5];
        Behavior behavior;
    };
    struct LifeRecord {
        Animal body;
        int64_t death_step;
        int active_neurons, active_synapses, modulatory_neurons, plastic_synapses;
        float mean_abs_weight_change;
    };
    struct BiteEvent {
        int64_t step;
        int attacker_id, prey_id;
        float x, y;
    };
    struct PlasticityCohort {
        int id, assigned, frozen, initial_age, initial_offspring;
        int final_age, final_offspring, alive, death_cause;
        int neurons, synapses, generation;
        float initial_energy, initial_health;
        int64_t start_step, observed_step, death_step;
    };
    struct BehaviorTotals {
        uint64_t opportunities[3][2][3], accepted[3][2][3];
        uint64_t empty_steps[3][2], empty_eat[3][2];
        double approach_progress[3][3], approach_travel[3][3];
    };
    struct Header {
        char magic[8];
        uint32_t version, endian;
        uint64_t layout;
        int64_t step, checkpoint_every;
        unsigned long long seed;
        int initial, trace_animal;
        float food_renewal;
        uint32_t flags, neuron_count, synapse_count;
        uint64_t csv_size, behavior_size, lives_size;
        char output[512];
    };
    out[0] = sizeof(Animal);
    out[1] = offsetof(Animal, alive);
    out[2] = offsetof(Animal, neuron_count);
    out[3] = offsetof(Animal, synapse_count);
    out[4] = sizeof(BehaviorTotals);
    out[5] = sizeof(LifeRecord);
    out[6] = sizeof(BiteEvent);
    out[7] = sizeof(PlasticityCohort);
    out[8] = sizeof(Behavior);
    out[9] = sizeof(Header);
}

// CheckpointNeuron (main.cu:164-171) laid out for a ring of HISTORY slots
// (main.cu:46 HISTORY = D + 1): out = {sizeof at 9 slots, sizeof at 17,
// offsetof spike_history}.
}  // extern "C"
template <int HISTORY> struct CheckpointNeuronOf {
    int slot, owner, next, port;
    ns_neuron_gene gene;
    ns_neuron_state state;
    float pending_exc[4], pending_inh;
    unsigned long long mod_event_tick[6];
    unsigned char spike_history[HISTORY];
};
extern "C" {
void ora_checkpoint_neuron_layout(uint64_t *out) {
    out[0] = sizeof(CheckpointNeuronOf<9>);
    out[1] = sizeof(CheckpointNeuronOf<17>);
    out[2] = offsetof(CheckpointNeuronOf<17>, spike_history);
}

}  // extern "C"
