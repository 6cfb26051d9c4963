// Device pool layout and per-element physiology for the tick kernels.
//
// Pools are structure-of-arrays in HBM: field k of element i lives at
// base[k * capacity + i], so a wavefront touching 64 consecutive neurons or
// synapses reads each field as one contiguous 256-byte run.  Brains occupy
// contiguous segments: neurons in genome order, synapses sorted by
// (post, original edge index).  That sort turns each neuron's incoming
// synapses into one contiguous run whose order is the reference's
// accumulation order (ascending edge index, gym_adapter/core.cuh:37-47,
// gym_adapter/gpu.cu:40-47), so a sequential per-neuron sum over the run
// rounds exactly like the reference's CPU loop — no float atomics.
//
// The arithmetic below restates src/physiology.cuh (cited per function) for
// register-resident state.  Expression shapes and evaluation order are kept
// operation for operation; the file must be compiled with -ffp-contract=off.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ns_math.h"

namespace nsa {

// ---- neuron gene fields (float, stride = neuron capacity) -------------------
enum NeuronGeneField {
    NG_CAP, NG_LEAK, NG_REST, NG_THR, NG_SLOPE, NG_A, NG_TAUW, NG_B, NG_RESET,
    NG_ADP_SPIKE, NG_ADP_TAU, NG_RBASE, NG_RGAIN, NG_REVENT, NG_RDECAY, NG_FIELDS
};
// extra genes for LP (type 2) and Prinz (type 3): lp_scale[4], prinz_g[8], ambient
enum { NX_LP = 0, NX_PRINZ = 4, NX_AMBIENT = 12, NX_FIELDS = 13 };
// ---- neuron state fields (float) --------------------------------------------
enum NeuronStateField {
    NS_V, NS_W, NS_ADP, NS_INH, NS_DEPOL, NS_LAS, NS_D0, NS_D1, NS_D2, NS_D3, NS_RNS,
    NS_FIELDS
};
constexpr int kCellFields = 16;   // union LP/Prinz/LFS
// ---- neuron flag bits --------------------------------------------------------
constexpr uint32_t NF_TYPE = 0xffu;
constexpr uint32_t NF_ACTIVE = 1u << 8;
constexpr uint32_t NF_RECEPTOR = 1u << 9;   // sensory input through receptor events
constexpr uint32_t NF_WORLD = 1u << 10;     // inactive neurons keep their ring slot
constexpr int NF_RING_SHIFT = 11;           // bits 11-15: world ring length - 1
constexpr uint32_t NF_RING = 31u << NF_RING_SHIFT;

// Class of a neuron slot inside a pack: cortical AdEx (0), receptor-driven
// (1), other AdEx (2), LFS (3), inactive (4).  Packs order their slots by
// class (stable), so a wavefront mostly runs one integrator.
constexpr int kSlotClasses = 5;
__host__ __device__ inline int slot_class(uint32_t f) {
    const uint32_t ty = f & NF_TYPE;
    return !(f & NF_ACTIVE) ? 4 : ty == 0 ? 0 : ty == 5 ? 1 : ty == 6 ? 3 : 2;
}

// ---- synapse gene fields (float) ---------------------------------------------
enum SynapseGeneField { SG_W0, SG_U, SG_REC, SG_FAC, SG_SREC, SG_SDEP, SG_FIELDS };
// ---- synapse state fields (float) --------------------------------------------
enum SynapseStateField {
    SS_WEIGHT, SS_UTIL, SS_PTRACE, SS_QTRACE, SS_ELTP, SS_ELTD, SS_DA_AGE, SS_SPINE,
    SS_SAGE, SS_PRE_AGE, SS_POST_AGE, SS_SLOW, SS_FIELDS
};
// ---- synapse flag word ----------------------------------------------------------
//  bits 0-7 delay | 8-10 plasticity | 11-13 modulator receptors | 14-20 sites
//  bits 24-31: route byte = branch(2) | inh(1) | modulator(2) | valid | stp | pair
constexpr uint32_t SF_DELAY = 0xffu;
constexpr int SF_PLAST_SHIFT = 8, SF_MODREC_SHIFT = 11, SF_SITES_SHIFT = 14, SF_ROUTE_SHIFT = 24;
constexpr uint32_t R_BRANCH = 3u, R_INH = 1u << 2, R_MOD_SHIFT = 3, R_MOD = 3u << 3;
constexpr uint32_t R_VALID = 1u << 5, R_STP = 1u << 6, R_PAIR = 1u << 7;

constexpr unsigned kModNE = 1, kMod5HT = 2, kModDA = 4;
// Global modulator byte (Pool::modbits) of a cell without sensory input whose
// last tick of the previous call changed no state (k_resident_lds; every other
// writer of the byte stores modulator bits only, which clears it).
constexpr unsigned kStillCarry = 0x80u;
constexpr float kDepolGateMv = 2.3f;   // profile-2 LTP gate on the post's recent depolarization

// World-rule ring length (src/main.cu:45-46 HISTORY = D + 1): 9 slots for the
// reference's delay bound D = 8, D + 1 for a bound of 9..16 (the 1-16 ms
// delays of the 256 x 256 world).  An inactive neuron keeps its ring slot
// (main.cu:718), so its shift register rotates: the bit of HISTORY ticks ago
// becomes the bit of this tick.
constexpr int kWorldHistory = 9;
__host__ __device__ inline int world_ring(int max_delay) {
    return max_delay + 1 > kWorldHistory ? max_delay + 1 : kWorldHistory;
}
__host__ __device__ inline uint32_t world_flags(int max_delay) {
    return NF_WORLD | (uint32_t(world_ring(max_delay) - 1) << NF_RING_SHIFT);
}
__host__ __device__ inline uint32_t world_rotate(uint32_t h, uint32_t flags) {
    return (h << 1) | ((h >> ((flags & NF_RING) >> NF_RING_SHIFT)) & 1u);
}

// ---------------------------------------------------------------------------
// Initial states (physiology.cuh:179-202, 393-404), shared by the device reset
// kernels and the host-side genome handles.
// ---------------------------------------------------------------------------
__host__ __device__ inline uint32_t mix_seed(uint32_t seed) {   // physiology.cuh:183-186
    uint32_t x = seed ? seed : 1;
    x ^= x >> 16; x *= 0x7feb352du;
    x ^= x >> 15; x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Cell union of a fresh neuron of `type` (LP: lp_cell.cuh initial values;
// Prinz: prinz_cell.cuh; LFS: voltage at rest, threshold at the gene's);
// returns its initial voltage.
__host__ __device__ inline float initial_cell(uint32_t type, float rest, float thr, float *cell) {
    for (int k = 0; k < 16; ++k) cell[k] = 0.0f;
    float v = rest;
    if (type == 2) {
        const float lp[16] = {-56, 0, .2f, 0, 1, 0, 0, 1, 0, 0, 1, 1, 0, .04f, .01f, -56};
        for (int k = 0; k < 16; ++k) cell[k] = lp[k];
        v = cell[15] - 9.342518882f;
    } else if (type == 3) {
        cell[0] = -50.0f; cell[1] = .05f;
        for (int k = 0; k < 4; ++k) cell[9 + k] = 1.0f;
        v = cell[0];
    } else if (type == 6) {
        cell[0] = rest;
        cell[1] = thr;
    }
    return v;
}

// xorshift generator state of a fresh neuron or synapse seeded `seed` (never
// 0: physiology.cuh:188, 399).
__host__ __device__ inline uint32_t initial_rng(uint32_t seed) {
    const uint32_t x = mix_seed(seed);
    return x ? x : 1;
}

// Ready mask of `sites` release sites (physiology.cuh:402-403).
__host__ __device__ inline uint64_t initial_ready(int sites) {
    return sites == 64 ? ~uint64_t(0) : (uint64_t(1) << sites) - 1;
}

struct Pool {
    int64_t ncap, scap;
    // neurons
    const float *ng;          // NG_FIELDS x ncap
    const float *nx;          // NX_FIELDS x ncap
    const uint32_t *nflags;   // ncap
    const int32_t *ninput;    // absolute input index or -1
    const int32_t *nin;       // ncap + 1: first incoming synapse position
    const int32_t *nbrain;    // owning brain
    float *ns;                // NS_FIELDS x ncap
    float *cell;              // kCellFields x ncap
    int32_t *total;
    uint32_t *rrng;
    uint32_t *hist;           // spike shift register: bit k = spike k+1 ticks ago
    uint8_t *spike;
    uint8_t *modbits;         // per-tick modulator bits (plain form)
    uint64_t *stamps;         // 6 x ncap, src/main.cu:118-120 mod_event_tick
    // synapses
    const float *sg;          // SG_FIELDS x scap
    const uint32_t *sflags;
    const int32_t *spre, *spost, *sbrain;
    float *ss;                // SS_FIELDS x scap
    int64_t *last;
    uint64_t *ready;
    uint32_t *srng;
    uint8_t *arrived;
    float *released;          // per-tick release (plain form)
    // brains
    const int32_t *bn0, *bnc, *bs0, *bsc, *bin0, *bout0;
    int64_t *btick;
    uint8_t *brun;            // brain advances this call
    uint8_t *bfrozen;         // plasticity and homeostasis frozen (world)
    int32_t *bstatus;
    const int32_t *motor;     // absolute neuron position per output slot
    int outputs;              // outputs per brain
    unsigned long long *events;   // arrivals: [0] resident launches, [1] per-tick
                                  // launches; [2] neuron ticks the resident launch skipped
};

// ---------------------------------------------------------------------------
// Neuron registers
// ---------------------------------------------------------------------------
struct NeuronRegs {
    float v, w, adp, inh, depol, d0, d1, d2, d3, rns;
    int32_t total;
    uint32_t rrng;
};

struct NeuronConst {
    float cap, leak, rest, thr, slope, a, tauw, b, reset, adp_spike;
    float leak_slope;       // g.leak_ns * g.slope_mv (physiology.cuh:355)
    float adp_decay;        // expf(-h / afterdepol_tau_ms) when tau > 0
    float rbase, rgain, revent;
    float rdecay;           // expf(-1 / receptor_decay_ms)
    uint32_t flags;
};

__device__ inline float clamp_range(float x, float lo, float hi) {   // physiology.cuh:78-80
    return fminf(hi, fmaxf(lo, x));
}

// Correctly rounded x / c from rc = RN_double(1/c).  The double product is
// within 2^-52 (relative) of x/c, while a quotient of two floats that is not
// itself a float midpoint lies at least 2^-49 (relative) from every float
// rounding midpoint, so rounding the product to float gives RN(x/c) for all
// normal, infinite and zero results.  Nonzero results below 2^-126 (where
// exact subnormal midpoints exist) take the IEEE division.  Checked for
// every float x by tests/test_division.py.
__device__ inline float div_const(float x, float c, double rc) {
    const double q = (double)x * rc;
    float r = (float)q;
    if (__builtin_expect(__builtin_fabs(q) < 0x1p-126 && x != 0.0f, 0)) r = x / c;
    return r;
}

// Correctly rounded 5 / d for d >= 1: reciprocal estimate refined by two
// Newton steps in double (error ~2^-52 from any start within a few float
// ulps), times 5, rounded once to float; same argument as div_const.
__device__ inline float div5(float d) {
    if (!(d >= 1.0f && d < 3.0e38f)) return 5.0f / d;
    const double dd = d;
    double r = (double)__builtin_amdgcn_rcpf(d);
    double e = __builtin_fma(-dd, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-dd, r, 1.0);
    r = __builtin_fma(r, e, r);
    return (float)(5.0 * r);
}

constexpr double kRcp04 = 1.0 / double(0.4f);

__device__ inline void load_neuron(const Pool &p, int64_t i, NeuronRegs &s, NeuronConst &g,
                                   const uint64_t *etab) {
    const int64_t N = p.ncap;
    s.v = p.ns[NS_V * N + i];
    s.w = p.ns[NS_W * N + i];
    s.adp = p.ns[NS_ADP * N + i];
    s.inh = p.ns[NS_INH * N + i];
    s.depol = p.ns[NS_DEPOL * N + i];
    s.d0 = p.ns[NS_D0 * N + i];
    s.d1 = p.ns[NS_D1 * N + i];
    s.d2 = p.ns[NS_D2 * N + i];
    s.d3 = p.ns[NS_D3 * N + i];
    s.rns = p.ns[NS_RNS * N + i];
    s.total = p.total[i];
    s.rrng = p.rrng[i];
    g.cap = p.ng[NG_CAP * N + i];
    g.leak = p.ng[NG_LEAK * N + i];
    g.rest = p.ng[NG_REST * N + i];
    g.thr = p.ng[NG_THR * N + i];
    g.slope = p.ng[NG_SLOPE * N + i];
    g.a = p.ng[NG_A * N + i];
    g.tauw = p.ng[NG_TAUW * N + i];
    g.b = p.ng[NG_B * N + i];
    g.reset = p.ng[NG_RESET * N + i];
    g.adp_spike = p.ng[NG_ADP_SPIKE * N + i];
    g.leak_slope = g.leak * g.slope;
    float tau = p.ng[NG_ADP_TAU * N + i];
    g.adp_decay = tau > 0 ? ns_expf(-0.25f / tau, etab) : 1.0f;
    g.flags = p.nflags[i];
    if (g.flags & NF_RECEPTOR) {
        g.rbase = p.ng[NG_RBASE * N + i];
        g.rgain = p.ng[NG_RGAIN * N + i];
        g.revent = p.ng[NG_REVENT * N + i];
        g.rdecay = ns_expf(-1.0f / p.ng[NG_RDECAY * N + i], etab);
    } else {
        g.rbase = g.rgain = g.revent = 0; g.rdecay = 1;
    }
    // tau > 0 is tested per substep in the reference (physiology.cuh:368); a
    // negative adp_tau gives "no decay" there too, so the flag below encodes it.
    if (!(tau > 0)) g.flags |= 1u << 31;
}

__device__ inline void store_neuron(const Pool &p, int64_t i, const NeuronRegs &s) {
    const int64_t N = p.ncap;
    p.ns[NS_V * N + i] = s.v;
    p.ns[NS_W * N + i] = s.w;
    p.ns[NS_ADP * N + i] = s.adp;
    p.ns[NS_INH * N + i] = s.inh;
    p.ns[NS_DEPOL * N + i] = s.depol;
    p.ns[NS_D0 * N + i] = s.d0;
    p.ns[NS_D1 * N + i] = s.d1;
    p.ns[NS_D2 * N + i] = s.d2;
    p.ns[NS_D3 * N + i] = s.d3;
    p.ns[NS_RNS * N + i] = s.rns;
    p.total[i] = s.total;
    p.rrng[i] = s.rrng;
}

// physiology.cuh:206-219 (receptor_current), called only for neurons with a
// sensory input and the receptor flag.
__device__ inline float receptor_drive(NeuronRegs &s, const NeuronConst &g, float level) {
    if (g.revent <= 0) return 0;
    float rate = clamp_range(g.rbase + g.rgain * level, 0, 1000);
    uint32_t r = s.rrng;
    r ^= r << 13; r ^= r >> 17; r ^= r << 5;
    s.rrng = r;
    float u = (r >> 8) * (1.0f / 16777216.0f);
    s.rns *= g.rdecay;
    if (u < rate * (1.0f / 1000.0f)) s.rns += g.revent;
    return s.rns * (0.0f - s.v);
}

// Dendritic conductances below this give branch_io exactly +0 (see adex_tick).
constexpr float kBranchQuiet = 0x1p-37f;

// physiology.cuh:226-230
__device__ inline float branch_io(float c, const uint64_t *etab) {
    float regen = div5(1.0f + ns_expf(div_const(-(c - 4.0f), 0.4f, kRcp04), etab));
    return fminf(11.0f, c + regen - kBranchOffset);
}

// physiology.cuh:290-374, AdEx branch (types 0, 1, 4, 5).  Returns the spike.
// One exact shortcut: branch_io(c) is exactly +0 for 0 <= c < 2^-37: c - 4
// rounds to -4, so the regenerative term equals its value at c = 0, which is
// the subtracted offset (a float in [2^-13, 2^-12), whose half ulp is 2^-37),
// and c + offset rounds to offset.  Such a dendrite adds +0 to a sum that is
// never -0, so it is skipped.  The adaptation update runs in every substep,
// as in the reference (a branch around it costs more than the division it
// would save, and it is what a zero or nan time constant needs).
//
// still (when not null) receives whether the tick ended with v and w as it
// began them, without a spike, with the conductances, the inhibition and a
// decaying afterdepolarization all +0 throughout (checked on entry: they then
// stay +0).  The tick is a function of its starting state and its inputs, so
// the same inputs then give the same tick again, whatever v and w did
// between its substeps.
__device__ inline int adex_tick(NeuronRegs &s, const NeuronConst &g, float current,
                                const uint64_t *etab, bool *still = nullptr) {
    const float h = 0.25f;
    const bool cortical = (g.flags & NF_TYPE) == 0;
    const bool adp_on = !(g.flags & (1u << 31));
    int spike = 0;
    bool same = true;
    const uint32_t v_start = __float_as_uint(s.v), w_start = __float_as_uint(s.w);
    if (still)
        same = (__float_as_uint(s.d0) | __float_as_uint(s.d1) | __float_as_uint(s.d2) |
                __float_as_uint(s.d3) | __float_as_uint(s.inh) |
                (adp_on ? __float_as_uint(s.adp) : 0u)) == 0u;
#pragma unroll
    for (int step = 0; step < 4; ++step) {
        float v = s.v;
        s.w += h * (g.a * (v - g.rest) - s.w) / g.tauw;
        float x = clamp_range((v - g.thr) / g.slope, -20.0f, 10.0f);
        float soma = 0;
        if (cortical) {
            if (!(s.d0 >= 0.0f && s.d0 < kBranchQuiet)) soma += branch_io(s.d0, etab);
            if (!(s.d1 >= 0.0f && s.d1 < kBranchQuiet)) soma += branch_io(s.d1, etab);
            if (!(s.d2 >= 0.0f && s.d2 < kBranchQuiet)) soma += branch_io(s.d2, etab);
            if (!(s.d3 >= 0.0f && s.d3 < kBranchQuiet)) soma += branch_io(s.d3, etab);
        } else {
            soma += s.d0; soma += s.d1; soma += s.d2; soma += s.d3;
        }
        // x is clamped to [-20, 10] (a nan v clamps too): ns_expf's range
        float dv = g.leak * (g.rest - v) + g.leak_slope * ns_expf_finite(x, etab) +
                   soma * (0.0f - v) + s.inh * (-75.0f - v) + current + s.adp - s.w -
                   0.0f * (v - -68.0f);
        s.v = v + h * dv / g.cap;
        if (s.v >= 0.0f) {
            spike = 1;
            ++s.total;
            s.v = g.reset;
            s.w += g.b;
            s.adp += g.adp_spike;
        }
        if (adp_on) s.adp *= g.adp_decay;
        s.d0 *= kDecayExcQuarter; s.d1 *= kDecayExcQuarter;
        s.d2 *= kDecayExcQuarter; s.d3 *= kDecayExcQuarter;
        s.inh *= kDecayInhQuarter;
    }
    if (still)
        *still = same && !spike && __float_as_uint(s.v) == v_start && __float_as_uint(s.w) == w_start;
    return spike;
}

// physiology.cuh:258-287 LFS cell (type 6); cell state in global memory
// (cell[k * ncap + i]: voltage, threshold, activated[3], open[3]).
__device__ inline int lfs_tick(NeuronRegs &s, const NeuronConst &g, float *cell, int64_t N,
                               int64_t i, float current, float exc, float inh) {
    const float w[3] = {250.0f, 10.0f, 4.0f};
    const float open_ms[3] = {10.0f, 40.0f, 750.0f};
    const float close_ms[3] = {50.0f, 750.0f, 2000.0f};
    const float decay[3] = {kDecayInhMs, kDecayOpen40, kDecayOpen750};
    float volt = cell[0 * N + i], thr = cell[1 * N + i];
    float act[3], open[3];
    for (int j = 0; j < 3; ++j) { act[j] = cell[(2 + j) * N + i]; open[j] = cell[(5 + j) * N + i]; }
    thr = g.thr + (thr - g.thr) * kDecayInhMs;
    float under = 0;
    for (int j = 0; j < 3; ++j) {
        open[j] += act[j] / open_ms[j] - open[j] / close_ms[j];
        act[j] *= decay[j];
        under += w[j] * open[j];
    }
    float v = volt;
    float cur = g.leak * (g.rest - v) + exc * (10.0f - v) + inh * (-60.0f - v) +
                under * (-80.0f - v) + current;
    volt = v + cur / g.cap;
    int spike = volt >= thr;
    if (spike) {
        thr = 200.0f;
        for (int j = 0; j < 3; ++j) act[j] += 1.0f;
        ++s.total;
    }
    s.v = volt;
    cell[0 * N + i] = volt;
    cell[1 * N + i] = thr;
    for (int j = 0; j < 3; ++j) { cell[(2 + j) * N + i] = act[j]; cell[(5 + j) * N + i] = open[j]; }
    return spike;
}

// physiology.cuh:290-302 prelude and the per-type dispatch of neuron_advance
// for the register-resident types (0, 1, 4, 5, 6).  Types 2 and 3 run in
// slow_neuron_tick.  Returns the spike.
//
// still (when not null): whether the tick changed no state (AdEx types only;
// an LFS cell reports false).  With the same inputs, a tick that changed no
// state changes none the next time either: the tick is a function of the
// state and its inputs.
__device__ inline int neuron_tick(NeuronRegs &s, const NeuronConst &g, float current, float e0,
                                  float e1, float e2, float e3, float inh, float *cell, int64_t N,
                                  int64_t i, const uint64_t *etab, bool *still = nullptr) {
    s.d0 += e0; s.d1 += e1; s.d2 += e2; s.d3 += e3;
    float sub = fmaxf(0.0f, fminf(s.v, -50.0f) - g.rest);
    const uint32_t depol_in = __float_as_uint(s.depol);
    s.depol += 0.1f * (sub - s.depol);
    s.inh += inh;
    if ((g.flags & NF_TYPE) == 6) {
        float e = 0;
        e += s.d0; e += s.d1; e += s.d2; e += s.d3;
        int spike = lfs_tick(s, g, cell, N, i, current, e, s.inh);
        s.d0 *= kDecayExcMs; s.d1 *= kDecayExcMs; s.d2 *= kDecayExcMs; s.d3 *= kDecayExcMs;
        s.inh *= kDecayInhMs;
        if (still) *still = false;
        return spike;
    }
    const int spike = adex_tick(s, g, current, etab, still);
    if (still) *still = *still && __float_as_uint(s.depol) == depol_in;
    return spike;
}

// homeostatic_advance, physiology.cuh:243-253.
__device__ inline float homeostasis_las(float las, float elapsed_ms, float observed_hz) {
    float ratio = clamp_range(observed_hz / 0.4f, 0.0f, 2.65f);
    float change = ratio < 1.0f ? kLog192 * (1.0f - ratio) : kLog070 * ((ratio - 1.0f) / 1.65f);
    return clamp_range(las + change * (elapsed_ms / 172800000.0f), -3.0f, 3.0f);
}

// ---------------------------------------------------------------------------
// LP (type 2) and Prinz (type 3): lp_cell.cuh, prinz_cell.cuh.  State in
// global memory; these run in their own kernel so their register demand does
// not touch the resident path.
// ---------------------------------------------------------------------------
__device__ inline float lp_sig(float v, float mid, float slope, const uint64_t *et) {
    return 1.0f / (1.0f + ns_expf((v - mid) / slope, et));
}
__device__ inline float lp_ratio(float num, float den, float lim) {
    return fabsf(den) < 1e-5f ? lim : num / den;
}

// lp_cell.cuh:41-96
__device__ inline void lp_rhs(const float *x, float ina, const float *sc, float *d,
                              const uint64_t *et) {
    const float va = x[0], vs = x[15];
    const float na = x[1] * x[1] * x[1] * x[2] * 7.15f * (va - 50.0f);
    const float kd = x[8] * x[8] * x[8] * x[8] * 1.43f * (va + 72.0f);
    const float mk = .2613041227f * x[14] * (va + 80.0f);
    const float aleak = sc[3] * .001004288227f * (va + 50.0f);
    const float coup = .004017089073f * (vs - va);
    d[0] = (-na - kd - mk - aleak + coup) / .1148312178f;
    const float ev = ns_expf(vs / 11.4866f, et);
    const float ghk = fabsf(vs) < .001f ? 11.4866f * (x[13] - 15123.5f)
                                       : vs * (x[13] * ev - 15123.5f) / (ev - 1.0f);
    const float cal = sc[0] * (x[3] * x[4] * .1247796343f + x[5] * .07142141077f) * .00155982f *
                      ghk * .9f;
    const float cak = sc[1] * x[6] * x[7] * 1.663612026f * (vs + 72.0f) * .9f;
    const float amix = lp_sig(vs, 55.0f, -15.0f, et);
    const float ak = x[9] * x[9] * x[9] * .8008791415f * (vs + 72.0f) *
                     (amix * x[10] + (1.0f - amix) * x[11]) * .9f;
    const float ih = sc[2] * x[12] * .0126874385f * (vs + 20.0f) * .9f;
    const float sleak = sc[3] * .0006901584324f * (vs + 50.0f);
    d[15] = (-cal - cak - ak - ih - sleak - coup + (ina - 2.0f) * .01222505828f) / .0543909161f;
    float a = lp_ratio(-16.64f - .32f * va, ns_expf(-13.0f - va / 4.0f, et) - 1.0f, 1.28f);
    float b = lp_ratio(.28f * va + 7.0f, ns_expf(va / 5.0f + 5.0f, et) - 1.0f, 1.4f);
    d[1] = a * (1.0f - x[1]) - b * x[1];
    a = .128f * ns_expf(-2.66666666667f - va / 18.0f, et);
    b = 4.0f / (ns_expf(-5.0f - va / 5.0f, et) + 1.0f);
    d[2] = a * (1.0f - x[2]) - b * x[2];
    a = lp_ratio(-1.6f - .032f * va, ns_expf(-10.0f - va / 5.0f, et) - 1.0f, .16f);
    b = .5f * ns_expf(-1.375f - va / 40.0f, et);
    d[8] = a * (1.0f - x[8]) - b * x[8];
    d[3] = (lp_sig(vs, 15.0f, -9.8f, et) - x[3]) * .045f;
    d[4] = (lp_sig(vs, -40.0f, 3.2f, et) - x[4]) * .02f * lp_sig(vs, -15.0f, -10.0f, et);
    d[5] = (lp_sig(vs, 29.1336f, -4.43083f, et) - x[5]) * .0608644f;
    const float ca = x[13];
    const float tgt = lp_sig(vs, -.6f * ca, -23.0f, et) * lp_sig(vs, -16.0f - .6f * ca, -5.0f, et) *
                      ca / (2.5f + ca);
    d[6] = (tgt - x[6]) * .6f;
    d[7] = (.7f / (.6f + ca) - x[7]) * .035f;
    d[9] = (lp_sig(vs, -12.0f, -26.0f, et) - x[9]) * .14f;
    const float ainf = lp_sig(vs, -62.0f, 6.0f, et);
    d[10] = (ainf - x[10]) * .05f;
    d[11] = (ainf - x[11]) * .0036f * lp_sig(vs, -40.0f, -12.0f, et);
    d[12] = (lp_sig(vs, -70.0f, 8.0f, et) - x[12]) * .00033f / lp_sig(vs, -110.0f, -21.6f, et);
    d[13] = -.05049639866f * cal - .01729930764f * (ca - .02f);
    d[14] = (lp_sig(va, -26.9916223f, -5.957441281f, et) - x[14]) * .0001386796618f *
            lp_sig(va, -60.57875745f, -13.25674673f, et);
}

// lp_cell.cuh:98-113
__device__ inline void lp_advance(float *x, float ina, const float *sc, const uint64_t *et) {
    const float dt = .05f;
    for (int step = 0; step < 20; ++step) {
        float k1[16], k2[16], k3[16], k4[16], t[16];
        lp_rhs(x, ina, sc, k1, et);
        for (int j = 0; j < 16; ++j) t[j] = x[j] + .5f * dt * k1[j];
        lp_rhs(t, ina, sc, k2, et);
        for (int j = 0; j < 16; ++j) t[j] = x[j] + .5f * dt * k2[j];
        lp_rhs(t, ina, sc, k3, et);
        for (int j = 0; j < 16; ++j) t[j] = x[j] + dt * k3[j];
        lp_rhs(t, ina, sc, k4, et);
        for (int j = 0; j < 16; ++j) x[j] += dt * (k1[j] + 2 * k2[j] + 2 * k3[j] + k4[j]) / 6.0f;
    }
}

__device__ inline float pz_gate(float v, float a, float b, const uint64_t *et) {   // prinz_cell.cuh:41-43
    return 1.0f / (1.0f + ns_expf((v + a) / b, et));
}
__device__ inline float pz_gate2(float v, float a, float b, float c, float d, const uint64_t *et) {
    return 1.0f / (ns_expf((v + a) / b, et) + ns_expf((v + c) / d, et));
}

// prinz_cell.cuh:44-108; c = {v, ca, m[7], h[4]}
__device__ inline void pz_substep(float *c, const float *gg, float ext, float clamp_ns,
                                  float clamp_rev, const uint64_t *et, const double *lt) {
    const float dt = .025f;
    float *m = c + 2, *h = c + 9;
    const float v = c[0], ca = fmaxf(c[1], 0.000001f);
    const float eca = (8.31451e3f * 283.0f / (2.0f * 96485.3415f)) * ns_logf(3000.0f / ca, lt);
    const float cond[8] = {gg[0] * m[0] * m[0] * m[0] * h[0], gg[1] * m[1] * m[1] * m[1] * h[1],
                           gg[2] * m[2] * m[2] * m[2] * h[2], gg[3] * m[3] * m[3] * m[3] * h[3],
                           gg[4] * m[4] * m[4] * m[4] * m[4], gg[5] * m[5] * m[5] * m[5] * m[5],
                           gg[6] * m[6], gg[7]};
    const float rev[8] = {50, eca, eca, -80, -80, -80, -20, -50};
    float gsum = clamp_ns;
    float wsum = ext + clamp_ns * clamp_rev;
    for (int i = 0; i < 8; ++i) { gsum += cond[i]; wsum += cond[i] * rev[i]; }
    if (fabsf(gsum) < .000001f)
        c[0] = v + dt * wsum / 628.3f;
    else {
        const float target = wsum / gsum;
        c[0] = target + (v - target) * ns_expf(-dt * gsum / 628.3f, et);
    }
    const float ica = (cond[1] + cond[2]) * (v - eca) / 1000000.0f;
    const float ca_t = .05f - 14961.0f * ica;
    c[1] = fmaxf(.000001f, ca_t + (ca - ca_t) * ns_expf(-dt / 200.0f, et));
    const float mi[7] = {pz_gate(v, 25.5f, -5.29f, et), pz_gate(v, 27.1f, -7.2f, et),
                         pz_gate(v, 33, -8.1f, et), pz_gate(v, 27.2f, -8.7f, et),
                         (ca / (ca + 3)) * pz_gate(v, 28.3f, -12.6f, et),
                         pz_gate(v, 12.3f, -11.8f, et), pz_gate(v, 75, 5.5f, et)};
    const float mt[7] = {2.64f - 2.52f * pz_gate(v, 120, -25, et),
                         43.4f - 42.6f * pz_gate(v, 68.1f, -20.5f, et),
                         2.8f + 14 * pz_gate2(v, 27, 10, 70, -13, et),
                         23.2f - 20.8f * pz_gate(v, 32.9f, -15.2f, et),
                         180.6f - 150.2f * pz_gate(v, 46, -22.7f, et),
                         14.4f - 12.8f * pz_gate(v, 28.3f, -19.2f, et),
                         2.0f / (ns_expf(-14.59f - .086f * v, et) + ns_expf(-1.87f + .0701f * v, et))};
    for (int i = 0; i < 7; ++i) m[i] = mi[i] + (m[i] - mi[i]) * ns_expf(-dt / mt[i], et);
    const float hi[4] = {pz_gate(v, 48.9f, 5.18f, et), pz_gate(v, 32.1f, 5.5f, et),
                         pz_gate(v, 60, 6.2f, et), pz_gate(v, 56.9f, 4.9f, et)};
    const float ht[4] = {(1.34f / (1 + ns_expf((v + 62.9f) / -10, et))) *
                             (1.5f + 1 / (1 + ns_expf((v + 34.9f) / 3.6f, et))),
                         210 - 179.6f * pz_gate(v, 55, -16.9f, et),
                         120 + 300 * pz_gate2(v, 55, 9, 65, -16, et),
                         77.2f - 58.4f * pz_gate(v, 38.9f, -26.5f, et)};
    for (int i = 0; i < 4; ++i) h[i] = hi[i] + (h[i] - hi[i]) * ns_expf(-dt / ht[i], et);
}

// ---------------------------------------------------------------------------
// Synapses
// ---------------------------------------------------------------------------
__device__ inline float rng_uniform24(uint32_t &x) {    // physiology.cuh:380-391
    x ^= x << 13; x ^= x >> 17; x ^= x << 5;
    return (x >> 8) * (1.0f / 16777216.0f);
}

// synapse_release, physiology.cuh:425-454.  STP state in global memory,
// touched only on arrival.  `weight` is the current synaptic weight.
// synapse_release (physiology.cuh:425-454) on one synapse's release state,
// wherever it lives (pool arrays or a workgroup's LDS copy).  sdep is read only
// when the slow pool is on, as the reference does.
__device__ inline float stp_release(uint32_t flags, float weight, int64_t tick, float U,
                                    float rec, float fac, float srec, const float *sdep,
                                    float &util, float &slow, int64_t &last, uint64_t &ready,
                                    uint32_t &rng, const uint64_t *etab) {
    float p_recover = 0;
    if (last > -1000000) {
        float dt = (tick - last) * 1.0f;
        p_recover = 1.0f - ns_expf(-dt / rec, etab);
        if (fac > 0) util *= ns_expf(-dt / fac, etab);
        if (srec > 0) slow = 1.0f - (1.0f - slow) * ns_expf(-dt / srec, etab);
    }
    util = fac > 0 ? util + U * (1.0f - util) : U;
    const int n = (flags >> SF_SITES_SHIFT) & 127;
    int out = 0;
    for (int k = 0; k < n; ++k) {
        const uint64_t bit = uint64_t(1) << k;
        if (!(ready & bit) && rng_uniform24(rng) < p_recover) ready |= bit;
        if ((ready & bit) && rng_uniform24(rng) < util) {
            ready &= ~bit;
            ++out;
        }
    }
    float y = weight * out / n * slow;
    if (srec > 0) slow *= 1.0f - *sdep;
    last = tick;
    return y;
}

__device__ inline float release_at(const Pool &p, int64_t q, uint32_t flags, float weight,
                                   int64_t tick, const uint64_t *etab) {
    const int64_t S = p.scap;
    float util = p.ss[SS_UTIL * S + q], slow = p.ss[SS_SLOW * S + q];
    int64_t last = p.last[q];
    uint64_t ready = p.ready[q];
    uint32_t rng = p.srng[q];
    const float y = stp_release(flags, weight, tick, p.sg[SG_U * S + q], p.sg[SG_REC * S + q],
                                p.sg[SG_FAC * S + q], p.sg[SG_SREC * S + q], &p.sg[SG_SDEP * S + q],
                                util, slow, last, ready, rng, etab);
    p.ss[SS_UTIL * S + q] = util;
    p.ss[SS_SLOW * S + q] = slow;
    p.last[q] = last;
    p.ready[q] = ready;
    p.srng[q] = rng;
    return y;
}

// Plasticity registers: x/y hold the profile's two traces
//   profiles 1,2: pre_trace, post_trace   profile 3: eligible_ltp, eligible_ltd
//   profile 4: da_age, (unused)           profile 5: structural_age, spine_volume
struct PlastRegs {
    float weight, x, y, pre_age, post_age;
};

// State fields behind a profile's x and y registers (any profile value maps
// to a valid field).
__device__ inline int plast_x_field(int profile) {
    return profile == 3 ? SS_ELTP : profile == 4 ? SS_DA_AGE : profile == 5 ? SS_SAGE : SS_PTRACE;
}
__device__ inline int plast_y_field(int profile) {
    return profile == 3 ? SS_ELTD : profile == 5 ? SS_SPINE : SS_QTRACE;
}

__device__ inline void load_plast(const Pool &p, int64_t q, int profile, PlastRegs &r) {
    const int64_t S = p.scap;
    r.weight = p.ss[SS_WEIGHT * S + q];
    r.pre_age = p.ss[SS_PRE_AGE * S + q];
    r.post_age = p.ss[SS_POST_AGE * S + q];
    r.x = p.ss[plast_x_field(profile) * S + q];
    r.y = p.ss[plast_y_field(profile) * S + q];
}

__device__ inline void store_plast(const Pool &p, int64_t q, int profile, const PlastRegs &r) {
    const int64_t S = p.scap;
    p.ss[SS_WEIGHT * S + q] = r.weight;
    p.ss[SS_PRE_AGE * S + q] = r.pre_age;
    p.ss[SS_POST_AGE * S + q] = r.post_age;
    p.ss[plast_x_field(profile) * S + q] = r.x;
    if (profile != 4) p.ss[plast_y_field(profile) * S + q] = r.y;
}

__device__ inline float da_envelope(float age, const uint64_t *et) {   // physiology.cuh:408-412
    if (age < 0 || age > 20000.0f) return 0;
    float x = age / 2000.0f;
    return x * x * x * x * ns_expf(4.0f * (1.0f - x), et);
}
__device__ inline float spine_window(float age, const uint64_t *et) {  // physiology.cuh:416-420
    if (age < 0 || age > 5000.0f) return 0;
    float x = age / 1000.0f;
    return x * x * x * ns_expf(3.0f * (1.0f - x), et);
}

// synapse_plasticity_step, physiology.cuh:459-555 (all receptors available);
// called only for valid excitatory pairs with plasticity != 0.
__device__ inline void plasticity_step(PlastRegs &r, int profile, unsigned modrec, float w0,
                                       bool arrived, bool post_spike, float depol, unsigned mod,
                                       const uint64_t *et) {
    const float w_max = 2.0f * w0;
    if (profile == 5) {
        if (r.x >= 0) {
            r.x += 1.0f;
            if (r.x > 5000.0f) r.x = -1;
        }
        r.pre_age += 1.0f;
        if (arrived) r.pre_age = 0;
        if (post_spike && r.pre_age <= 15.0f && r.x < 0) r.x = 0;
        if ((mod & kModDA) && (modrec & kModDA))
            r.y = clamp_range(r.y + 0.025f * spine_window(r.x, et), 1.0f, 2.0f);
        return;
    }
    if (profile == 4) {
        if (r.x >= 0) {
            r.x += 1.0f;
            if (r.x > 20000.0f) r.x = -1;
        }
        r.pre_age += 1.0f;
        if (arrived) r.pre_age = 0;
        if (post_spike && r.pre_age <= 15.0f && (r.x < 0 || r.x > 30.0f)) {
            r.x = 0;
            r.weight -= 0.005383f * w0;
        }
        if ((mod & kModDA) && (modrec & kModDA)) {
            r.weight += 0.00453f * da_envelope(r.x, et) * w0;
            r.x = -1;
        }
        r.weight = clamp_range(r.weight, 0.0f, w_max);
        return;
    }
    if (profile == 3) {
        r.x *= kDecayEligLtp;
        r.y *= kDecayEligLtd;
        r.pre_age += 1.0f;
        r.post_age += 1.0f;
        if (arrived) {
            if (r.post_age <= 15.0f) r.y += 1.0f;
            r.pre_age = 0;
        }
        if (post_spike) {
            if (r.pre_age <= 15.0f) r.x += 1.0f;
            r.post_age = 0;
        }
        if ((mod & kModNE) && (modrec & kModNE)) {
            r.weight += 0.0048f * r.x * w0;
            r.x = 0;
        }
        if ((mod & kMod5HT) && (modrec & kMod5HT)) {
            r.weight -= 0.0075f * r.y * w0;
            r.y = 0;
        }
        r.weight = clamp_range(r.weight, 0.0f, w_max);
        return;
    }
    r.x *= kDecayLtp;
    r.y *= kDecayLtd;
    r.pre_age += 1.0f;
    r.post_age += 1.0f;
    if (arrived) {
        if (r.post_age <= (profile == 2 ? 50.0f : 20.0f)) r.weight *= 1.0f - 0.004f * r.y;
        r.x += 1.0f;
        r.pre_age = 0;
    }
    if (post_spike) {
        if (r.pre_age <= 20.0f && (profile == 1 || depol >= kDepolGateMv))
            r.weight += 0.016f * r.x * (2.0f * w0 - r.weight);
        r.y += 1.0f;
        r.post_age = 0;
    }
    r.weight = clamp_range(r.weight, 0.0f, w_max);
}

}  // namespace nsa
