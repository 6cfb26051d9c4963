// Tick kernels for gfx950 (MI355X).
//
// Two execution forms over the same structure-of-arrays pools, bit-identical
// to each other and to the reference's CPU tick order (gym_adapter/core.cuh:
// 30-75, src/main.cu:659-801 and 2702-2728):
//
//  * plain per-tick form — one launch each of delivery, neuron and
//    plasticity per tick over the global pools (the reference's global-slot
//    kernels, with the float atomics replaced by a per-neuron sequential sum
//    over its incoming CSR run);
//  * resident form — one launch per call: each workgroup owns a pack of
//    whole brains and runs every tick with neuron state and plasticity
//    traces in registers, the spike rings, releases and routing in LDS, and
//    two-phase delivery: arrivals are compacted into an LDS queue with
//    wavefront ballots, then short-term plasticity runs densely over the
//    queue.  Brains are independent, so no cross-workgroup synchronisation
//    is needed inside a call.
#include <atomic>

#include "ns_kernels.h"

namespace nsa {

__device__ const uint64_t g_exp_tab[32] = NS_EXP2_TABLE;
__device__ const double g_log_tab[32] = NS_LOGF_TABLE;

__device__ inline bool need_ok(const Pool &p, int32_t b, uint8_t need) {
    return (p.brun[b] & need) == need;
}

// ---------------------------------------------------------------------------
// Reset: initial_neuron_state / initial_synapse_state (physiology.cuh:179-202,
// 393-404) seeded as gym_adapter/gpu.cu:18-23.  need = 1: brains with brun&1.
// ---------------------------------------------------------------------------
// Initial state of neuron i of a brain reset with `seed` (initial_neuron_state,
// physiology.cuh:179-202, and the cell initial values of lp_cell.cuh /
// prinz_cell.cuh); `orig` is its original genome index.
__device__ inline void reset_neuron(const Pool &p, int64_t i, uint32_t seed, uint32_t orig) {
    const int64_t N = p.ncap;
    const uint32_t type = p.nflags[i] & NF_TYPE;
    for (int k = 0; k < NS_FIELDS; ++k) p.ns[k * N + i] = 0.0f;
    float cell[kCellFields];
    const float v = initial_cell(type, p.ng[NG_REST * N + i], p.ng[NG_THR * N + i], cell);
    for (int k = 0; k < kCellFields; ++k) p.cell[k * N + i] = cell[k];
    p.ns[NS_V * N + i] = v;
    p.rrng[i] = initial_rng(seed + orig * 7919u);
    p.total[i] = 0;
    p.spike[i] = 0;
    p.hist[i] = 0;
    p.modbits[i] = 0;
    for (int k = 0; k < 6; ++k) p.stamps[k * N + i] = 0;
}

// initial_synapse_state (physiology.cuh:393-404) of synapse q.
__device__ inline void reset_synapse(const Pool &p, int64_t q, uint32_t seed, uint32_t orig) {
    const int64_t S = p.scap;
    float init[SS_FIELDS] = {};
    init[SS_WEIGHT] = p.sg[SG_W0 * S + q];
    init[SS_DA_AGE] = -1; init[SS_SPINE] = 1; init[SS_SAGE] = -1;
    init[SS_PRE_AGE] = 1000000; init[SS_POST_AGE] = 1000000; init[SS_SLOW] = 1;
    for (int k = 0; k < SS_FIELDS; ++k) p.ss[k * S + q] = init[k];
    p.ready[q] = initial_ready((p.sflags[q] >> SF_SITES_SHIFT) & 127);
    p.last[q] = -1000000;
    p.srng[q] = initial_rng(seed + orig * 104729u);
    p.arrived[q] = 0;
    p.released[q] = 0;
}

// Reset of every brain whose run flag is set (seeds per pool brain).
__global__ void k_reset_neurons(Pool p, int64_t n, const uint32_t *bseed, const uint32_t *orig) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t b = p.nbrain[i];
    if (!(p.brun[b] & 1)) return;
    reset_neuron(p, i, bseed[b], orig[i]);
}

__global__ void k_reset_synapses(Pool p, int64_t n, const uint32_t *bseed, const uint32_t *orig) {
    int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= n) return;
    int32_t b = p.sbrain[q];
    if (!(p.brun[b] & 1)) return;
    reset_synapse(p, q, bseed[b], orig[q]);
}

// Reset of the listed slots only: list[0..count) slot indices, list[count..2
// count) their seeds.  gridDim.y workgroups per slot share its whole capacity
// (a slot of a few thousand neurons is not left to one workgroup), and the
// first sets its tick counter back to zero.
__global__ void k_reset_slots(Pool p, const int32_t *list, int count, const uint32_t *norig,
                              const uint32_t *sorig) {
    const int b = list[blockIdx.x];
    const uint32_t seed = uint32_t(list[count + blockIdx.x]);
    const int64_t n0 = p.bn0[b], n1 = n0 + p.bnc[b];
    const int64_t s0 = p.bs0[b], s1 = s0 + p.bsc[b];
    const int64_t first = int64_t(blockIdx.y) * blockDim.x + threadIdx.x;
    const int64_t stride = int64_t(gridDim.y) * blockDim.x;
    for (int64_t i = n0 + first; i < n1; i += stride) reset_neuron(p, i, seed, norig[i]);
    for (int64_t q = s0 + first; q < s1; q += stride) reset_synapse(p, q, seed, sorig[q]);
    if (blockIdx.y == 0 && threadIdx.x == 0) p.btick[b] = 0;
}

// Slot placement: each entry copies `bytes` from the staged image (device
// copy of the pinned staging buffer) to one pool array range.  One workgroup
// per entry; 4-byte words when both ends and the length allow it.
__global__ void k_scatter(const ScatterEntry *entries, int n, const unsigned char *base) {
    const int e = blockIdx.x;
    if (e >= n) return;
    const ScatterEntry en = entries[e];
    unsigned char *dst = reinterpret_cast<unsigned char *>(en.dst);
    const unsigned char *src = base + en.src;
    if (((en.dst | en.src | en.bytes) & 3u) == 0) {
        uint32_t *d = reinterpret_cast<uint32_t *>(dst);
        const uint32_t *s4 = reinterpret_cast<const uint32_t *>(src);
        const uint64_t words = en.bytes / 4, step = blockDim.x;
        uint64_t i = threadIdx.x;
        // four independent loads in flight per thread before their stores
        for (; i + 3 * step < words; i += 4 * step) {
            const uint32_t a = s4[i], b = s4[i + step], c = s4[i + 2 * step], e4 = s4[i + 3 * step];
            d[i] = a;
            d[i + step] = b;
            d[i + 2 * step] = c;
            d[i + 3 * step] = e4;
        }
        for (; i < words; i += step) d[i] = s4[i];
    } else {
        for (uint64_t i = threadIdx.x; i < en.bytes; i += blockDim.x) dst[i] = src[i];
    }
}

// Stable counting sort of one pack's neuron slots by slot_class, the cells
// that carry a settled mark (kStillCarry) after the others, as k_resident_lds
// orders them after a launch: per chunk of 1024 slots, each key's ballot gives
// a slot its rank within its wavefront and the wavefront's count; earlier
// wavefronts and chunks add their counts.
constexpr int kOrderThreads = 1024;
__device__ inline int order_key(const uint32_t *nflags, const uint8_t *modbits, int64_t i) {
    return slot_class(nflags[i]) + ((modbits[i] & kStillCarry) ? kSlotClasses : 0);
}
__global__ __launch_bounds__(kOrderThreads) void k_pack_order(const uint32_t *nflags,
                                                              const uint8_t *modbits,
                                                              const Pack *packs,
                                                              const int32_t *list,
                                                              uint16_t *perm) {
    constexpr int W = kOrderThreads / 64, K = 2 * kSlotClasses;
    __shared__ int total[K], base[K], wcnt[W][K];
    const Pack pk = packs[list[blockIdx.x]];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid < K) total[tid] = 0;
    __syncthreads();
    for (int j = tid; j < pk.nn; j += kOrderThreads) atomicAdd(&total[order_key(nflags, modbits, pk.n0 + j)], 1);
    __syncthreads();
    if (tid == 0) {
        int s = 0;
        for (int c = 0; c < K; ++c) { base[c] = s; s += total[c]; }
    }
    __syncthreads();
    for (int j0 = 0; j0 < pk.nn; j0 += kOrderThreads) {   // uniform trip count
        const int j = j0 + tid;
        const int c = j < pk.nn ? order_key(nflags, modbits, pk.n0 + j) : -1;
        int rank = 0;
        for (int k = 0; k < K; ++k) {
            const uint64_t m = __ballot(c == k);
            if (lane == 0) wcnt[wave][k] = __popcll(m);
            if (c == k) rank = __popcll(m & ((uint64_t(1) << lane) - 1));
        }
        __syncthreads();
        if (c >= 0) {
            int pos = base[c] + rank;
            for (int w = 0; w < wave; ++w) pos += wcnt[w][c];
            perm[pk.n0 + pos] = uint16_t(j);
        }
        __syncthreads();
        if (tid < K) {
            int s = 0;
            for (int w = 0; w < W; ++w) s += wcnt[w][tid];
            base[tid] += s;
        }
        __syncthreads();
    }
}

__global__ void k_set_ticks(Pool p, int nb, int64_t add, int64_t set, uint8_t need) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb || !need_ok(p, b, need)) return;
    p.btick[b] = set >= 0 ? set : p.btick[b] + add;
}

__global__ void k_fill_f32(float *x, int64_t n, float v) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}
__global__ void k_fill_i32(int32_t *x, int64_t n, int32_t v) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}
// The clears before a step in one launch: outputs to +0, statuses to 0.
__global__ void k_clear_step(float *out, int64_t nout, int32_t *status, int64_t nstatus) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < nout) out[i] = 0.0f;
    if (i < nstatus) status[i] = 0;
}

// Motor spike counts of a world step added to the caller's totals
// (main.cu:743 adds each spike as it happens; the sum is the same).
__global__ void k_add_counts(int32_t *total, const float *counts, int64_t n) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) total[i] += int32_t(counts[i]);
}

// ---------------------------------------------------------------------------
// Plain per-tick form
// ---------------------------------------------------------------------------

// Delivery (src/main.cu:659-708; gym_adapter/core.cuh:37-47): arrival test on
// the spike shift register, short-term release on arrival.  Returns arrival.
__device__ inline unsigned deliver_one(const Pool &p, int64_t q, int32_t b, int t,
                                       const uint64_t *etab = g_exp_tab) {
    uint32_t flags = p.sflags[q];
    uint32_t route = flags >> SF_ROUTE_SHIFT;
    float rel = 0;
    unsigned arr = 0;
    if (route & R_VALID) {
        int64_t tick = p.btick[b] + t;
        int d = flags & SF_DELAY;
        if (tick >= d) arr = (p.hist[p.spre[q]] >> (d - 1)) & 1u;
        if (arr) {
            float w = p.ss[SS_WEIGHT * p.scap + q];
            rel = (route & R_STP) ? release_at(p, q, flags, w, tick, etab) : w;
        }
    }
    p.arrived[q] = uint8_t(arr);
    p.released[q] = rel;
    return arr;
}

__global__ void k_deliver(Pool p, int64_t s_lo, int64_t s_hi, int t, uint8_t need) {
    int64_t q = s_lo + int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    unsigned arr = 0;
    if (q < s_hi) {
        int32_t b = p.sbrain[q];
        if (need_ok(p, b, need)) arr = deliver_one(p, q, b, t);
    }
    uint64_t m = __ballot(arr);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(&p.events[1], (unsigned long long)__popcll(m));
}

// Sequential per-neuron reduction of this tick's releases in ascending edge
// order (gym_adapter/gpu.cu:40-47), with the world's modulation stamps
// (src/main.cu:669-673).  Returns the modulator bits.
__device__ inline unsigned gather_inputs(const Pool &p, int64_t i, int64_t tick, float &e0,
                                         float &e1, float &e2, float &e3, float &inh,
                                         const uint64_t *etab) {
    unsigned mod = 0;
    bool have_scale = false;
    float scale = 0;
    const int32_t q1 = p.nin[i + 1];
    for (int32_t q = p.nin[i]; q < q1; ++q) {
        float r = p.released[q];
        if (r == 0.0f) continue;
        uint32_t route = p.sflags[q] >> SF_ROUTE_SHIFT;
        if (route & R_MOD) {
            if (r > 0) {
                unsigned m = (route >> R_MOD_SHIFT) & 3u;
                mod |= 1u << (m - 1);
                uint64_t &st = p.stamps[((m - 1) * 2 + (tick & 1)) * p.ncap + i];
                if (st < uint64_t(tick + 1)) st = uint64_t(tick + 1);
            }
        } else if (route & R_INH) {
            inh += r;
        } else {
            if (!have_scale) {
                scale = ns_expf(p.ns[NS_LAS * p.ncap + i], etab);
                have_scale = true;
            }
            float v = r * scale;
            switch (route & R_BRANCH) {
            case 0: e0 += v; break;
            case 1: e1 += v; break;
            case 2: e2 += v; break;
            default: e3 += v; break;
            }
        }
    }
    return mod;
}

__device__ inline float sensory_current(const Pool &p, int64_t i, NeuronRegs &s,
                                        const NeuronConst &g, const float *input, float gain) {
    int32_t idx = p.ninput[i];
    if (idx < 0) return 0.0f;
    float level = input[idx];
    return (g.flags & NF_RECEPTOR) ? receptor_drive(s, g, level) : gain * level;
}

// neuron_step (src/main.cu:710-746) / the neuron loop of core.cuh:49-58, for
// the register types (0, 1, 4, 5, 6) and inactive cells of any type.
// `arrived` = false: no incoming synapse of neuron i arrived this tick, so
// every release it would sum is +0 and its sums are +0 without reading them
// (sums start at +0; +0 + +0 = +0) and no modulator bit is set.
//
// still (when not null) receives whether this tick changed no state of an
// AdEx cell without sensory input (neuron_tick), so that its next tick with
// no arrival is the identity.
__device__ inline void neuron_one(const Pool &p, int64_t i, int32_t b, int t, const float *input,
                                  float gain, bool arrived = true, bool *still = nullptr,
                                  const uint64_t *etab = g_exp_tab) {
    uint32_t flags = p.nflags[i];
    if (!(flags & NF_ACTIVE)) {
        uint32_t h = p.hist[i];
        if (flags & NF_WORLD) {
            p.hist[i] = world_rotate(h, flags);
        } else {
            p.hist[i] = h << 1;
            p.spike[i] = 0;
        }
        return;
    }
    int64_t tick = p.btick[b] + t;
    // the state and gene loads go out before the input gather's chain
    NeuronRegs s;
    NeuronConst g;
    load_neuron(p, i, s, g, etab);
    float e0 = 0, e1 = 0, e2 = 0, e3 = 0, inh = 0;
    const unsigned mod = arrived ? gather_inputs(p, i, tick, e0, e1, e2, e3, inh, etab) : 0u;
    float current = sensory_current(p, i, s, g, input, gain);
    int spike = neuron_tick(s, g, current, e0, e1, e2, e3, inh, p.cell, p.ncap, i, etab, still);
    if (still) *still = *still && p.ninput[i] < 0 && (flags & NF_TYPE) != 6;
    store_neuron(p, i, s);
    p.spike[i] = uint8_t(spike);
    p.hist[i] = (p.hist[i] << 1) | uint32_t(spike);
    p.modbits[i] = uint8_t(mod);
    if (!isfinite(s.v)) p.bstatus[b] = -1;
}

__global__ void k_neuron(Pool p, int64_t n_lo, int64_t n_hi, int t, uint8_t need,
                         const float *input, float gain) {
    int64_t i = n_lo + int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n_hi) return;
    int32_t b = p.nbrain[i];
    if (!need_ok(p, b, need)) return;
    uint32_t type = p.nflags[i] & NF_TYPE;
    if ((type == 2 || type == 3) && (p.nflags[i] & NF_ACTIVE)) return;
    neuron_one(p, i, b, t, input, gain);
}

// Types 2 (LP, lp_cell.cuh) and 3 (Prinz, prinz_cell.cuh): physiology.cuh:313-347.
__global__ void k_neuron_slow(Pool p, const int32_t *list, int count, int t, uint8_t need,
                              const float *input, float gain) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    int64_t i = list[k];
    int32_t b = p.nbrain[i];
    if (!need_ok(p, b, need)) return;
    uint32_t flags = p.nflags[i];
    const int64_t N = p.ncap;
    if (!(flags & NF_ACTIVE)) return;   // inactive cells run in k_neuron
    int64_t tick = p.btick[b] + t;
    float e0 = 0, e1 = 0, e2 = 0, e3 = 0, inh_in = 0;
    unsigned mod = gather_inputs(p, i, tick, e0, e1, e2, e3, inh_in, g_exp_tab);
    NeuronRegs s;
    NeuronConst g;
    load_neuron(p, i, s, g, g_exp_tab);
    float current = sensory_current(p, i, s, g, input, gain);
    s.d0 += e0; s.d1 += e1; s.d2 += e2; s.d3 += e3;
    float sub = fmaxf(0.0f, fminf(s.v, -50.0f) - g.rest);
    s.depol += 0.1f * (sub - s.depol);
    s.inh += inh_in;
    float c[kCellFields];
    for (int j = 0; j < kCellFields; ++j) c[j] = p.cell[j * N + i];
    float e = 0;
    e += s.d0; e += s.d1; e += s.d2; e += s.d3;
    int spike = 0;
    if ((flags & NF_TYPE) == 2) {
        float before = c[0];
        float ext = current + e * (0.0f - s.v) + s.inh * (-75.0f - s.v) - 0.0f * (s.v - -68.0f);
        float sc[4];
        for (int j = 0; j < 4; ++j) sc[j] = p.nx[(NX_LP + j) * N + i];
        lp_advance(c, ext / 1000.0f, sc, g_exp_tab);
        s.v = c[15] - 9.342518882f;
        spike = before < 0 && c[0] >= 0;
    } else {
        float gg[8];
        for (int j = 0; j < 8; ++j) gg[j] = p.nx[(NX_PRINZ + j) * N + i];
        float amb = p.nx[NX_AMBIENT * N + i];
        float ext = current + e * (0.0f - s.v) + s.inh * (-75.0f - s.v);
        bool fired = false;
        for (int j = 0; j < 40; ++j) {
            const float before = c[0];
            pz_substep(c, gg, ext, amb + 0.0f, -68.0f, g_exp_tab, g_log_tab);
            fired |= before < 0 && c[0] >= 0;
        }
        spike = fired;
        s.v = c[0];
    }
    s.total += spike;
    s.d0 *= kDecayExcMs; s.d1 *= kDecayExcMs; s.d2 *= kDecayExcMs; s.d3 *= kDecayExcMs;
    s.inh *= kDecayInhMs;
    for (int j = 0; j < kCellFields; ++j) p.cell[j * N + i] = c[j];
    store_neuron(p, i, s);
    p.spike[i] = uint8_t(spike);
    p.hist[i] = (p.hist[i] << 1) | uint32_t(spike);
    p.modbits[i] = uint8_t(mod);
    if (!isfinite(s.v)) p.bstatus[b] = -1;
}

// Plasticity (src/main.cu:748-782; core.cuh:60-66) of synapse q of brain b.
__device__ inline void plasticity_one(const Pool &p, int64_t q, int32_t b,
                                      const uint64_t *etab = g_exp_tab) {
    if (p.bfrozen[b]) return;
    uint32_t flags = p.sflags[q];
    uint32_t route = flags >> SF_ROUTE_SHIFT;
    int profile = (flags >> SF_PLAST_SHIFT) & 7;
    if (!(route & R_VALID) || !(route & R_PAIR) || !profile) return;
    int32_t post = p.spost[q];
    PlastRegs r;
    load_plast(p, q, profile, r);
    plasticity_step(r, profile, (flags >> SF_MODREC_SHIFT) & 7, p.sg[SG_W0 * p.scap + q],
                    p.arrived[q] != 0, p.spike[post] != 0, p.ns[NS_DEPOL * p.ncap + post],
                    p.modbits[post], etab);
    store_plast(p, q, profile, r);
}

// Plasticity plus motor counting (core.cuh:68, main.cu:742-743) for threads
// below `slots`.
__global__ void k_plasticity(Pool p, int64_t s_lo, int64_t s_hi, int t, uint8_t need,
                             int slots, float *out) {
    int64_t gid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid < slots) {
        int32_t b = int32_t(gid / p.outputs);
        const int32_t mn = p.motor[gid];
        if (need_ok(p, b, need) && mn >= 0) out[p.bout0[b] + gid % p.outputs] += float(p.spike[mn]);
    }
    int64_t q = s_lo + gid;
    if (q >= s_hi) return;
    int32_t b = p.sbrain[q];
    if (!need_ok(p, b, need)) return;
    plasticity_one(p, q, b);
}

// Homeostasis at the end of a batched call (core.cuh:70-72): elapsed = ticks,
// rate = total_spikes * 1000 / max(ticks, tick after the call).
__global__ void k_homeostasis_gym(Pool p, int64_t n_lo, int64_t n_hi, int ticks, uint8_t need) {
    int64_t i = n_lo + int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n_hi) return;
    int32_t b = p.nbrain[i];
    if (!need_ok(p, b, need)) return;
    uint32_t flags = p.nflags[i];
    if (!(flags & NF_ACTIVE) || (flags & NF_TYPE) != 0) return;
    int64_t tick = p.btick[b] + ticks;
    float hz = p.total[i] * 1000.0f / float(ticks > tick ? int64_t(ticks) : tick);
    float &las = p.ns[NS_LAS * p.ncap + i];
    las = homeostasis_las(las, float(ticks), hz);
}

// homeostasis_step (src/main.cu:1287-1300): once per world step, elapsed 16
// ms, rate = total_spikes * 1000 / max(16, age * 16).
__global__ void k_homeostasis_world(Pool p, int64_t n, const int32_t *age) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t b = p.nbrain[i];
    if (!(p.brun[b] & 1) || p.bfrozen[b]) return;
    uint32_t flags = p.nflags[i];
    if (!(flags & NF_ACTIVE) || (flags & NF_TYPE) != 0) return;
    float elapsed = fmaxf(16.0f, age[b] * 16.0f);
    float hz = p.total[i] * 1000.0f / elapsed;
    float &las = p.ns[NS_LAS * p.ncap + i];
    las = homeostasis_las(las, 16.0f, hz);
}

// ---------------------------------------------------------------------------
// Resident form
// ---------------------------------------------------------------------------

// k plasticity steps of a pair synapse for ticks with no arrival, no spike of
// its post neuron and no modulator at it: plasticity_step with those inputs
// (physiology.cuh:459-555) decays the two traces, advances the ages by one
// and bounds the weight, each variable on its own.  So the steps are replayed
// variable by variable: each trace by its k multiplications in order; an age
// that is a whole number below 2^24 - 128 by one addition of k (every
// intermediate sum is then a float exactly, so the k additions of 1 give the
// same value), any other age by k additions; the weight by one bound, which
// is idempotent.  Profiles 4 and 5 count their dopamine / structural age up
// to a reset value instead of decaying a trace.
// Most quiet steps a synapse defers before it replays them (7 bits).
constexpr uint32_t kQuietMax = 0x7f;

__device__ inline float add_steps(float age, uint32_t k) {
    if (age >= 0.0f && age <= 16777088.0f && age == __builtin_truncf(age)) return age + float(k);
    for (uint32_t i = 0; i < k; ++i) age += 1.0f;
    return age;
}

__device__ inline float count_steps(float x, uint32_t k, float reset_above) {
    for (uint32_t i = 0; i < k && x >= 0.0f; ++i) {
        x += 1.0f;
        if (x > reset_above) x = -1.0f;
    }
    return x;
}

__device__ inline void replay_quiet_steps(PlastRegs &r, int profile, float w0, uint32_t k) {
    if (!k) return;
    r.pre_age = add_steps(r.pre_age, k);
    if (profile == 5) {
        r.x = count_steps(r.x, k, 5000.0f);
        return;
    }
    if (profile == 4) {
        r.x = count_steps(r.x, k, 20000.0f);
    } else {
        const float dx = profile == 3 ? kDecayEligLtp : kDecayLtp;
        const float dy = profile == 3 ? kDecayEligLtd : kDecayLtd;
        for (uint32_t i = 0; i < k; ++i) {
            r.x *= dx;
            r.y *= dy;
        }
        r.post_age = add_steps(r.post_age, k);
    }
    r.weight = clamp_range(r.weight, 0.0f, 2.0f * w0);
}


// Plasticity of tick t-1 of big-brain synapse q with its quiet steps
// deferred (the rule of k_resident_lds phase A): a pair synapse whose step
// has no input (no arrival at t-1, no spike of its post neuron, no modulator
// at it) and no arrival at t counts the step in `pend` instead; the next step
// with an input, the 127th deferred step and the call's last step replay the
// deferred ones first.
__device__ inline void plasticity_deferred(const Pool &p, int64_t q, int32_t b, bool arr_now, bool last,
                                           uint8_t &pend, const uint64_t *etab = g_exp_tab) {
    if (p.bfrozen[b]) return;
    const uint32_t flags = p.sflags[q];
    const int32_t post = p.spost[q];           // loaded with the flags word (q is in the pool)
    const bool arrived = p.arrived[q] != 0;
    const uint32_t route = flags >> SF_ROUTE_SHIFT;
    const int profile = (flags >> SF_PLAST_SHIFT) & 7;
    if (!(route & R_VALID) || !(route & R_PAIR) || !profile) return;
    const bool spiked = p.spike[post] != 0;
    const unsigned pm = p.modbits[post], modrec = (flags >> SF_MODREC_SHIFT) & 7;
    if (!last && !arrived && !spiked && !(pm & modrec) && !arr_now && pend < kQuietMax) {
        ++pend;
        return;
    }
    PlastRegs r;
    load_plast(p, q, profile, r);
    const float w0 = p.sg[SG_W0 * p.scap + q];
    replay_quiet_steps(r, profile, w0, pend);
    plasticity_step(r, profile, modrec, w0, arrived, spiked, p.ns[NS_DEPOL * p.ncap + post], pm, etab);
    store_plast(p, q, profile, r);
    pend = 0;
}

// Whether synapse q of brain b arrives at tick t of the call.
__device__ inline bool arrives(const Pool &p, int64_t q, int32_t b, int t, uint32_t &flags) {
    flags = p.sflags[q];
    const int32_t pre = p.spre[q];   // loaded with the flags word (q is in the pool)
    if (!((flags >> SF_ROUTE_SHIFT) & R_VALID)) return false;
    const int d = flags & SF_DELAY;
    return p.btick[b] + t >= d && ((p.hist[pre] >> (d - 1)) & 1u);
}

// deliver_one for a synapse whose arrival at tick t is already known (the
// arrival depends only on the flags word and the pre neuron's spike register,
// which plasticity does not change); the weight is read after the synapse's
// plasticity step of tick t-1, as in deliver_one.
__device__ inline unsigned deliver_known(const Pool &p, int64_t q, int32_t b, int t, bool arr,
                                         uint32_t flags, const uint64_t *etab) {
    float rel = 0;
    if (arr) {
        const float w = p.ss[SS_WEIGHT * p.scap + q];
        rel = ((flags >> SF_ROUTE_SHIFT) & R_STP) ? release_at(p, q, flags, w, p.btick[b] + t, etab) : w;
    }
    p.arrived[q] = uint8_t(arr);
    p.released[q] = rel;
    return arr;
}

// A brain too large for a pack: one workgroup runs all its ticks, with the
// per-tick phases separated by workgroup barriers instead of kernel launches
// and its state kept in global memory (L2-resident for the call).
//
// `lds` (hit_cap bytes; may be null) holds, when the brain fits, a byte per
// neuron and then a byte per synapse.  The neuron byte: bit 0, a synapse onto
// it arrived this tick (set by the delivery pass; the neuron pass skips the
// input gather of neurons without one); bit 1, its last tick changed no state
// and it is an AdEx cell without sensory input (see neuron_tick): with no
// arrival its tick is the identity, and the neuron pass only shifts its spike
// register.  The synapse byte counts its deferred quiet plasticity steps.
template <int WG>
__device__ void big_brain_ticks(const Pool &p, int32_t b, int ticks, const float *input,
                                float gain, float *out, int homeostasis, uint8_t *lds, int64_t hit_cap,
                                const uint64_t *etab) {
    __shared__ unsigned big_events, big_skipped;
    if (!(p.brun[b] & 1)) return;
    const int tid = threadIdx.x;
    const int64_t n0 = p.bn0[b], n1 = n0 + p.bnc[b];
    const int64_t s0 = p.bs0[b], s1 = s0 + p.bsc[b];
    if (!lds || (n1 - n0) + (s1 - s0) > hit_cap) lds = nullptr;
    uint8_t *hit = lds, *pend = lds ? lds + (n1 - n0) : nullptr;
    if (lds)
        for (int64_t i = tid; i < (n1 - n0) + (s1 - s0); i += WG) lds[i] = 0;
    if (tid == 0) big_events = big_skipped = 0;
    __syncthreads();
    unsigned arrivals = 0, skipped = 0;
    // Per tick: plasticity of tick t-1 and delivery of tick t in one synapse
    // pass (a synapse's plasticity precedes its next delivery in its own
    // thread, the reference's order), then the neurons of tick t.
    for (int t = 0; t < ticks; ++t) {
        for (int64_t q = s0 + tid; q < s1; q += WG) {
            unsigned arr;
            if (lds) {
                uint32_t flags;
                const bool now = arrives(p, q, b, t, flags);
                if (t > 0) plasticity_deferred(p, q, b, now, false, pend[q - s0], etab);
                arr = deliver_known(p, q, b, t, now, flags, etab);
            } else {
                if (t > 0) plasticity_one(p, q, b, etab);
                arr = deliver_one(p, q, b, t, etab);
            }
            if (arr && hit) hit[p.spost[q] - n0] |= 1;
            arrivals += arr;
        }
        __syncthreads();
        for (int64_t i = n0 + tid; i < n1; i += WG) {
            if (!hit) {
                neuron_one(p, i, b, t, input, gain, true, nullptr, etab);
                continue;
            }
            const uint8_t h = hit[i - n0];
            if (h == 2) {
                p.hist[i] <<= 1;
                p.spike[i] = 0;
                p.modbits[i] = 0;
                ++skipped;
                continue;
            }
            bool still = false;
            neuron_one(p, i, b, t, input, gain, h & 1, &still, etab);
            hit[i - n0] = still ? 2 : 0;
        }
        __syncthreads();
        for (int m = tid; m < p.outputs; m += WG) {
            const int32_t mn = p.motor[int64_t(b) * p.outputs + m];
            if (mn >= 0) out[p.bout0[b] + m] += float(p.spike[mn]);
        }
    }
    if (ticks > 0)
        for (int64_t q = s0 + tid; q < s1; q += WG) {
            if (lds) plasticity_deferred(p, q, b, false, true, pend[q - s0], etab);
            else plasticity_one(p, q, b, etab);
        }
    __syncthreads();
    if (homeostasis) {
        const int64_t tick = p.btick[b] + ticks;
        for (int64_t i = n0 + tid; i < n1; i += WG) {
            const uint32_t flags = p.nflags[i];
            if (!(flags & NF_ACTIVE) || (flags & NF_TYPE) != 0) continue;
            const float hz = p.total[i] * 1000.0f / float(ticks > tick ? int64_t(ticks) : tick);
            p.ns[NS_LAS * p.ncap + i] = homeostasis_las(p.ns[NS_LAS * p.ncap + i], float(ticks), hz);
        }
    }
    if (arrivals) atomicAdd(&big_events, arrivals);
    if (skipped) atomicAdd(&big_skipped, skipped);
    __syncthreads();
    if (tid == 0) {
        p.btick[b] += ticks;
        if (big_events) atomicAdd(&p.events[0], (unsigned long long)big_events);
        if (big_skipped) atomicAdd(&p.events[2], (unsigned long long)big_skipped);
    }
}

template <int WG, int KN, int KS, int MINW>
__global__ __launch_bounds__(WG, MINW) void k_resident(Pool p, const Pack *packs, const int32_t *big,
                                                 int nbig, const uint16_t *perm, int ticks,
                                                 const float *input, float gain, float *out,
                                                 int homeostasis) {
    constexpr int NMAX = WG * KN, SMAX = WG * KS;
    __shared__ uint64_t etab[32];
    __shared__ uint32_t hist_l[NMAX];
    __shared__ float depol_l[NMAX];
    __shared__ uint8_t mod_l[NMAX];
    __shared__ uint16_t in_l[NMAX + 1];
    __shared__ float rel_l[SMAX];
    __shared__ uint8_t route_l[SMAX];
    __shared__ uint8_t sbr_l[SMAX];
    __shared__ uint16_t queue_l[SMAX];
    __shared__ int64_t btick_l[kPackBrains];
    __shared__ uint8_t brun_l[kPackBrains], bfrozen_l[kPackBrains];
    __shared__ uint16_t mot_l[kPackMotors];
    __shared__ uint16_t cnt_l[kPackMotors];
    __shared__ int qlen;
    __shared__ unsigned events_l;

    if (int(blockIdx.x) < nbig) {
        if (threadIdx.x < 32) etab[threadIdx.x] = g_exp_tab[threadIdx.x];   // ordered by the first barrier
        big_brain_ticks<WG>(p, big[blockIdx.x], ticks, input, gain, out, homeostasis, nullptr, 0, etab);
        return;
    }
    const Pack pk = packs[blockIdx.x - nbig];
    const int tid = threadIdx.x;
    {   // a pack whose brains all sit out this call (free slots, masks) does nothing
        int any = 0;
        for (int b = tid; b < pk.nb; b += WG) any |= p.brun[pk.b0 + b] & 1;
        if (!__syncthreads_or(any)) return;
    }
    const int64_t N = p.ncap, S = p.scap;
    const int64_t n0 = pk.n0, s0 = pk.s0;

    if (tid < 32) etab[tid] = g_exp_tab[tid];
    for (int b = tid; b < pk.nb; b += WG) {
        btick_l[b] = p.btick[pk.b0 + b];
        brun_l[b] = p.brun[pk.b0 + b] & 1;
        bfrozen_l[b] = p.bfrozen[pk.b0 + b];
    }
    for (int n = tid; n < pk.nn; n += WG) {
        hist_l[n] = p.hist[n0 + n];
        in_l[n] = uint16_t(p.nin[n0 + n] - s0);
        mod_l[n] = 0;
        depol_l[n] = 0;
    }
    if (tid == 0) {
        in_l[pk.nn] = uint16_t(pk.ns);
        qlen = 0;
        events_l = 0;
    }
    for (int q = tid; q < pk.ns; q += WG) {
        route_l[q] = uint8_t(p.sflags[s0 + q] >> SF_ROUTE_SHIFT);
        sbr_l[q] = uint8_t(p.sbrain[s0 + q] - pk.b0);
    }
    for (int m = tid; m < pk.nm; m += WG) {
        const int32_t mn = p.motor[int64_t(pk.b0) * p.outputs + m];
        mot_l[m] = mn >= 0 ? uint16_t(mn - n0) : kNoMotor;
        cnt_l[m] = 0;
    }
    __syncthreads();

    // Neuron registers.  Slot (k, tid) holds the pack's neuron perm[k*WG+tid]:
    // the pack's neurons sorted by cell type, so wavefronts run one membrane
    // model at a time.
    NeuronRegs nr[KN];
    NeuronConst nc[KN];
    float nscale[KN], ncur[KN];
    int nbr[KN], nid[KN];
    bool nlive[KN];
#pragma unroll
    for (int k = 0; k < KN; ++k) {
        const int slot = k * WG + tid;
        nlive[k] = false;
        nbr[k] = 0;
        nscale[k] = 0;
        ncur[k] = 0;
        nid[k] = slot < pk.nn ? int(perm[n0 + slot]) : pk.nn;
        const int n = nid[k];
        if (slot < pk.nn) {
            const int64_t i = n0 + n;
            nbr[k] = p.nbrain[i] - pk.b0;
            load_neuron(p, i, nr[k], nc[k], etab);
            nlive[k] = brun_l[nbr[k]] != 0;
            nscale[k] = ns_expf(p.ns[NS_LAS * N + i], etab);
            int32_t idx = p.ninput[i];
            ncur[k] = idx >= 0 ? input[idx] : -1.0f;   // level, or -1 for no input
            if (idx < 0) nc[k].flags |= 1u << 30;
        }
    }
    // Synapse registers.
    uint32_t sfl[KS], spp[KS];
    float sw0[KS];
    PlastRegs sp[KS];
    unsigned arrmask = 0;
#pragma unroll
    for (int k = 0; k < KS; ++k) {
        const int q = k * WG + tid;
        sfl[k] = 0;
        spp[k] = 0;
        sw0[k] = 0;
        sp[k] = PlastRegs{0, 0, 0, 0, 0};
        if (q < pk.ns) {
            const int64_t gq = s0 + q;
            uint32_t f = p.sflags[gq];
            if (!brun_l[sbr_l[q]]) f &= ~(R_VALID << SF_ROUTE_SHIFT);
            sfl[k] = f;
            spp[k] = uint32_t(p.spre[gq] - n0) | (uint32_t(p.spost[gq] - n0) << 16);
            sw0[k] = p.sg[SG_W0 * S + gq];
            int profile = (f >> SF_PLAST_SHIFT) & 7;
            load_plast(p, gq, profile, sp[k]);
        }
    }

    for (int t = 0; t < ticks; ++t) {
        // Phase A: plasticity of tick t-1, then arrivals of tick t.
#pragma unroll
        for (int k = 0; k < KS; ++k) {
            const int q = k * WG + tid;
            const uint32_t f = sfl[k];
            const uint32_t route = f >> SF_ROUTE_SHIFT;
            bool arrive_stp = false;
            if (route & R_VALID) {
                const int profile = (f >> SF_PLAST_SHIFT) & 7;
                const int post = spp[k] >> 16, pre = spp[k] & 0xffff;
                if (t > 0 && (route & R_PAIR) && profile && !bfrozen_l[sbr_l[q]])
                    plasticity_step(sp[k], profile, (f >> SF_MODREC_SHIFT) & 7, sw0[k],
                                    (arrmask >> k) & 1, hist_l[post] & 1, depol_l[post],
                                    mod_l[post], etab);
                const int64_t tick = btick_l[sbr_l[q]] + t;
                const int d = f & SF_DELAY;
                const bool arr = tick >= d && ((hist_l[pre] >> (d - 1)) & 1u);
                arrmask = (arrmask & ~(1u << k)) | (unsigned(arr) << k);
                rel_l[q] = arr ? sp[k].weight : 0.0f;
                arrive_stp = arr && (route & R_STP);
            } else if (q < pk.ns) {
                rel_l[q] = 0.0f;
                arrmask &= ~(1u << k);
            }
            // wavefront ballot compaction of arrivals that need short-term release
            const uint64_t m = __ballot(arrive_stp);
            if (m) {
                const int lane = tid & 63;
                int base = 0;
                if (lane == 0) base = atomicAdd(&qlen, __popcll(m));
                base = __shfl(base, 0);
                if (arrive_stp) queue_l[base + __popcll(m & ((uint64_t(1) << lane) - 1))] = uint16_t(q);
            }
        }
        unsigned my_arr = __popc(arrmask);
        for (int off = 32; off > 0; off >>= 1) my_arr += __shfl_xor(my_arr, off);
        if ((tid & 63) == 0 && my_arr) atomicAdd(&events_l, my_arr);
        __syncthreads();
        // Phase A2: short-term release over the compacted arrivals.
        {
            const int n_q = qlen;
            for (int j = tid; j < n_q; j += WG) {
                const int q = queue_l[j];
                const int64_t gq = s0 + q;
                const int64_t tick = btick_l[sbr_l[q]] + t;
                rel_l[q] = release_at(p, gq, p.sflags[gq], rel_l[q], tick, etab);
            }
        }
        __syncthreads();
        // Phase B: neurons.
        if (tid == 0) qlen = 0;
#pragma unroll
        for (int k = 0; k < KN; ++k) {
            const int n = nid[k];
            if (!nlive[k]) continue;
            const uint32_t flags = nc[k].flags;
            if (!(flags & NF_ACTIVE)) {
                const uint32_t h = hist_l[n];
                hist_l[n] = (flags & NF_WORLD) ? world_rotate(h, flags) : h << 1;
                continue;
            }
            const int64_t tick = btick_l[nbr[k]] + t;
            float e0 = 0, e1 = 0, e2 = 0, e3 = 0, inh = 0;
            unsigned mod = 0;
            const int q1 = in_l[n + 1];
            // Every release is >= +0 and the sums start at +0, so adding +0
            // for the other routes leaves them bit-identical: branch-free.
            for (int q = in_l[n]; q < q1; ++q) {
                const float r = rel_l[q];
                const uint32_t route = route_l[q];
                if ((route & R_MOD) && r > 0) {
                    const unsigned m = (route >> R_MOD_SHIFT) & 3u;
                    mod |= 1u << (m - 1);
                    uint64_t &st = p.stamps[((m - 1) * 2 + (tick & 1)) * N + n0 + n];
                    if (st < uint64_t(tick + 1)) st = uint64_t(tick + 1);
                }
                const bool exc = !(route & (R_MOD | R_INH));
                const float v = exc ? r * nscale[k] : 0.0f;
                const uint32_t br = route & R_BRANCH;
                inh += (route & R_MOD) ? 0.0f : ((route & R_INH) ? r : 0.0f);
                e0 += br == 0 ? v : 0.0f;
                e1 += br == 1 ? v : 0.0f;
                e2 += br == 2 ? v : 0.0f;
                e3 += br == 3 ? v : 0.0f;
            }
            float current = 0.0f;
            if (!(flags & (1u << 30)))
                current = (flags & NF_RECEPTOR) ? receptor_drive(nr[k], nc[k], ncur[k])
                                                : gain * ncur[k];
            const int spike = neuron_tick(nr[k], nc[k], current, e0, e1, e2, e3, inh, p.cell, N,
                                          n0 + n, etab);
            hist_l[n] = (hist_l[n] << 1) | uint32_t(spike);
            depol_l[n] = nr[k].depol;
            mod_l[n] = uint8_t(mod);
            if (!isfinite(nr[k].v)) p.bstatus[pk.b0 + nbr[k]] = -1;
        }
        __syncthreads();
        // Motor spike counts of tick t (core.cuh:68).
        for (int m = tid; m < pk.nm; m += WG)
            if (brun_l[m / p.outputs] && mot_l[m] != kNoMotor) cnt_l[m] += hist_l[mot_l[m]] & 1u;
    }
    // Plasticity of the last tick.
    if (ticks > 0) {
#pragma unroll
        for (int k = 0; k < KS; ++k) {
            const int q = k * WG + tid;
            const uint32_t f = sfl[k];
            const uint32_t route = f >> SF_ROUTE_SHIFT;
            const int profile = (f >> SF_PLAST_SHIFT) & 7;
            if ((route & R_VALID) && (route & R_PAIR) && profile && !bfrozen_l[sbr_l[q]]) {
                const int post = spp[k] >> 16;
                plasticity_step(sp[k], profile, (f >> SF_MODREC_SHIFT) & 7, sw0[k],
                                (arrmask >> k) & 1, hist_l[post] & 1, depol_l[post], mod_l[post],
                                etab);
            }
        }
    }
    // Write back.
#pragma unroll
    for (int k = 0; k < KN; ++k) {
        const int n = nid[k];
        if (n >= pk.nn || !brun_l[nbr[k]]) continue;
        const int64_t i = n0 + n;
        if (homeostasis && (nc[k].flags & NF_ACTIVE) && (nc[k].flags & NF_TYPE) == 0) {
            const int64_t tick = btick_l[nbr[k]] + ticks;
            const float hz = nr[k].total * 1000.0f / float(ticks > tick ? int64_t(ticks) : tick);
            p.ns[NS_LAS * N + i] = homeostasis_las(p.ns[NS_LAS * N + i], float(ticks), hz);
        }
        if (nc[k].flags & NF_ACTIVE) {
            store_neuron(p, i, nr[k]);
            p.spike[i] = uint8_t(hist_l[n] & 1u);
            p.modbits[i] = mod_l[n];
        } else if (!(nc[k].flags & NF_WORLD)) {
            p.spike[i] = 0;
        }
        p.hist[i] = hist_l[n];
    }
#pragma unroll
    for (int k = 0; k < KS; ++k) {
        const int q = k * WG + tid;
        if (q >= pk.ns || !brun_l[sbr_l[q]]) continue;
        const int64_t gq = s0 + q;
        const int profile = (sfl[k] >> SF_PLAST_SHIFT) & 7;
        store_plast(p, gq, profile, sp[k]);
        p.arrived[gq] = uint8_t((arrmask >> k) & 1);
    }
    for (int m = tid; m < pk.nm; m += WG) {
        const int b = m / p.outputs;
        // outputs start at zero (run_pool clears them); a brain that did not
        // run may share its output row with a slot's new occupant
        if (brun_l[b]) out[p.bout0[pk.b0 + b] + m % p.outputs] = float(cnt_l[m]);
    }
    for (int b = tid; b < pk.nb; b += WG)
        if (brun_l[b]) p.btick[pk.b0 + b] = btick_l[b] + ticks;
    __syncthreads();
    if (tid == 0 && events_l) atomicAdd(&p.events[0], (unsigned long long)events_l);
}

// ---------------------------------------------------------------------------
// Resident form, synapse state in LDS.
//
// Same tick order as k_resident, but every per-synapse value (weight and
// plasticity traces, genes, flags, endpoints, this tick's arrival) lives in
// dynamically sized LDS and threads stride over the pack's synapses.  Thread
// registers then hold only neuron state, which keeps the kernel under 128
// VGPRs.  PLASTIC = false drops the trace arrays and the initial weight (the
// plasticity bound) for pools without plastic synapses.  A tick's releases
// take no per-synapse room: an arrival without short-term plasticity
// releases its weight, and the short-term releases sit in a compact array of
// kRelCap entries named by the arrival code.
//
// Per tick: phase A (synapse-parallel) applies plasticity of tick t-1, detects
// the arrivals of tick t, marks the post neuron of every arrival and queues
// the arrivals that take the short-term release (wave ballot, one LDS atomic
// per wave); phase A2 runs those releases over the compacted queue; phase B
// (neuron-parallel) sums the inputs of marked neurons in ascending synapse
// order and integrates every neuron.  A neuron with no arrival this tick has
// only +0 releases, so its sums are +0 without reading them.
//
// Quiet plasticity steps are deferred, exactly.  A pair synapse whose step for
// tick t-1 has no input (no arrival, no spike of its post, no modulator at
// it) only decays its traces, ages and bounds its weight, and nothing reads
// that state before its next step with an input, its next arrival (whose
// release reads the weight) or the write-back.  So such steps are counted in
// the synapse's LDS word and replayed there, in tick order, by the same
// operations: the reference's quiet-interval shortcut (physiology.cuh:559-586,
// main.cu:1049-1106) without its closed form, so bit for bit the per-tick
// result.
// ---------------------------------------------------------------------------
template <bool PLASTIC>
__host__ __device__ constexpr int lds_synapse_bytes() {
    return PLASTIC ? 4 * 8 + 2 : 4 * 3 + 2;
}

// A synapse's endpoints and brain within its pack share one LDS word:
// pre in bits 0-10, post in bits 11-21, brain in bits 22-28.
constexpr int kPpPost = 11, kPpBrain = 22;
constexpr uint32_t kPpMask = 0x7ffu;

// A synapse's 16-bit LDS word (ai_l): in bits 0-8 its arrival code for the
// tick: no arrival; an arrival whose release is its weight (no short-term
// plasticity); k + 1 for the k-th queued short-term arrival (release in
// crel_l[k]); or a queued arrival past the kRelCap entries, whose release its
// post neuron's thread computes.  In bits 9-15, how many of its plasticity
// steps are deferred (see phase A).
constexpr uint32_t kArrNone = 0, kArrDirect = 0x101, kArrSpill = 0x102, kArrMask = 0x1ff;
constexpr int kRelCap = 256;
constexpr int kPendShift = 9;
constexpr uint32_t kPendMax = kQuietMax;

// A neuron's LDS modulator byte (mod_l) holds the tick's modulator bits
// (kModBits); in bit 6, whether a synapse onto it arrived this tick (set in
// phase A, where every writer stores the same byte, and read in phase B,
// whose store of the byte clears it); and in bit 7, whether its recent
// depolarization passes the profile-2 gate: plasticity reads the
// depolarization only through that comparison.
constexpr unsigned kModBits = kModNE | kMod5HT | kModDA, kArrivalMark = 0x40u, kDepolGate = 0x80u;
// NeuronConst flag bit (k_resident_lds): an AdEx cell without sensory input,
// whose tick may be skipped when its last one changed no state.
constexpr uint32_t kSkipCell = 1u << 29;
__device__ inline float gate_depol(unsigned m) { return (m & kDepolGate) ? kDepolGateMv : 0.0f; }

template <int WG, int KN, int MINW, bool PLASTIC>
__global__ __launch_bounds__(WG, MINW) void k_resident_lds(Pool p, const Pack *packs,
                                                         const int32_t *big, int nbig,
                                                         uint16_t *perm, int ticks,
                                                         const float *input, float gain,
                                                         float *out, int homeostasis, int smax) {
    constexpr int NMAX = WG * KN;
    static_assert(NMAX <= int(kPpMask) + 1 && kPackBrains <= 128, "pack indexes fit the endpoint word");
    extern __shared__ __align__(16) unsigned char lds_syn[];
    __shared__ uint64_t etab[32];
    __shared__ uint32_t hist_l[NMAX];
    __shared__ uint8_t mod_l[NMAX];          // modulator bits | kArrivalMark | kDepolGate
    __shared__ uint16_t in_l[NMAX + 1];
    __shared__ int64_t btick_l[kPackBrains];
    __shared__ uint8_t brun_l[kPackBrains], bfrozen_l[kPackBrains];
    __shared__ uint16_t mot_l[kPackMotors];
    __shared__ uint16_t cnt_l[kPackMotors];
    __shared__ int qlen;
    __shared__ unsigned events_l, skipped_l;
    __shared__ float crel_l[kRelCap];          // releases of the tick's short-term arrivals

    if (int(blockIdx.x) < nbig) {
        // the pack's synapse arrays are free in a big brain's workgroup; the
        // exponential's table is read from LDS, ordered by the first barrier
        if (threadIdx.x < 32) etab[threadIdx.x] = g_exp_tab[threadIdx.x];
        big_brain_ticks<WG>(p, big[blockIdx.x], ticks, input, gain, out, homeostasis, lds_syn,
                            int64_t(smax) * lds_synapse_bytes<PLASTIC>(), etab);
        return;
    }
    const Pack pk = packs[blockIdx.x - nbig];
    const int tid = threadIdx.x;
    if (tid < 32) etab[tid] = g_exp_tab[tid];   // ordered by the barrier below
    {   // the pack's brains; a pack whose brains all sit out this call (free
        // slots, masks) does nothing
        int any = 0;
        for (int b = tid; b < pk.nb; b += WG) {
            const uint8_t run = p.brun[pk.b0 + b] & 1;
            btick_l[b] = p.btick[pk.b0 + b];
            brun_l[b] = run;
            bfrozen_l[b] = p.bfrozen[pk.b0 + b];
            any |= run;
        }
        if (!__syncthreads_or(any)) return;
    }
    const int64_t N = p.ncap, S = p.scap;
    const int64_t n0 = pk.n0, s0 = pk.s0;
    const int ns = pk.ns;

    // synapse arrays (structure of arrays, smax entries each)
    float *w_l = reinterpret_cast<float *>(lds_syn);
    float *w0_l = PLASTIC ? w_l + smax : w_l;   // initial weight: plasticity's bound only
    uint32_t *fl_l = reinterpret_cast<uint32_t *>(PLASTIC ? w0_l + smax : w_l + smax);
    uint32_t *pp_l = fl_l + smax;
    float *x_l = reinterpret_cast<float *>(pp_l + smax);
    float *y_l = PLASTIC ? x_l + smax : x_l;
    float *pa_l = PLASTIC ? y_l + smax : x_l;
    float *qa_l = PLASTIC ? pa_l + smax : x_l;
    uint16_t *ai_l = reinterpret_cast<uint16_t *>(PLASTIC ? qa_l + smax : x_l);   // arrival code

    // Setup issues its loads in as few dependent rounds as it can, with no
    // barrier between them (the run flags are already in LDS): slot order,
    // neuron and synapse fields, then neuron registers with the flag-
    // dependent synapse traces, then the input level.
    int nid[KN];
#pragma unroll
    for (int k = 0; k < KN; ++k) {
        const int slot = k * WG + tid;
        nid[k] = slot < pk.nn ? int(perm[n0 + slot]) : pk.nn;
    }
    for (int n = tid; n < pk.nn; n += WG) {
        hist_l[n] = p.hist[n0 + n];
        in_l[n] = uint16_t(p.nin[n0 + n] - s0);
        mod_l[n] = 0;
    }
    if (tid == 0) {
        in_l[pk.nn] = uint16_t(ns);
        qlen = 0;
        events_l = 0;
        skipped_l = 0;
    }
    for (int m = tid; m < pk.nm; m += WG) {
        const int32_t mn = p.motor[int64_t(pk.b0) * p.outputs + m];
        mot_l[m] = mn >= 0 ? uint16_t(mn - n0) : kNoMotor;
        cnt_l[m] = 0;
    }
    // Synapse fields that do not depend on the flag word, SU synapses per
    // thread per round (a slot past the pack reloads the round's first
    // synapse; stores are guarded).  A synapse of a brain that sits out loses
    // its valid bit.
    constexpr int SU = 4;
    for (int q0 = tid; q0 < ns; q0 += SU * WG) {
        uint32_t f[SU], pp[SU];
        int b[SU];
        float w0[SU], w[SU], pa[SU], qa[SU];
#pragma unroll
        for (int u = 0; u < SU; ++u) {
            const int64_t gq = s0 + (q0 + u * WG < ns ? q0 + u * WG : q0);
            b[u] = p.sbrain[gq] - pk.b0;
            f[u] = p.sflags[gq];
            pp[u] = uint32_t(p.spre[gq] - n0) | (uint32_t(p.spost[gq] - n0) << kPpPost);
            if (PLASTIC) w0[u] = p.sg[SG_W0 * S + gq];
            w[u] = p.ss[SS_WEIGHT * S + gq];
            if (PLASTIC) {
                pa[u] = p.ss[SS_PRE_AGE * S + gq];
                qa[u] = p.ss[SS_POST_AGE * S + gq];
            }
        }
#pragma unroll
        for (int u = 0; u < SU; ++u) {
            const int q = q0 + u * WG;
            if (q >= ns) break;
            fl_l[q] = brun_l[b[u]] ? f[u] : f[u] & ~(R_VALID << SF_ROUTE_SHIFT);
            pp_l[q] = pp[u] | (uint32_t(b[u]) << kPpBrain);
            if (PLASTIC) w0_l[q] = w0[u];
            w_l[q] = w[u];
            ai_l[q] = uint16_t(kArrNone);
            if (PLASTIC) {
                pa_l[q] = pa[u];
                qa_l[q] = qa[u];
            }
        }
    }

    // Neuron registers, slots in type-sorted order (see k_resident).
    NeuronRegs nr[KN];
    NeuronConst nc[KN];
    float nscale[KN], ncur[KN];
    int nbr[KN], nidx[KN];
    bool nlive[KN];
    uint32_t carried = 0;   // bit k: kStillCarry of slot k's cell
#pragma unroll
    for (int k = 0; k < KN; ++k) {
        const int n = nid[k];
        nlive[k] = false;
        nbr[k] = 0;
        nscale[k] = 0;
        nidx[k] = -1;
        if (n < pk.nn) {
            const int64_t i = n0 + n;
            nbr[k] = p.nbrain[i] - pk.b0;
            load_neuron(p, i, nr[k], nc[k], etab);
            nscale[k] = p.ns[NS_LAS * N + i];
            nidx[k] = p.ninput[i];
            carried |= uint32_t((p.modbits[i] & kStillCarry) != 0) << k;
        }
    }
    // The profile-selected traces of each plastic synapse (the flag word
    // names the fields), thread-private LDS entries written above.
    if (PLASTIC) {
        for (int q0 = tid; q0 < ns; q0 += SU * WG) {
            float x[SU], y[SU];
#pragma unroll
            for (int u = 0; u < SU; ++u) {
                const int q = q0 + u * WG < ns ? q0 + u * WG : q0;
                const int profile = (fl_l[q] >> SF_PLAST_SHIFT) & 7;
                x[u] = p.ss[plast_x_field(profile) * S + s0 + q];
                y[u] = p.ss[plast_y_field(profile) * S + s0 + q];
            }
#pragma unroll
            for (int u = 0; u < SU; ++u) {
                const int q = q0 + u * WG;
                if (q >= ns) break;
                x_l[q] = x[u];
                y_l[q] = y[u];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < KN; ++k) {
        ncur[k] = 0;
        if (nid[k] < pk.nn) {
            nlive[k] = brun_l[nbr[k]] != 0;
            nscale[k] = ns_expf(nscale[k], etab);
            ncur[k] = nidx[k] >= 0 ? input[nidx[k]] : -1.0f;
            if (nidx[k] < 0) nc[k].flags |= 1u << 30;
            // an AdEx cell without sensory input: its only inputs are synaptic,
            // in this call and the next
            if (nidx[k] < 0 && (nc[k].flags & NF_TYPE) != 6) nc[k].flags |= kSkipCell;
            else carried &= ~(1u << k);
        }
    }
    __syncthreads();

    unsigned my_events = 0, my_skipped = 0;
    // bit k: neuron slot k's last tick (of this call, or of the previous one
    // for a cell without input) changed no state (see phase B); and whether
    // any tick of this call changed it (else write-back skips it)
    uint32_t still = carried, dirty = 0;
    uint32_t bad = 0;   // bit k: a tick of this call left slot k's v nonfinite
    for (int t = 0; t < ticks; ++t) {
        // Phase A: plasticity of tick t-1, arrivals of tick t; the arrivals
        // that take a short-term release are noted (one bit per round) and
        // released by the same thread after the sweep.
        uint32_t stp_rounds = 0;
        for (int base = 0, round = 0; base < ns; base += WG, ++round) {
            const int q = base + tid;
            if (q < ns) {
                const uint32_t f = fl_l[q];
                const uint32_t route = f >> SF_ROUTE_SHIFT;
                if (route & R_VALID) {
                    const uint32_t pp = pp_l[q];
                    const int pre = pp & kPpMask, post = (pp >> kPpPost) & kPpMask;
                    const int b = pp >> kPpBrain;
                    const int64_t tick = btick_l[b] + t;
                    const int d = f & SF_DELAY;
                    const bool arr = tick >= d && ((hist_l[pre] >> (d - 1)) & 1u);
                    uint32_t pend = 0;
                    if (PLASTIC) {
                        const int profile = (f >> SF_PLAST_SHIFT) & 7;
                        const uint32_t word = ai_l[q];
                        pend = word >> kPendShift;
                        // the step of tick t-1, deferred when it has no input
                        // and no arrival at t reads the weight next.  The
                        // test runs in every lane (a wave holds pair synapses
                        // almost always); only the step itself is a branch.
                        const bool pair = t > 0 && (route & R_PAIR) && profile && !bfrozen_l[b];
                        const unsigned pm = mod_l[post];
                        const unsigned modrec = (f >> SF_MODREC_SHIFT) & 7;
                        const bool arrived = (word & kArrMask) != kArrNone;
                        const bool spiked = hist_l[post] & 1;
                        const bool input = arrived || spiked || (pm & kModBits & modrec) || arr ||
                                           pend == kPendMax;
                        if (pair && input) {
                            PlastRegs r{w_l[q], x_l[q], y_l[q], pa_l[q], qa_l[q]};
                            const float w0 = w0_l[q];
                            replay_quiet_steps(r, profile, w0, pend);
                            plasticity_step(r, profile, modrec, w0, arrived, spiked, gate_depol(pm),
                                            pm & kModBits, etab);
                            w_l[q] = r.weight; x_l[q] = r.x; y_l[q] = r.y;
                            pa_l[q] = r.pre_age; qa_l[q] = r.post_age;
                        }
                        pend = pair ? (input ? 0u : pend + 1) : pend;
                    }
                    // short-term arrivals take their code below
                    ai_l[q] = uint16_t((pend << kPendShift) | (arr ? kArrDirect : kArrNone));
                    if (arr) mod_l[post] |= kArrivalMark;
                    my_events += arr;
                    if (arr && (route & R_STP)) stp_rounds |= 1u << round;
                }
            }
        }
        // Short-term releases (physiology.cuh:425-454) of the arrivals this
        // thread found: a few per pack and tick, so each is released where it
        // was found (its weight is final: plasticity of tick t-1 ran above, in
        // this thread) rather than after a barrier of its own.  A release
        // past the kRelCap entries is left to its post neuron's thread.
        for (uint32_t rounds = stp_rounds; rounds; rounds &= rounds - 1) {
            const int q = __builtin_ctz(rounds) * WG + tid;
            const uint32_t keep = ai_l[q] & ~kArrMask;   // deferred steps
            const int pos = atomicAdd(&qlen, 1);
            if (pos < kRelCap) {
                const int64_t tick = btick_l[pp_l[q] >> kPpBrain] + t;
                crel_l[pos] = release_at(p, s0 + q, fl_l[q], w_l[q], tick, etab);
                ai_l[q] = uint16_t(keep | uint32_t(pos + 1));
            } else {
                ai_l[q] = uint16_t(keep | kArrSpill);
            }
        }
        __syncthreads();
        // Phase B: neurons.
        if (tid == 0) qlen = 0;
#pragma unroll
        for (int k = 0; k < KN; ++k) {
            const int n = nid[k];
            if (!nlive[k]) continue;
            const uint32_t flags = nc[k].flags;
            if (!(flags & NF_ACTIVE)) {
                const uint32_t h = hist_l[n];
                hist_l[n] = (flags & NF_WORLD) ? world_rotate(h, flags) : h << 1;
                continue;
            }
            const bool arrived = (mod_l[n] & kArrivalMark) != 0;
            if (((still >> k) & 1u) && !arrived) {
                // Its last tick changed no state and no synapse onto it
                // arrives: its inputs are those of the last tick (+0 sums, no
                // sensory current), so this tick is the same identity: no
                // spike, no modulator bits.
                hist_l[n] <<= 1;
                mod_l[n] = uint8_t(mod_l[n] & kDepolGate);
                ++my_skipped;
                continue;
            }
            const int64_t tick = btick_l[nbr[k]] + t;
            float e0 = 0, e1 = 0, e2 = 0, e3 = 0, inh = 0;
            unsigned mod = 0;
            const int q1 = arrived ? int(in_l[n + 1]) : 0;
            for (int q = in_l[n]; q < q1; ++q) {
                // this tick's release: +0 without an arrival, the weight
                // without short-term plasticity, else the queued release
                const uint32_t a = ai_l[q] & kArrMask;
                float r = a == kArrNone ? 0.0f : a == kArrDirect ? w_l[q] : 0.0f;
                if (a != kArrNone && a != kArrDirect)
                    r = a == kArrSpill ? release_at(p, s0 + q, fl_l[q], w_l[q], tick, etab)
                                       : crel_l[a - 1];
                const uint32_t route = fl_l[q] >> SF_ROUTE_SHIFT;
                if ((route & R_MOD) && r > 0) {
                    const unsigned mm = (route >> R_MOD_SHIFT) & 3u;
                    mod |= 1u << (mm - 1);
                    uint64_t &st = p.stamps[((mm - 1) * 2 + (tick & 1)) * N + n0 + n];
                    if (st < uint64_t(tick + 1)) st = uint64_t(tick + 1);
                }
                const bool exc = !(route & (R_MOD | R_INH));
                const float v = exc ? r * nscale[k] : 0.0f;
                const uint32_t br = route & R_BRANCH;
                inh += (route & R_MOD) ? 0.0f : ((route & R_INH) ? r : 0.0f);
                e0 += br == 0 ? v : 0.0f;
                e1 += br == 1 ? v : 0.0f;
                e2 += br == 2 ? v : 0.0f;
                e3 += br == 3 ? v : 0.0f;
            }
            float current = 0.0f;
            if (!(flags & (1u << 30)))
                current = (flags & NF_RECEPTOR) ? receptor_drive(nr[k], nc[k], ncur[k])
                                                : gain * ncur[k];
            bool same = false;
            const int spike = neuron_tick(nr[k], nc[k], current, e0, e1, e2, e3, inh, p.cell, N,
                                          n0 + n, etab, &same);
            const bool unchanged = same && (flags & kSkipCell);
            still = (still & ~(1u << k)) | (uint32_t(unchanged) << k);
            dirty |= uint32_t(!unchanged) << k;
            hist_l[n] = (hist_l[n] << 1) | uint32_t(spike);
            mod_l[n] = uint8_t(mod | (nr[k].depol >= kDepolGateMv ? kDepolGate : 0u));
            bad |= uint32_t(!isfinite(nr[k].v)) << k;
        }
        __syncthreads();
        for (int m = tid; m < pk.nm; m += WG)
            if (brun_l[m / p.outputs] && mot_l[m] != kNoMotor) cnt_l[m] += hist_l[mot_l[m]] & 1u;
    }
    // Plasticity of the last tick.
    if (PLASTIC && ticks > 0) {
        for (int q = tid; q < ns; q += WG) {
            const uint32_t f = fl_l[q];
            const uint32_t route = f >> SF_ROUTE_SHIFT;
            const int profile = (f >> SF_PLAST_SHIFT) & 7;
            const uint32_t pp = pp_l[q];
            if ((route & R_VALID) && (route & R_PAIR) && profile && !bfrozen_l[pp >> kPpBrain]) {
                const int post = (pp >> kPpPost) & kPpMask;
                const uint32_t word = ai_l[q];
                PlastRegs r{w_l[q], x_l[q], y_l[q], pa_l[q], qa_l[q]};
                const float w0 = w0_l[q];
                replay_quiet_steps(r, profile, w0, word >> kPendShift);
                plasticity_step(r, profile, (f >> SF_MODREC_SHIFT) & 7, w0, (word & kArrMask) != kArrNone,
                                hist_l[post] & 1, gate_depol(mod_l[post]), mod_l[post] & kModBits,
                                etab);
                w_l[q] = r.weight; x_l[q] = r.x; y_l[q] = r.y; pa_l[q] = r.pre_age; qa_l[q] = r.post_age;
            }
        }
    }
    // Write back.
#pragma unroll
    for (int k = 0; k < KN; ++k) {
        const int n = nid[k];
        if (n >= pk.nn || !brun_l[nbr[k]]) continue;
        const int64_t i = n0 + n;
        if (homeostasis && (nc[k].flags & NF_ACTIVE) && (nc[k].flags & NF_TYPE) == 0) {
            const int64_t tick = btick_l[nbr[k]] + ticks;
            const float hz = nr[k].total * 1000.0f / float(ticks > tick ? int64_t(ticks) : tick);
            p.ns[NS_LAS * N + i] = homeostasis_las(p.ns[NS_LAS * N + i], float(ticks), hz);
        }
        if (nc[k].flags & NF_ACTIVE) {
            if ((bad >> k) & 1u) p.bstatus[pk.b0 + nbr[k]] = -1;
            if ((dirty >> k) & 1u) store_neuron(p, i, nr[k]);   // else the loaded state, unchanged
            p.spike[i] = uint8_t(hist_l[n] & 1u);
            // a cell without input that ends the call settled carries it over
            const bool carry = ((still >> k) & 1u) && (nc[k].flags & (1u << 30));
            p.modbits[i] = uint8_t((mod_l[n] & kModBits) | (carry ? kStillCarry : 0u));
        } else if (!(nc[k].flags & NF_WORLD)) {
            p.spike[i] = 0;
        }
        p.hist[i] = hist_l[n];
    }
    for (int q = tid; q < ns; q += WG) {
        if (!brun_l[pp_l[q] >> kPpBrain]) continue;
        const int64_t gq = s0 + q;
        if (PLASTIC) {
            // only the synapses plasticity runs on change their traces
            const uint32_t f = fl_l[q], route = f >> SF_ROUTE_SHIFT;
            const int profile = (f >> SF_PLAST_SHIFT) & 7;
            if ((route & R_VALID) && (route & R_PAIR) && profile && !bfrozen_l[pp_l[q] >> kPpBrain]) {
                PlastRegs r{w_l[q], x_l[q], y_l[q], pa_l[q], qa_l[q]};
                store_plast(p, gq, profile, r);
            }
        }
        p.arrived[gq] = uint8_t((ai_l[q] & kArrMask) != kArrNone);
    }
    for (int m = tid; m < pk.nm; m += WG) {
        const int b = m / p.outputs;
        // outputs start at zero (run_pool clears them); a brain that did not
        // run may share its output row with a slot's new occupant
        if (brun_l[b]) out[p.bout0[pk.b0 + b] + m % p.outputs] = float(cnt_l[m]);
    }
    for (int b = tid; b < pk.nb; b += WG)
        if (brun_l[b]) p.btick[pk.b0 + b] = btick_l[b] + ticks;
    for (int off = 32; off > 0; off >>= 1) {
        my_events += __shfl_xor(my_events, off);
        my_skipped += __shfl_xor(my_skipped, off);
    }
    if ((tid & 63) == 0 && my_events) atomicAdd(&events_l, my_events);
    if ((tid & 63) == 0 && my_skipped) atomicAdd(&skipped_l, my_skipped);
    // Slot order for the next launch: by cell class as the build orders it,
    // cells with sensory input (never skipped) first, then the unsettled
    // cells without input, then the slots whose last tick changed no state
    // (and those of brains that sat out), so that cells that skip share
    // wavefronts, now and once the unsettled ones settle.  The order only
    // assigns neurons to threads.
    {
        constexpr int W = WG / 64, K = 3 * kSlotClasses;
        const int lane = tid & 63, wave = tid >> 6;
        int *cnt = reinterpret_cast<int *>(lds_syn);   // [KN][W][K], then totals and bases
        int *tot = cnt + KN * W * K, *base = tot + K;
        int key[KN], rank[KN];
        __syncthreads();   // the synapse arrays are no longer read
#pragma unroll
        for (int k = 0; k < KN; ++k) {
            key[k] = -1;
            rank[k] = 0;
            if (nid[k] < pk.nn)
                key[k] = slot_class(nc[k].flags) +
                         ((((still >> k) & 1u) || !nlive[k]) ? 2 * kSlotClasses
                          : (nc[k].flags & (1u << 30)) ? kSlotClasses : 0);
            for (int c = 0; c < K; ++c) {
                const uint64_t m = __ballot(key[k] == c);
                if (lane == 0) cnt[(k * W + wave) * K + c] = __popcll(m);
                if (key[k] == c) rank[k] = __popcll(m & ((uint64_t(1) << lane) - 1));
            }
        }
        __syncthreads();
        if (tid < K) {
            int s = 0;
            for (int j = 0; j < KN * W; ++j) {
                const int x = cnt[j * K + tid];
                cnt[j * K + tid] = s;
                s += x;
            }
            tot[tid] = s;
        }
        __syncthreads();
        if (tid == 0)
            for (int c = 0, s = 0; c < K; ++c) {
                base[c] = s;
                s += tot[c];
            }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KN; ++k)
            if (key[k] >= 0)
                perm[n0 + base[key[k]] + cnt[(k * W + wave) * K + key[k]] + rank[k]] = uint16_t(nid[k]);
    }
    __syncthreads();
    if (tid == 0 && events_l) atomicAdd(&p.events[0], (unsigned long long)events_l);
    if (tid == 0 && skipped_l) atomicAdd(&p.events[2], (unsigned long long)skipped_l);
}

// ---------------------------------------------------------------------------
// Device math digest (diagnostic)
// ---------------------------------------------------------------------------
__global__ void k_math_digest(int fn, uint64_t lo, uint64_t hi, int64_t chunk,
                              unsigned long long *digests) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t u = lo + uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; u < hi; u += stride) {
        const float x = u2f(uint32_t(u));
        const float y = fn == 0 ? ns_expf(x, g_exp_tab) : ns_logf(x, g_log_tab);
        const uint32_t bits = y != y ? 0x7fc00000u : f2u(y);
        const unsigned long long term = (unsigned long long)bits * (unsigned long long)(u | 1u);
        atomicAdd(&digests[(u - lo) / uint64_t(chunk)], term);
    }
}

void launch_math_digest(int fn, uint64_t lo, uint64_t hi, int64_t chunk,
                        unsigned long long *digests, hipStream_t st) {
    hipLaunchKernelGGL(k_math_digest, dim3(256 * 64), dim3(256), 0, st, fn, lo, hi, chunk, digests);
}

// ---------------------------------------------------------------------------
// Launch wrappers
// ---------------------------------------------------------------------------
static inline unsigned blocks_for(int64_t n, int threads) {
    return unsigned((n + threads - 1) / threads);
}

void launch_reset(const Pool &p, int64_t nn, int64_t ns, int nb, const uint32_t *bseed,
                  const uint32_t *norig, const uint32_t *sorig, hipStream_t st) {
    if (nn) hipLaunchKernelGGL(k_reset_neurons, dim3(blocks_for(nn, 256)), dim3(256), 0, st, p, nn, bseed, norig);
    if (ns) hipLaunchKernelGGL(k_reset_synapses, dim3(blocks_for(ns, 256)), dim3(256), 0, st, p, ns, bseed, sorig);
    hipLaunchKernelGGL(k_set_ticks, dim3(blocks_for(nb, 256)), dim3(256), 0, st, p, nb, int64_t(0), int64_t(0), uint8_t(1));
}

void launch_reset_slots(const Pool &p, const int32_t *list, int count, const uint32_t *norig,
                        const uint32_t *sorig, hipStream_t st) {
    if (count > 0)
        hipLaunchKernelGGL(k_reset_slots, dim3(count, 4), dim3(256), 0, st, p, list, count, norig, sorig);
}

void launch_scatter(const ScatterEntry *entries, int n, const unsigned char *base, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_scatter, dim3(n), dim3(256), 0, st, entries, n, base);
}

void launch_pack_order(const uint32_t *nflags, const uint8_t *modbits, const Pack *packs,
                       const int32_t *list, int count, uint16_t *perm, hipStream_t st) {
    if (count > 0)
        hipLaunchKernelGGL(k_pack_order, dim3(count), dim3(kOrderThreads), 0, st, nflags, modbits,
                           packs, list, perm);
}

void launch_fill(float *x, int64_t n, float v, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_fill_f32, dim3(blocks_for(n, 256)), dim3(256), 0, st, x, n, v);
}
void launch_fill_i32(int32_t *x, int64_t n, int32_t v, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(n, 256)), dim3(256), 0, st, x, n, v);
}
void launch_clear_step(float *out, int64_t nout, int32_t *status, int64_t nstatus, hipStream_t st) {
    const int64_t n = std::max(nout, nstatus);
    if (n) hipLaunchKernelGGL(k_clear_step, dim3(blocks_for(n, 256)), dim3(256), 0, st, out, nout, status, nstatus);
}

void launch_add_counts(int32_t *total, const float *counts, int64_t n, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_add_counts, dim3(blocks_for(n, 256)), dim3(256), 0, st, total, counts, n);
}

void launch_per_tick(const Pool &p, const PlainRange &r, int ticks, uint8_t need,
                     const float *input, float gain, float *out, int homeostasis,
                     hipStream_t st) {
    const int T = 256;
    const int64_t ns = r.s_hi - r.s_lo, nn = r.n_hi - r.n_lo;
    const int64_t plast_threads = ns > r.slots ? ns : r.slots;
    for (int t = 0; t < ticks; ++t) {
        if (ns) hipLaunchKernelGGL(k_deliver, dim3(blocks_for(ns, T)), dim3(T), 0, st, p, r.s_lo, r.s_hi, t, need);
        if (nn) hipLaunchKernelGGL(k_neuron, dim3(blocks_for(nn, T)), dim3(T), 0, st, p, r.n_lo, r.n_hi, t, need, input, gain);
        if (r.slow_count)
            hipLaunchKernelGGL(k_neuron_slow, dim3(blocks_for(r.slow_count, 64)), dim3(64), 0, st, p, r.slow, r.slow_count, t, need, input, gain);
        if (plast_threads)
            hipLaunchKernelGGL(k_plasticity, dim3(blocks_for(plast_threads, T)), dim3(T), 0, st, p, r.s_lo, r.s_hi, t, need, r.slots, out);
    }
    if (homeostasis && nn)
        hipLaunchKernelGGL(k_homeostasis_gym, dim3(blocks_for(nn, T)), dim3(T), 0, st, p, r.n_lo, r.n_hi, ticks, need);
    hipLaunchKernelGGL(k_set_ticks, dim3(blocks_for(r.brains, T)), dim3(T), 0, st, p, r.brains, int64_t(ticks), int64_t(-1), need);
}

void launch_world_homeostasis(const Pool &p, int64_t nn, const int32_t *age, hipStream_t st) {
    if (nn) hipLaunchKernelGGL(k_homeostasis_world, dim3(blocks_for(nn, 256)), dim3(256), 0, st, p, nn, age);
}

template <int WG, int KN, int KS, int MINW>
static hipError_t launch_resident_t(const Pool &p, const Pack *packs, int npacks,
                                    const int32_t *big, int nbig, const uint16_t *perm, int ticks,
                                    const float *input, float gain, float *out, int homeostasis,
                                    hipStream_t st) {
    hipLaunchKernelGGL((k_resident<WG, KN, KS, MINW>), dim3(nbig + npacks), dim3(WG), 0, st, p, packs,
                       big, nbig, perm, ticks, input, gain, out, homeostasis);
    return hipGetLastError();
}

template <int WG, int KN, int MINW, bool PLASTIC>
static hipError_t launch_lds_t(const Pool &p, const Pack *packs, int npacks, const int32_t *big,
                               int nbig, uint16_t *perm, int ticks, const float *input,
                               float gain, float *out, int homeostasis, int smax, hipStream_t st) {
    const size_t bytes = size_t(smax) * lds_synapse_bytes<PLASTIC>() + 16;
    // dynamic LDS limit raised so far for this instance, per device (threads
    // that launch on one device at the same time raise it the same way)
    constexpr int kMaxDevices = 64;
    static std::atomic<size_t> allowed[kMaxDevices];
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices || bytes > allowed[dev].load(std::memory_order_relaxed)) {
        e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_resident_lds<WG, KN, MINW, PLASTIC>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes));
        if (e != hipSuccess) return e;
        if (dev >= 0 && dev < kMaxDevices) {
            size_t cur = allowed[dev].load(std::memory_order_relaxed);
            while (cur < bytes && !allowed[dev].compare_exchange_weak(cur, bytes)) {}
        }
    }
    hipLaunchKernelGGL((k_resident_lds<WG, KN, MINW, PLASTIC>), dim3(nbig + npacks), dim3(WG), bytes,
                       st, p, packs, big, nbig, perm, ticks, input, gain, out, homeostasis, smax);
    return hipGetLastError();
}

// Explicit instantiations so the device pass emits every tile's kernel.
// MINW is the launch bound's minimum waves per SIMD (it caps VGPRs).
#define NS_INSTANTIATE(WG, KN, KS, MINW)                                                       \
    template __global__ void k_resident<WG, KN, KS, MINW>(Pool, const Pack *, const int32_t *, int, \
                                                          const uint16_t *, int, const float *,  \
                                                          float, float *, int);
#define NS_INSTANTIATE_LDS(WG, KN, MINW)                                                        \
    template __global__ void k_resident_lds<WG, KN, MINW, true>(                                 \
        Pool, const Pack *, const int32_t *, int, uint16_t *, int, const float *, float,        \
        float *, int, int);                                                                     \
    template __global__ void k_resident_lds<WG, KN, MINW, false>(                                \
        Pool, const Pack *, const int32_t *, int, uint16_t *, int, const float *, float,        \
        float *, int, int);
NS_INSTANTIATE(256, 1, 4, 1)
NS_INSTANTIATE(256, 1, 4, 2)
NS_INSTANTIATE(256, 1, 4, 3)
NS_INSTANTIATE(256, 2, 4, 1)
NS_INSTANTIATE(256, 2, 8, 1)
NS_INSTANTIATE(512, 1, 4, 1)
NS_INSTANTIATE(512, 1, 4, 2)
NS_INSTANTIATE(512, 2, 4, 1)
NS_INSTANTIATE(512, 1, 8, 1)
NS_INSTANTIATE(1024, 1, 4, 1)
NS_INSTANTIATE(1024, 1, 8, 1)
NS_INSTANTIATE(128, 2, 8, 1)
NS_INSTANTIATE(128, 1, 4, 4)
NS_INSTANTIATE_LDS(1024, 1, 1)
NS_INSTANTIATE_LDS(1024, 2, 1)
NS_INSTANTIATE_LDS(512, 1, 1)
NS_INSTANTIATE_LDS(512, 2, 1)
NS_INSTANTIATE_LDS(256, 1, 1)
NS_INSTANTIATE_LDS(256, 2, 1)

// Host-only table: clang would otherwise promote this const array of host
// function pointers into device constant memory.
#if !defined(__HIP_DEVICE_COMPILE__)
#define NS_TILE(WG, KN, KS, MINW) {WG, KN, KS, MINW, 0, 0, &launch_resident_t<WG, KN, KS, MINW>, nullptr, nullptr}
#define NS_TILE_LDS(WG, KN, MINW, SP, SS)                                                      \
    {WG, KN, 0, MINW, SP, SS, nullptr, &launch_lds_t<WG, KN, MINW, true>,                      \
     &launch_lds_t<WG, KN, MINW, false>}
// phase A of k_resident_lds notes short-term arrivals in one bit per round
// of WG synapses: at most 32 rounds per pack
constexpr bool lds_rounds_fit(int wg, int sp, int ss) { return (sp + wg - 1) / wg <= 32 && (ss + wg - 1) / wg <= 32; }
static_assert(lds_rounds_fit(1024, 4448, 10784) && lds_rounds_fit(1024, 4224, 10272) &&
              lds_rounds_fit(1024, 3936, 9536) && lds_rounds_fit(1024, 2048, 3072) &&
              lds_rounds_fit(512, 1792, 2944) && lds_rounds_fit(512, 4448, 10784) &&
              lds_rounds_fit(256, 896, 1408) && lds_rounds_fit(256, 1792, 2944),
              "LDS tile with more than 32 synapse rounds");
const TileShape kTiles[] = {
    NS_TILE(256, 1, 4, 1),
    NS_TILE(256, 1, 4, 2),
    NS_TILE(256, 1, 4, 3),
    NS_TILE(256, 2, 4, 1),
    NS_TILE(256, 2, 8, 1),
    NS_TILE(512, 1, 4, 1),
    NS_TILE(512, 1, 4, 2),
    NS_TILE(512, 2, 4, 1),
    NS_TILE(512, 1, 8, 1),
    NS_TILE(1024, 1, 4, 1),
    NS_TILE(1024, 1, 8, 1),
    NS_TILE(128, 2, 8, 1),
    NS_TILE(128, 1, 4, 4),
    NS_TILE_LDS(1024, 1, 1, 4448, 10784),
    // two neurons per thread (2,048 per pack): spills about 85 registers per
    // lane to scratch, yet keeps brains of 1,025-2,048 neurons out of the
    // one-workgroup-per-brain path (8% faster on the 4M-neuron Pareto mix)
    NS_TILE_LDS(1024, 2, 1, 4224, 10272),
    // the same tile with about 6% lighter packs: a launch takes whole rounds
    // of one pack per CU, and where a pool's packs fill a round only in part
    // (2,048 brains at 20 synapses per neuron) lighter packs finish sooner
    NS_TILE_LDS(1024, 2, 1, 3936, 9536),
    NS_TILE_LDS(1024, 1, 1, 2048, 3072),
    NS_TILE_LDS(512, 1, 1, 1792, 2944),
    NS_TILE_LDS(512, 2, 1, 4448, 10784),
    NS_TILE_LDS(256, 1, 1, 896, 1408),
    NS_TILE_LDS(256, 2, 1, 1792, 2944),
};
const int kNumTiles = int(sizeof(kTiles) / sizeof(kTiles[0]));
#endif

}  // namespace nsa
