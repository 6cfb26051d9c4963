// Batched-episode surface of libns_amd.so: genome handles (ns_create, a
// drop-in for gym_adapter/native.cpp:46-94) and the device batch (ns_gpu_*,
// a drop-in for gym_adapter/gpu.cu:66-123), plus the pool builder shared by
// every surface.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <thread>
#include <set>
#include <tuple>

#include "ns_host.h"

namespace nsa {

thread_local std::string g_error;
void set_error(const std::string &msg) { g_error = msg; }

// ---------------------------------------------------------------------------
// Neuron profiles (src/physiology.cuh:82-176, prinz_cell.cuh:24-28).
// ---------------------------------------------------------------------------
static ns_neuron_gene blank_gene() {
    ns_neuron_gene g;
    std::memset(&g, 0, sizeof g);
    g.sensory_channel = -1;
    g.receptor_decay_ms = 3;
    g.target_affinity = 1;
    return g;
}

static void set_membrane(ns_neuron_gene &g, const float (&v)[11]) {
    g.capacitance_pf = v[0]; g.leak_ns = v[1]; g.rest_mv = v[2]; g.threshold_mv = v[3];
    g.slope_mv = v[4]; g.adapt_ns = v[5]; g.adapt_tau_ms = v[6]; g.spike_adapt_pa = v[7];
    g.reset_mv = v[8]; g.afterdepol_spike_pa = v[9]; g.afterdepol_tau_ms = v[10];
}

ns_neuron_gene profile_gene(int type) {
    static const float kPyramidal[11] = {104, 4.3f, -65, -52, 0.8f, -0.8f, 88, 65, -53, 0, 0};
    static const float kInterneuron[11] = {59, 2.9f, -62, -42, 3.0f, 1.8f, 16, 61, -54, 0, 0};
    ns_neuron_gene g = blank_gene();
    g.active = 1;
    const bool fast_spiking = type == 1 || type == 2 || type == 3;
    set_membrane(g, fast_spiking ? kInterneuron : kPyramidal);
    g.inhibitory_transmitter = fast_spiking ? 1 : 0;
    g.type = uint8_t(type > 6 ? 6 : type);
    switch (g.type) {
    case 0:
        g.reset_mv = -68.0f; g.afterdepol_spike_pa = 120.0f; g.afterdepol_tau_ms = 45.0f;
        break;
    case 2:
        for (float &x : g.lp_scale) x = 1;
        break;
    case 3: {
        const float pm[8] = {400 * 628.0f, 2.5f * 628, 6 * 628.0f, 50 * 628.0f,
                             10 * 628.0f, 100 * 628.0f, .01f * 628, 0};
        std::memcpy(g.prinz_g, pm, sizeof pm);
        break;
    }
    case 4:
        g.capacitance_pf = 8.7f; g.leak_ns = 8.7f / 12.0f; g.rest_mv = -60.0f;
        g.threshold_mv = -40.0f; g.reset_mv = -60.0f; g.adapt_ns = 0; g.spike_adapt_pa = 0;
        break;
    case 5:
        g.capacitance_pf = 10.0f; g.leak_ns = 0.5f; g.rest_mv = -60.0f; g.threshold_mv = -40.0f;
        g.reset_mv = -60.0f; g.adapt_ns = 0; g.spike_adapt_pa = 0;
        g.receptor_baseline_hz = 7.0f; g.receptor_gain_hz = 43.0f;
        g.receptor_event_ns = 2.0f; g.receptor_decay_ms = 3.0f;
        break;
    case 6:
        g.capacitance_pf = 1200.0f; g.leak_ns = 1000.0f / 65.6f; g.rest_mv = -45.1f;
        g.threshold_mv = -51.8f;
        break;
    default:
        break;
    }
    return g;
}

// gym_adapter/native.cpp:46-82, with the delay bound generalised.
HostBrain *parse_gym_brain(int count, int edges, int channels, int outputs, const float *ng,
                           const float *sg, const int *motors, int max_delay) {
    if (count < 1 || edges < 0 || channels < 1 || channels > 65535 || outputs < 1)
        throw std::runtime_error("Invalid brain dimensions");
    if (max_delay < 1 || max_delay > NS_MAX_DELAY) throw std::runtime_error("Invalid delay bound");
    auto *b = new HostBrain;
    try {
        b->channels = channels;
        b->max_delay = max_delay;
        for (int i = 0; i < count; ++i) {
            const float *p = ng + size_t(i) * 10;
            if (p[0] < 0 || p[0] > 6 || p[1] < -1 || p[1] >= channels || p[2] < 0 || p[2] > 1 ||
                p[3] < 0 || p[3] > 1 || p[4] < 0 || p[4] > 3)
                throw std::runtime_error("Invalid neuron gene");
            ns_neuron_gene g = profile_gene(int(uint8_t(p[0])));
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
                p[6] > 1 || p[7] <= 0 || p[8] < 0 || p[9] < 0 || p[9] > 5 || p[10] < 0 ||
                p[10] > 7)
                throw std::runtime_error("Invalid synapse gene");
            ns_synapse_gene g;
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

static inline bool excitatory_cell(unsigned t) { return t == 0 || t == 4 || t == 5; }

// ---------------------------------------------------------------------------
// Pool builder
// ---------------------------------------------------------------------------
namespace {
struct Bin {
    int nn = 0, ns = 0, nb = 0;
    std::vector<int> members;
};
}  // namespace

const TileShape *find_tile(const Tile &t) {
    for (int k = 0; k < kNumTiles; ++k) {
        const TileShape &s = kTiles[k];
        if (s.threads == t.threads && s.neurons_per_thread == t.kn && s.synapses_per_thread == t.ks &&
            s.min_waves == t.minw && (t.ks > 0 || s.lds_plastic == t.lds))
            return &s;
    }
    return nullptr;
}

// A brain plastic pair exists: a valid excitatory-to-excitatory synapse with a
// plasticity profile (the pairs the plastic kernels keep traces for).
bool has_plastic_pair(const HostBrain &b) {
    for (const auto &g : b.edges) {
        const ns_neuron_gene &a = b.genes[g.pre], &z = b.genes[g.post];
        if (g.plasticity && g.active && a.active && z.active && !a.inhibitory_transmitter &&
            excitatory_cell(a.type) && !a.modulator && !z.inhibitory_transmitter &&
            excitatory_cell(z.type))
            return true;
    }
    return false;
}

bool has_slow_cell(const HostBrain &b) {
    for (const auto &g : b.genes)
        if (g.type == 2 || g.type == 3) return true;
    return false;
}

// Structure-of-arrays image of one slot: the occupant (or nothing) laid out
// at pool positions n0.., s0.. in a slot of cap_n neurons and cap_s synapses.
// Neurons past the occupant are inactive with empty incoming runs; synapses
// past it are invalid and point at the slot's first neuron, so the tail never
// changes a result.
struct SlotImage {
    int cap_n = 0, cap_s = 0;
    std::vector<float> ng, nx, sg;
    std::vector<uint32_t> nflags, norig, sflags, sorig;
    std::vector<int32_t> ninput, nin, spre, spost, motor, syn_perm;
    int32_t bout0 = 0;

    void make(const BrainSpec *spec, int64_t n0, int64_t s0, int cn, int cs, int outs) {
        cap_n = cn; cap_s = cs;
        ng.assign(size_t(NG_FIELDS) * cn, 0.0f);
        nx.assign(size_t(NX_FIELDS) * cn, 0.0f);
        nflags.assign(cn, 0u); norig.assign(cn, 0u);
        ninput.assign(cn, -1); nin.assign(cn, int32_t(s0));
        sg.assign(size_t(SG_FIELDS) * cs, 0.0f);
        sflags.assign(cs, 1u); sorig.assign(cs, 0u);          // delay 1, invalid route
        spre.assign(cs, int32_t(n0)); spost.assign(cs, int32_t(n0));
        syn_perm.assign(cs, -1);
        motor.assign(outs, -1);
        bout0 = 0;
        if (!spec || !spec->brain) return;
        const HostBrain &b = *spec->brain;
        const int nc = int(b.genes.size()), sc = int(b.edges.size());
        bout0 = spec->output_base;
        // main.cu:718,742-743: under the world rules an inactive neuron counts
        // no spikes (its register bit 0 is a kept ring slot, not a spike)
        for (int j = 0; j < outs; ++j)
            motor[j] = b.motors[j] >= 0 && (!spec->world || b.genes[b.motors[j]].active)
                           ? int32_t(n0 + b.motors[j]) : -1;
        for (int i = 0; i < nc; ++i) {
            const ns_neuron_gene &g = b.genes[i];
            const float f[NG_FIELDS] = {g.capacitance_pf, g.leak_ns, g.rest_mv, g.threshold_mv,
                                        g.slope_mv, g.adapt_ns, g.adapt_tau_ms, g.spike_adapt_pa,
                                        g.reset_mv, g.afterdepol_spike_pa, g.afterdepol_tau_ms,
                                        g.receptor_baseline_hz, g.receptor_gain_hz,
                                        g.receptor_event_ns, g.receptor_decay_ms};
            for (int k = 0; k < NG_FIELDS; ++k) ng[size_t(k) * cn + i] = f[k];
            for (int k = 0; k < 4; ++k) nx[size_t(NX_LP + k) * cn + i] = g.lp_scale[k];
            for (int k = 0; k < 8; ++k) nx[size_t(NX_PRINZ + k) * cn + i] = g.prinz_g[k];
            nx[size_t(NX_AMBIENT) * cn + i] = g.prinz_ambient_leak_ns;
            nflags[i] = uint32_t(g.type) | (g.active ? NF_ACTIVE : 0) |
                        (b.receptor[i] ? NF_RECEPTOR : 0) | (spec->world ? world_flags(b.max_delay) : 0);
            ninput[i] = b.sensory[i] >= 0 ? spec->input_base + b.sensory[i] : -1;
            norig[i] = b.orig_n.empty() ? uint32_t(i) : b.orig_n[i];
        }
        // synapses sorted by (post, edge index): each neuron's incoming run is
        // contiguous and in the reference's accumulation order
        std::vector<int> perm(sc);
        std::iota(perm.begin(), perm.end(), 0);
        std::stable_sort(perm.begin(), perm.end(),
                         [&](int x, int y) { return b.edges[x].post < b.edges[y].post; });
        std::vector<int> incoming(nc + 1, 0);
        for (int j = 0; j < sc; ++j) incoming[b.edges[j].post + 1]++;
        for (int i = 0; i < nc; ++i) incoming[i + 1] += incoming[i];
        for (int i = 0; i < cn; ++i) nin[i] = int32_t(s0 + (i < nc ? incoming[i] : sc));
        for (int k = 0; k < sc; ++k) {
            const int j = perm[k];
            const ns_synapse_gene &g = b.edges[j];
            const ns_neuron_gene &a = b.genes[g.pre], &z = b.genes[g.post];
            const float f[SG_FIELDS] = {g.initial_weight_ns, g.utilization, g.recovery_ms,
                                        g.facilitation_ms, g.slow_recovery_ms, g.slow_depletion};
            for (int m = 0; m < SG_FIELDS; ++m) sg[size_t(m) * cs + k] = f[m];
            uint32_t route = uint32_t(g.branch & 3);
            if (a.modulator) route |= uint32_t(a.modulator & 3) << R_MOD_SHIFT;
            else if (a.inhibitory_transmitter) route |= R_INH;
            if (g.active && a.active && z.active) route |= R_VALID;
            if (!a.inhibitory_transmitter || a.modulator) route |= R_STP;
            if (!a.inhibitory_transmitter && excitatory_cell(a.type) && !a.modulator &&
                !z.inhibitory_transmitter && excitatory_cell(z.type))
                route |= R_PAIR;
            sflags[k] = uint32_t(g.delay) | (uint32_t(g.plasticity & 7) << SF_PLAST_SHIFT) |
                        (uint32_t(g.mod_receptors & 7) << SF_MODREC_SHIFT) |
                        (uint32_t(site_count(g)) << SF_SITES_SHIFT) | (route << SF_ROUTE_SHIFT);
            spre[k] = int32_t(n0 + g.pre);
            spost[k] = int32_t(n0 + g.post);
            sorig[k] = b.orig_s.empty() ? uint32_t(j) : b.orig_s[j];
            syn_perm[k] = j;
        }
    }
};

// Type-sorted slot order of one pack (cortical AdEx (type 0), receptor cells
// (5), other AdEx (1, 4), LFS (6), then inactive cells): waves of one type
// take the same branches.
// Host form of k_pack_order (used when a pool is built).
static void pack_order(const Pack &pk, const uint32_t *nflags, uint16_t *perm) {
    int start[kSlotClasses + 1] = {};     // stable counting sort by class
    for (int j = 0; j < pk.nn; ++j) start[slot_class(nflags[j]) + 1]++;
    for (int c = 0; c < kSlotClasses; ++c) start[c + 1] += start[c];
    for (int j = 0; j < pk.nn; ++j) perm[start[slot_class(nflags[j])]++] = uint16_t(j);
}

void DevicePool::build(const std::vector<BrainSpec> &specs, int outs, const Tile &t,
                       hipStream_t st) {
    arena.release();
    stream = st;
    tile = t;
    outputs = outs;
    nb = int(specs.size());
    ++generation;
    const TileShape *shape = find_tile(t);
    if (!shape) throw std::runtime_error("Tile shape not compiled");
    plastic = force_plastic;
    nbatch = 0;
    for (const auto &spec : specs) {
        if (!spec.brain) continue;
        if (has_plastic_pair(*spec.brain)) plastic = true;
        nbatch = std::max(nbatch, spec.batch_index + 1);
    }
    const int nmax = t.nmax();
    const int smax = t.ks > 0 ? t.threads * t.ks : (plastic ? shape->lds_plastic : shape->lds_static);
    smax_lds = t.ks > 0 ? 0 : smax;
    const int brains_per_pack = std::min(kPackBrains, std::max(1, kPackMotors / std::max(1, outs)));

    // 1. Slot sizes and classes: slots whose brain has LP/Prinz cells take the
    // per-tick kernels; slots too large for a pack get a workgroup of their
    // own in the resident launch; the rest are packed several per workgroup.
    std::vector<int> capn(nb), caps(nb);
    std::vector<uint8_t> slow(nb, 0), eligible(nb, 0), large(nb, 0);
    for (int i = 0; i < nb; ++i) {
        const BrainSpec &spec = specs[i];
        const int on = spec.brain ? int(spec.brain->genes.size()) : 0;
        const int os = spec.brain ? int(spec.brain->edges.size()) : 0;
        capn[i] = std::max(std::max(on, spec.cap_n), 1);
        caps[i] = std::max(os, spec.cap_s);
        slow[i] = spec.brain && has_slow_cell(*spec.brain);
        eligible[i] = !slow[i] && capn[i] <= nmax && caps[i] <= smax && outs <= kPackMotors;
        large[i] = !slow[i] && !eligible[i];
    }
    // 2. First-fit-decreasing packing of resident slots by cost.
    std::vector<int> order;
    for (int i = 0; i < nb; ++i)
        if (eligible[i]) order.push_back(i);
    auto cost = [&](int i) { return int64_t(capn[i]) * 24 + int64_t(caps[i]); };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost(a) > cost(b); });
    std::vector<Bin> bins;
    // the last 1,024 bins stay open: on a 32,768-brain Pareto mix, 64 open
    // bins left 8% more packs than the volume bound, 1,024 leave 2%
    const size_t window = 1024;
    size_t open_from = 0;
    for (int i : order) {
        const int nc = capn[i], sc = caps[i];
        bool placed = false;
        for (size_t k = open_from; k < bins.size(); ++k) {
            Bin &bin = bins[k];
            if (bin.nn + nc <= nmax && bin.ns + sc <= smax && bin.nb < brains_per_pack) {
                bin.nn += nc; bin.ns += sc; bin.nb += 1; bin.members.push_back(i);
                placed = true;
                break;
            }
        }
        if (!placed) {
            bins.emplace_back();
            Bin &bin = bins.back();
            bin.nn = nc; bin.ns = sc; bin.nb = 1; bin.members.push_back(i);
            if (bins.size() - open_from > window) ++open_from;
        }
    }
    // 3. Pool order: per-tick slots, then large slots (largest first, so
    // their long-running workgroups start early), then bins.
    std::vector<int> pool_order, large_order;
    for (int i = 0; i < nb; ++i)
        if (slow[i]) pool_order.push_back(i);
        else if (large[i]) large_order.push_back(i);
    fallback_brains = int(pool_order.size());
    std::stable_sort(large_order.begin(), large_order.end(),
                     [&](int a, int b) { return cost(a) > cost(b); });
    nbig = int(large_order.size());
    for (int i : large_order) pool_order.push_back(i);
    for (const Bin &bin : bins)
        for (int i : bin.members) pool_order.push_back(i);

    // 4. Positions.
    bn0.assign(nb, 0); bnc.assign(nb, 0); bs0.assign(nb, 0); bsc.assign(nb, 0);
    occ_n.assign(nb, 0); occ_s.assign(nb, 0);
    batch_of.assign(nb, -1); pool_of.assign(nbatch, -1); fallback.assign(nb, 0);
    slot_kind.assign(nb, 2); pack_of.assign(nb, -1);
    nn = 0; ns = 0;
    for (int pb = 0; pb < nb; ++pb) {
        const int i = pool_order[pb];
        const BrainSpec &spec = specs[i];
        bn0[pb] = int32_t(nn); bnc[pb] = capn[i];
        bs0[pb] = int32_t(ns); bsc[pb] = caps[i];
        nn += bnc[pb]; ns += bsc[pb];
        if (spec.brain) {
            occ_n[pb] = int32_t(spec.brain->genes.size());
            occ_s[pb] = int32_t(spec.brain->edges.size());
            batch_of[pb] = spec.batch_index;
            pool_of[spec.batch_index] = pb;
        }
        fallback[pb] = pb < fallback_brains;
        slot_kind[pb] = pb < fallback_brains ? 0 : pb < fallback_brains + nbig ? 1 : 2;
    }
    if (nn >= (int64_t(1) << 30) || ns >= (int64_t(1) << 30))
        throw std::runtime_error("Pool exceeds 2^30 elements");

    // 5. Host structure-of-arrays images, one slot at a time, in arrays with
    // the headroom of a growable pool past the last slot (inactive neurons,
    // invalid synapses until append_slots uses it).
    int64_t head_n = 0, head_s = 0;
    int head_slots = 0;
    if (growable) {
        int max_n = 1, max_s = 1;
        for (int pb = 0; pb < nb; ++pb) {
            max_n = std::max(max_n, bnc[pb]);
            max_s = std::max(max_s, bsc[pb]);
        }
        head_n = std::max<int64_t>(nn / 16, 8 * int64_t(max_n));
        head_s = std::max<int64_t>(ns / 16, 8 * int64_t(max_s));
        head_slots = std::max(64, nb / 16);
    }
    const int64_t N = std::max<int64_t>(nn + head_n, 1), S = std::max<int64_t>(ns + head_s, 1);
    std::vector<float> ng(NG_FIELDS * N, 0.0f), nx(NX_FIELDS * N, 0.0f);
    std::vector<uint32_t> nflags(N, 0), norig(N, 0);
    std::vector<int32_t> ninput(N, -1), nin(N + 1, 0), nbrain(N, 0), slow_pos;
    std::vector<float> sg(SG_FIELDS * S, 0.0f);
    std::vector<uint32_t> sflags(S, 0), sorig(S, 0);
    std::vector<int32_t> spre(S, 0), spost(S, 0), sbrain(S, 0), motor(int64_t(nb) * outs, 0);
    std::vector<int32_t> bout0(nb, 0);
    syn_perm.assign(S, 0);
    SlotImage img;
    for (int pb = 0; pb < nb; ++pb) {
        const BrainSpec &spec = specs[pool_order[pb]];
        const int64_t n0 = bn0[pb], s0 = bs0[pb];
        const int cn = bnc[pb], cs = bsc[pb];
        img.make(&spec, n0, s0, cn, cs, outs);
        bout0[pb] = img.bout0;
        for (int j = 0; j < outs; ++j) motor[int64_t(pb) * outs + j] = img.motor[j];
        for (int i = 0; i < cn; ++i) {
            const int64_t n = n0 + i;
            for (int k = 0; k < NG_FIELDS; ++k) ng[k * N + n] = img.ng[size_t(k) * cn + i];
            for (int k = 0; k < NX_FIELDS; ++k) nx[k * N + n] = img.nx[size_t(k) * cn + i];
            nflags[n] = img.nflags[i];
            ninput[n] = img.ninput[i];
            nin[n] = img.nin[i];
            nbrain[n] = pb;
            norig[n] = img.norig[i];
            const uint32_t ty = img.nflags[i] & NF_TYPE;
            if (i < occ_n[pb] && (ty == 2 || ty == 3)) slow_pos.push_back(int32_t(n));
        }
        for (int k = 0; k < cs; ++k) {
            const int64_t q = s0 + k;
            for (int m = 0; m < SG_FIELDS; ++m) sg[m * S + q] = img.sg[size_t(m) * cs + k];
            sflags[q] = img.sflags[k];
            spre[q] = img.spre[k];
            spost[q] = img.spost[k];
            sbrain[q] = pb;
            sorig[q] = img.sorig[k];
            syn_perm[q] = img.syn_perm[k];
        }
    }
    nin[nn] = int32_t(ns);
    for (int64_t n = nn; n < N; ++n) nin[n + 1] = int32_t(ns);
    h_nflags = nflags;
    h_bout0 = bout0;

    // 6. Device arrays.
    Pool p{};
    p.ncap = N;
    p.scap = S;
    auto up = [&](auto *dst, const auto &src) {
        NS_HIP(hipMemcpy(dst, src.data(), sizeof(src[0]) * src.size(), hipMemcpyHostToDevice));
    };
    float *d_ng = arena.alloc<float>(NG_FIELDS * N); up(d_ng, ng); p.ng = d_ng;
    float *d_nx = arena.alloc<float>(NX_FIELDS * N); up(d_nx, nx); p.nx = d_nx;
    uint32_t *d_nflags = arena.alloc<uint32_t>(N); up(d_nflags, nflags); p.nflags = d_nflags;
    int32_t *d_ninput = arena.alloc<int32_t>(N); up(d_ninput, ninput); p.ninput = d_ninput;
    int32_t *d_nin = arena.alloc<int32_t>(N + 1); up(d_nin, nin); p.nin = d_nin;
    int32_t *d_nbrain = arena.alloc<int32_t>(N); up(d_nbrain, nbrain); p.nbrain = d_nbrain;
    p.ns = arena.alloc<float>(NS_FIELDS * N);
    p.cell = arena.alloc<float>(kCellFields * N);
    p.total = arena.alloc<int32_t>(N);
    p.rrng = arena.alloc<uint32_t>(N);
    p.hist = arena.alloc<uint32_t>(N);
    p.spike = arena.alloc<uint8_t>(N);
    p.modbits = arena.alloc<uint8_t>(N);
    p.stamps = arena.alloc<uint64_t>(6 * N);
    float *d_sg = arena.alloc<float>(SG_FIELDS * S); up(d_sg, sg); p.sg = d_sg;
    uint32_t *d_sflags = arena.alloc<uint32_t>(S); up(d_sflags, sflags); p.sflags = d_sflags;
    int32_t *d_spre = arena.alloc<int32_t>(S); up(d_spre, spre); p.spre = d_spre;
    int32_t *d_spost = arena.alloc<int32_t>(S); up(d_spost, spost); p.spost = d_spost;
    int32_t *d_sbrain = arena.alloc<int32_t>(S); up(d_sbrain, sbrain); p.sbrain = d_sbrain;
    p.ss = arena.alloc<float>(SS_FIELDS * S);
    p.last = arena.alloc<int64_t>(S);
    p.ready = arena.alloc<uint64_t>(S);
    p.srng = arena.alloc<uint32_t>(S);
    p.arrived = arena.alloc<uint8_t>(S);
    p.released = arena.alloc<float>(S);
    const int B = std::max(nb + head_slots, 1);
    nb_cap = B;
    int32_t *d_bn0 = arena.alloc<int32_t>(B); up(d_bn0, bn0); p.bn0 = d_bn0;
    int32_t *d_bnc = arena.alloc<int32_t>(B); up(d_bnc, bnc); p.bnc = d_bnc;
    int32_t *d_bs0 = arena.alloc<int32_t>(B); up(d_bs0, bs0); p.bs0 = d_bs0;
    int32_t *d_bsc = arena.alloc<int32_t>(B); up(d_bsc, bsc); p.bsc = d_bsc;
    p.bin0 = nullptr;
    int32_t *d_bout0 = arena.alloc<int32_t>(B); up(d_bout0, bout0); p.bout0 = d_bout0;
    p.btick = arena.alloc<int64_t>(B);
    p.brun = arena.alloc<uint8_t>(B);
    p.bfrozen = arena.alloc<uint8_t>(B);
    p.bstatus = arena.alloc<int32_t>(B);
    int32_t *d_motor = arena.alloc<int32_t>(int64_t(B) * std::max(outs, 1));
    up(d_motor, motor);
    p.motor = d_motor;
    p.outputs = outs;
    p.events = arena.alloc<unsigned long long>(3);
    NS_HIP(hipMemset(p.btick, 0, sizeof(int64_t) * B));
    NS_HIP(hipMemset(p.brun, 0, B));
    NS_HIP(hipMemset(p.bfrozen, 0, B));
    NS_HIP(hipMemset(p.bstatus, 0, sizeof(int32_t) * B));
    NS_HIP(hipMemset(p.events, 0, 3 * sizeof(unsigned long long)));
    NS_HIP(hipMemset(p.hist, 0, sizeof(uint32_t) * N));
    NS_HIP(hipMemset(p.stamps, 0, sizeof(uint64_t) * 6 * N));
    NS_HIP(hipMemset(p.released, 0, sizeof(float) * S));
    NS_HIP(hipMemset(p.modbits, 0, N));
    d_norig = arena.alloc<uint32_t>(N); up(d_norig, norig);
    d_sorig = arena.alloc<uint32_t>(S); up(d_sorig, sorig);
    d_bseed = arena.alloc<uint32_t>(B);
    NS_HIP(hipMemset(d_bseed, 0, sizeof(uint32_t) * B));
    slow_count = int(slow_pos.size());
    d_slow = arena.alloc<int32_t>(std::max(slow_count, 1));
    if (slow_count) up(d_slow, slow_pos);
    view = p;

    // 7. Large slots, packs (with type-sorted neuron slots) and per-tick ranges.
    std::vector<int32_t> big_list(std::max(nbig, 1), 0);
    for (int k = 0; k < nbig; ++k) big_list[k] = fallback_brains + k;
    big_cap = std::max(nbig + head_slots, 1);
    d_big = arena.alloc<int32_t>(big_cap);
    up(d_big, big_list);
    std::vector<uint16_t> h_perm(N, 0);
    packs.clear();
    open_pack = -1;
    int pb = fallback_brains + nbig;
    for (const Bin &bin : bins) {
        Pack pk{};
        pk.b0 = pb;
        pk.nb = bin.nb;
        pk.n0 = bn0[pb];
        pk.s0 = bs0[pb];
        pk.nn = bin.nn;
        pk.ns = bin.ns;
        pk.nm = bin.nb * outs;
        pack_order(pk, &nflags[pk.n0], &h_perm[pk.n0]);
        for (int k = 0; k < bin.nb; ++k) pack_of[pb + k] = int(packs.size());
        packs.push_back(pk);
        pb += bin.nb;
    }
    d_perm = arena.alloc<uint16_t>(N);
    up(d_perm, h_perm);
    packs_cap = std::max(int(packs.size()) + head_slots, 1);
    d_touched = arena.alloc<int32_t>(packs_cap);
    touched.clear();
    d_reset = arena.alloc<int32_t>(2 * int64_t(B));
    d_packs = arena.alloc<Pack>(packs_cap);
    if (!packs.empty()) up(d_packs, packs);
    const int64_t fb_n = fallback_brains ? int64_t(bn0[fallback_brains - 1]) + bnc[fallback_brains - 1] : 0;
    const int64_t fb_s = fallback_brains ? int64_t(bs0[fallback_brains - 1]) + bsc[fallback_brains - 1] : 0;
    fallback_range = PlainRange{0, fb_n, 0, fb_s, nb, fallback_brains * outs, d_slow, slow_count};
    all_range = PlainRange{0, nn, 0, ns, nb, nb * outs, d_slow, slow_count};
}

// ---------------------------------------------------------------------------
// Slot placement
// ---------------------------------------------------------------------------

// A kept one-word buffer for a staged copy.
static const void *keep_i32_once(std::vector<std::vector<unsigned char>> &keep, int32_t v) {
    keep.emplace_back(size_t(4));
    std::memcpy(keep.back().data(), &v, 4);
    return keep.back().data();
}

// Write one newborn's slot image into its reserved chunks (thread-safe across
// different slots: every write goes to the slot's own ranges).
static void fill_slot(DevicePool &d, const PlaceChunks &c, unsigned char *base) {
    const HostBrain *b = c.spec->brain;
    const int pb = c.pb;
    const int64_t n0 = d.bn0[pb], s0 = d.bs0[pb];
    const int cn = d.bnc[pb], cs = d.bsc[pb];
    const int nc = int(b->genes.size()), sc = int(b->edges.size());
    for (int k = 0; k < NG_FIELDS; ++k) {
        float *f = reinterpret_cast<float *>(base + c.ng[k]);
        for (int i = 0; i < nc; ++i) {
            const ns_neuron_gene &g = b->genes[i];
            const float v[NG_FIELDS] = {g.capacitance_pf, g.leak_ns, g.rest_mv, g.threshold_mv,
                                        g.slope_mv, g.adapt_ns, g.adapt_tau_ms, g.spike_adapt_pa,
                                        g.reset_mv, g.afterdepol_spike_pa, g.afterdepol_tau_ms,
                                        g.receptor_baseline_hz, g.receptor_gain_hz,
                                        g.receptor_event_ns, g.receptor_decay_ms};
            f[i] = v[k];
        }
    }
    std::memcpy(base + c.fl, &d.h_nflags[n0], 4u * cn);   // flags set by place()
    int32_t *inp = reinterpret_cast<int32_t *>(base + c.inp);
    for (int i = 0; i < nc; ++i) inp[i] = b->sensory[i] >= 0 ? c.spec->input_base + b->sensory[i] : -1;
    uint32_t *no = reinterpret_cast<uint32_t *>(base + c.no);
    for (int i = 0; i < nc; ++i) no[i] = b->orig_n.empty() ? uint32_t(i) : b->orig_n[i];
    // synapses by (post, edge index): a stable counting sort by post
    std::vector<int> start(nc + 1, 0), order(sc);
    for (int j = 0; j < sc; ++j) start[b->edges[j].post + 1]++;
    for (int i = 0; i < nc; ++i) start[i + 1] += start[i];
    int32_t *ni = reinterpret_cast<int32_t *>(base + c.ni);
    for (int i = 0; i < c.ninx; ++i) ni[i] = int32_t(s0 + start[i]);   // the first tail neuron bounds the last run
    for (int j = 0; j < sc; ++j) order[start[b->edges[j].post]++] = j;
    for (int m = 0; m < SG_FIELDS; ++m) {
        float *f = reinterpret_cast<float *>(base + c.sg[m]);
        for (int k = 0; k < sc; ++k) {
            const ns_synapse_gene &g = b->edges[order[k]];
            const float v[SG_FIELDS] = {g.initial_weight_ns, g.utilization, g.recovery_ms,
                                        g.facilitation_ms, g.slow_recovery_ms, g.slow_depletion};
            f[k] = v[m];
        }
    }
    uint32_t *sf = reinterpret_cast<uint32_t *>(base + c.sf);
    for (int k = 0; k < cs; ++k) {
        if (k >= sc) { sf[k] = 1u; d.syn_perm[s0 + k] = -1; continue; }   // invalid tail, delay 1
        const int j = order[k];
        const ns_synapse_gene &g = b->edges[j];
        const ns_neuron_gene &a = b->genes[g.pre], &z = b->genes[g.post];
        uint32_t route = uint32_t(g.branch & 3);
        if (a.modulator) route |= uint32_t(a.modulator & 3) << R_MOD_SHIFT;
        else if (a.inhibitory_transmitter) route |= R_INH;
        if (g.active && a.active && z.active) route |= R_VALID;
        if (!a.inhibitory_transmitter || a.modulator) route |= R_STP;
        if (!a.inhibitory_transmitter && excitatory_cell(a.type) && !a.modulator &&
            !z.inhibitory_transmitter && excitatory_cell(z.type))
            route |= R_PAIR;
        sf[k] = uint32_t(g.delay) | (uint32_t(g.plasticity & 7) << SF_PLAST_SHIFT) |
                (uint32_t(g.mod_receptors & 7) << SF_MODREC_SHIFT) |
                (uint32_t(site_count(g)) << SF_SITES_SHIFT) | (route << SF_ROUTE_SHIFT);
        d.syn_perm[s0 + k] = j;
    }
    int32_t *pre = reinterpret_cast<int32_t *>(base + c.pre);
    int32_t *post = reinterpret_cast<int32_t *>(base + c.post);
    uint32_t *so = reinterpret_cast<uint32_t *>(base + c.so);
    for (int k = 0; k < sc; ++k) {
        const ns_synapse_gene &g = b->edges[order[k]];
        pre[k] = int32_t(n0 + g.pre);
        post[k] = int32_t(n0 + g.post);
        so[k] = b->orig_s.empty() ? uint32_t(order[k]) : b->orig_s[order[k]];
    }
    int32_t *mo = reinterpret_cast<int32_t *>(base + c.mo);
    for (int j = 0; j < d.outputs; ++j) mo[j] = b->motors[j] >= 0 ? int32_t(n0 + b->motors[j]) : -1;
}

HostWorkers &HostWorkers::get() {
    static HostWorkers workers;
    return workers;
}

HostWorkers::HostWorkers() {
    // the host share of one GPU is about 16 CPUs; the caller works too
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int count = int(std::min(15u, hw > 1 ? hw - 1 : 1u));
    for (int i = 0; i < count; ++i) threads_.emplace_back([this] { loop(); });
}

HostWorkers::~HostWorkers() {
    {
        std::lock_guard<std::mutex> g(m_);
        stop_ = true;
    }
    wake_.notify_all();
    for (auto &th : threads_) th.join();
}

void HostWorkers::loop() {
    uint64_t seen = 0;
    for (;;) {
        const std::function<void(int)> *fn;
        int n;
        {
            std::unique_lock<std::mutex> g(m_);
            wake_.wait(g, [&] { return stop_ || round_ != seen; });
            if (stop_) return;
            seen = round_;
            fn = fn_;
            n = n_;
            ++busy_;
        }
        std::exception_ptr err;
        try {
            for (int i = next_.fetch_add(1); i < n; i = next_.fetch_add(1)) (*fn)(i);
        } catch (...) {
            err = std::current_exception();
            next_.store(n);   // the others stop taking indices
        }
        {
            std::lock_guard<std::mutex> g(m_);
            if (err && !error_) error_ = err;
            if (--busy_ == 0) done_.notify_all();
        }
    }
}

void HostWorkers::run(int n, const std::function<void(int)> &fn) {
    if (n <= 0) return;
    if (n < 8 || threads_.empty()) {
        for (int i = 0; i < n; ++i) fn(i);
        return;
    }
    std::lock_guard<std::mutex> one(run_lock_);
    {
        std::lock_guard<std::mutex> g(m_);
        fn_ = &fn;
        n_ = n;
        next_.store(0);
        ++round_;
    }
    wake_.notify_all();
    std::exception_ptr err;
    try {
        for (int i = next_.fetch_add(1); i < n; i = next_.fetch_add(1)) fn(i);
    } catch (...) {
        err = std::current_exception();
        next_.store(n);
    }
    // every index is taken; wait until the workers that joined this round
    // have finished theirs (a worker that wakes late finds none left)
    std::unique_lock<std::mutex> g(m_);
    done_.wait(g, [&] { return busy_ == 0; });
    fn_ = nullptr;
    if (!err) err = error_;
    error_ = nullptr;
    g.unlock();
    if (err) std::rethrow_exception(err);
}

GenomeReaper &GenomeReaper::get() {
    static GenomeReaper reaper;
    return reaper;
}

GenomeReaper::GenomeReaper() : thread_([this] { loop(); }) {}

GenomeReaper::~GenomeReaper() {
    {
        std::lock_guard<std::mutex> g(m_);
        stop_ = true;
    }
    wake_.notify_all();
    thread_.join();
}

void GenomeReaper::retire(std::vector<std::shared_ptr<const HostBrain>> &&genomes) {
    if (genomes.empty()) return;
    {
        std::lock_guard<std::mutex> g(m_);
        for (auto &p : genomes) pending_.push_back(std::move(p));
    }
    genomes.clear();
    wake_.notify_one();
}

void GenomeReaper::loop() {
    std::vector<std::shared_ptr<const HostBrain>> batch;
    for (;;) {
        {
            std::unique_lock<std::mutex> g(m_);
            wake_.wait(g, [&] { return stop_ || !pending_.empty(); });
            if (pending_.empty()) return;   // stopping, nothing left
            batch.swap(pending_);
        }
        batch.clear();   // the last references: the genomes are freed here
    }
}

void DevicePool::upload_stage(const Stager &sg, hipStream_t st) {
    const size_t bytes = sg.bytes();
    PinnedStaging::Buf &buf = staging.take(bytes);
    unsigned char *base = sg.write(buf.p);
    // slot images, in parallel over newborns (disjoint slots)
    HostWorkers::get().run(int(fill_list.size()),
                           [&](int j) { fill_slot(*this, fill_list[size_t(j)], base); });
    fill_list.clear();
    if (!copy_stream) {
        NS_HIP(hipStreamCreateWithFlags(&copy_stream, hipStreamNonBlocking));
        for (auto &d : dstage) {
            NS_HIP(hipEventCreateWithFlags(&d.uploaded, hipEventDisableTiming));
            NS_HIP(hipEventCreateWithFlags(&d.read, hipEventDisableTiming));
            NS_HIP(hipEventRecord(d.read, copy_stream));   // a first wait finds it complete
        }
    }
    stage_cur = stage_next;
    stage_next ^= 1;
    DeviceStage &d = dstage[stage_cur];
    if (d.cap < bytes) {
        ++dstage_grows;
        NS_HIP(hipEventSynchronize(d.read));   // no queued scatter reads the old buffer
        d.arena.release();
        d.cap = std::max(2 * bytes, kStageMinBytes);
        d.p = d.arena.alloc<unsigned char>(int64_t(d.cap));
    }
    NS_HIP(hipStreamWaitEvent(copy_stream, d.read, 0));
    NS_HIP(hipMemcpyAsync(d.p, buf.p, bytes, hipMemcpyHostToDevice, copy_stream));
    NS_HIP(hipEventRecord(buf.done, copy_stream));
    NS_HIP(hipEventRecord(d.uploaded, copy_stream));
    NS_HIP(hipStreamWaitEvent(st, d.uploaded, 0));
}

void DevicePool::scatter_range(int lo, int hi, hipStream_t st) {
    if (hi <= lo) return;
    DeviceStage &d = dstage[stage_cur];
    launch_scatter(reinterpret_cast<const ScatterEntry *>(d.p) + lo, hi - lo, d.p, st);
    NS_HIP(hipGetLastError());
    NS_HIP(hipEventRecord(d.read, st));
}

DevicePool::~DevicePool() {
    if (copy_stream) (void)hipStreamSynchronize(copy_stream);
    for (auto &d : dstage) {
        if (d.read) {
            (void)hipEventSynchronize(d.read);
            (void)hipEventDestroy(d.read);
        }
        if (d.uploaded) (void)hipEventDestroy(d.uploaded);
    }
    if (copy_stream) (void)hipStreamDestroy(copy_stream);
    arena.release();
}

void DevicePool::place(const std::vector<std::pair<int, const BrainSpec *>> &moves, Stager &sg) {
    if (moves.empty()) return;
    const int64_t N = view.ncap, S = view.scap;
    touched.clear();
    std::vector<PlaceChunks> chunks;
    auto off = [&](const void *dst, size_t bytes) { return sg.alloc(dst, bytes); };
    // 1. Slot tables and chunk reservations, in order.
    for (const auto &mv : moves) {
        const int pb = mv.first;
        const BrainSpec *spec = mv.second;
        // per-tick-class slots may be vacated, not filled (their LP/Prinz
        // neuron list is fixed at build)
        if (pb < 0 || pb >= nb || (slot_kind[pb] == 0 && spec && spec->brain))
            throw std::runtime_error("Slot cannot take a placement");
        const HostBrain *b = spec ? spec->brain : nullptr;
        if (batch_of[pb] >= 0 && pool_of[batch_of[pb]] == pb) pool_of[batch_of[pb]] = -1;
        if (!b) {
            // A vacated slot never runs (its run flag stays clear), so its
            // old image is left in place until a newborn overwrites it.
            occ_n[pb] = 0;
            occ_s[pb] = 0;
            batch_of[pb] = -1;
            continue;
        }
        const int64_t n0 = bn0[pb], s0 = bs0[pb];
        const int cn = bnc[pb], cs = bsc[pb];
        const int nc = int(b->genes.size()), sc = int(b->edges.size());
        if (nc > cn || sc > cs) throw std::runtime_error("Brain exceeds its slot");
        PlaceChunks c;
        c.pb = pb;
        c.spec = spec;
        c.ninx = std::min(cn, nc + 1);
        // The extra LP/Prinz genes are read only by per-tick-class slots,
        // which never take placements; a tail past the occupant needs only
        // its flags (inactive neurons, invalid synapses).
        for (int k = 0; k < NG_FIELDS; ++k) c.ng[k] = off(view.ng + k * N + n0, 4u * nc);
        c.fl = off(view.nflags + n0, 4u * cn);
        c.inp = off(view.ninput + n0, 4u * nc);
        c.no = off(d_norig + n0, 4u * nc);
        c.ni = off(view.nin + n0, 4u * c.ninx);
        for (int m = 0; m < SG_FIELDS; ++m) c.sg[m] = off(view.sg + m * S + s0, 4u * sc);
        c.sf = off(view.sflags + s0, 4u * cs);
        c.pre = off(view.spre + s0, 4u * sc);
        c.post = off(view.spost + s0, 4u * sc);
        c.so = off(d_sorig + s0, 4u * sc);
        c.mo = off(view.motor + int64_t(pb) * outputs, 4u * outputs);
        sg.add(view.bout0 + pb, &spec->output_base, 4);
        h_bout0[pb] = spec->output_base;
        occ_n[pb] = nc;
        occ_s[pb] = sc;
        batch_of[pb] = spec->batch_index;
        pool_of[spec->batch_index] = pb;
        if (slot_kind[pb] == 2) touched.push_back(pack_of[pb]);
        for (int i = 0; i < cn; ++i) {
            if (i < nc) {
                const ns_neuron_gene &g = b->genes[i];
                h_nflags[n0 + i] = uint32_t(g.type) | (g.active ? NF_ACTIVE : 0) |
                                   (b->receptor[i] ? NF_RECEPTOR : 0) | (spec->world ? world_flags(b->max_delay) : 0);
            } else {
                h_nflags[n0 + i] = 0;   // inactive tail
            }
        }
        chunks.push_back(c);
    }
    fill_list.insert(fill_list.end(), chunks.begin(), chunks.end());
    // 2. The packs whose slot orders must be recomputed (by order_touched,
    // on the device, once the images are in place).  The images themselves
    // are written by upload_stage, straight into pinned memory.
    std::sort(touched.begin(), touched.end());
    touched.erase(std::unique(touched.begin(), touched.end()), touched.end());
    if (!touched.empty()) sg.add(d_touched, touched.data(), 4u * touched.size());
}

bool DevicePool::append_slots(const std::vector<std::pair<int, int>> &caps,
                              const std::vector<std::pair<int, int>> &occ, Stager &sg,
                              std::vector<std::vector<unsigned char>> &keep, std::vector<int> &slots) {
    slots.clear();
    if (!growable || caps.empty()) return caps.empty();
    const int64_t N = view.ncap, S = view.scap;
    const int nmax = tile.nmax(), smax = tile.ks > 0 ? tile.threads * tile.ks : smax_lds;
    const int brains_per_pack = std::min(kPackBrains, std::max(1, kPackMotors / std::max(1, outputs)));
    // 1. Room: every appended slot may need a pack or an own-workgroup entry.
    int64_t need_n = 0, need_s = 0;
    for (const auto &c : caps) { need_n += c.first; need_s += c.second; }
    const int k = int(caps.size());
    if (nn + need_n > N || ns + need_s > S || nb + k > nb_cap || int(packs.size()) + k > packs_cap ||
        nbig + k > big_cap)
        return false;
    // 2. Slot tables; packable slots join the last appended pack while it has
    // room and still ends the pool (no own-workgroup slot appended after it).
    int open = open_pack >= 0 && open_pack == int(packs.size()) - 1 &&
                       packs[open_pack].b0 + packs[open_pack].nb == nb
                   ? open_pack : -1;
    std::vector<int> new_packs, new_big;
    if (open >= 0) new_packs.push_back(open);  // re-staged with its new extent
    for (int j = 0; j < k; ++j) {
        const int cn = caps[j].first, cs = caps[j].second;
        const int pb = nb++;
        const int64_t n0 = nn, s0 = ns;
        nn += cn;
        ns += cs;
        bn0.push_back(int32_t(n0)); bnc.push_back(cn);
        bs0.push_back(int32_t(s0)); bsc.push_back(cs);
        batch_of.push_back(-1); fallback.push_back(0);
        occ_n.push_back(0); occ_s.push_back(0);
        h_bout0.push_back(0);
        const bool packable = cn <= nmax && cs <= smax && outputs <= kPackMotors;
        slot_kind.push_back(uint8_t(packable ? 2 : 1));
        pack_of.push_back(-1);
        if (packable) {
            if (open >= 0 && packs[open].b0 + packs[open].nb == pb && packs[open].nn + cn <= nmax &&
                packs[open].ns + cs <= smax && packs[open].nb < brains_per_pack) {
                Pack &pk = packs[open];
                pk.nn += cn; pk.ns += cs; pk.nb += 1; pk.nm = pk.nb * outputs;
            } else {
                Pack pk{};
                pk.b0 = pb; pk.nb = 1; pk.n0 = int32_t(n0); pk.s0 = int32_t(s0);
                pk.nn = cn; pk.ns = cs; pk.nm = outputs;
                open = int(packs.size());
                packs.push_back(pk);
                new_packs.push_back(open);
            }
            pack_of[pb] = open;
        } else {
            new_big.push_back(pb);
        }
        for (int i = 0; i < cn; ++i) h_nflags[n0 + i] = 0;
        slots.push_back(pb);
        // 3. The image parts place() does not write: the slot's brain index on
        // every neuron and synapse, the neuron runs and endpoints past the
        // newborn (an empty run at the slot's end, the slot's first neuron),
        // an empty input and origin past it.
        const int nc = std::min(occ[j].first, cn), sc = std::min(occ[j].second, cs);
        auto keep_i32 = [&](int64_t count, int32_t v) {
            keep.emplace_back(size_t(4) * size_t(std::max<int64_t>(count, 0)));
            int32_t *x = reinterpret_cast<int32_t *>(keep.back().data());
            for (int64_t i = 0; i < count; ++i) x[i] = v;
            return static_cast<const void *>(x);
        };
        sg.add(view.nbrain + n0, keep_i32(cn, pb), 4u * cn);
        sg.add(view.sbrain + s0, keep_i32(cs, pb), 4u * cs);
        if (cn > nc + 1) sg.add(view.nin + n0 + nc + 1, keep_i32(cn - nc - 1, int32_t(s0 + sc)), 4u * (cn - nc - 1));
        if (cn > nc) {
            sg.add(view.ninput + n0 + nc, keep_i32(cn - nc, -1), 4u * (cn - nc));
            sg.add(d_norig + n0 + nc, keep_i32(cn - nc, 0), 4u * (cn - nc));
        }
        if (cs > sc) {
            sg.add(view.spre + s0 + sc, keep_i32(cs - sc, int32_t(n0)), 4u * (cs - sc));
            sg.add(view.spost + s0 + sc, keep_i32(cs - sc, int32_t(n0)), 4u * (cs - sc));
            sg.add(d_sorig + s0 + sc, keep_i32(cs - sc, 0), 4u * (cs - sc));
        }
        const int32_t tab[4] = {int32_t(n0), cn, int32_t(s0), cs};
        const void *const dst[4] = {view.bn0 + pb, view.bnc + pb, view.bs0 + pb, view.bsc + pb};
        for (int f = 0; f < 4; ++f) sg.add(dst[f], keep_i32(1, tab[f]), 4);
        keep.emplace_back(size_t(1), (unsigned char)0);
        sg.add(view.bfrozen + pb, keep.back().data(), 1);
        for (int64_t q = s0; q < s0 + cs; ++q) syn_perm[q] = -1;
    }
    // 4. Pack and own-workgroup entries (packs opened here, in full).
    for (int idx : new_packs) {
        keep.emplace_back(sizeof(Pack));
        std::memcpy(keep.back().data(), &packs[idx], sizeof(Pack));
        sg.add(d_packs + idx, keep.back().data(), sizeof(Pack));
    }
    for (int pb : new_big) {
        sg.add(d_big + nbig, keep_i32_once(keep, pb), 4);
        ++nbig;
    }
    open_pack = open;
    all_range.n_hi = nn;
    all_range.s_hi = ns;
    all_range.brains = nb;
    all_range.slots = nb * outputs;
    appended += k;
    ++generation;                              // launch sizes changed: graphs capture again
    return true;
}

void DevicePool::order_touched(hipStream_t st) {
    launch_pack_order(view.nflags, view.modbits, d_packs, d_touched, int(touched.size()), d_perm, st);
    NS_HIP(hipGetLastError());
    touched.clear();
}

void DevicePool::stage_state(int pb, const HostBrain &brain, const ns_neuron_state *in_n,
                             const uint32_t *hist, const uint64_t *stamps,
                             const ns_synapse_state *in_s, int64_t tick, Stager &sg,
                             std::vector<std::vector<unsigned char>> &keep) {
    const int64_t N = view.ncap, S = view.scap, n0 = bn0[pb], s0 = bs0[pb];
    const int nc = occ_n[pb], sc = occ_s[pb];
    if (nc != int(brain.genes.size()) || sc != int(brain.edges.size()))
        throw std::runtime_error("Slot occupant differs");
    // the slot's synapse order, as fill_slot lays it out: by (post, edge index)
    std::vector<int> start(nc + 1, 0), order(sc);
    for (int j = 0; j < sc; ++j) start[brain.edges[j].post + 1]++;
    for (int i = 0; i < nc; ++i) start[i + 1] += start[i];
    for (int j = 0; j < sc; ++j) order[start[brain.edges[j].post]++] = j;
    // one host buffer per field, staged as one range of the slot
    auto field = [&](void *dst_base, int64_t index, size_t elem, int count, auto fill) {
        if (!count) return;
        keep.emplace_back(elem * size_t(count));
        unsigned char *buf = keep.back().data();
        fill(buf);
        sg.add(static_cast<unsigned char *>(dst_base) + elem * size_t(index), buf, elem * size_t(count));
    };
    auto nsf = [&](int k, auto get) {
        field(view.ns, k * N + n0, 4, nc, [&](unsigned char *b) {
            float *f = reinterpret_cast<float *>(b);
            for (int i = 0; i < nc; ++i) f[i] = get(in_n[i]);
        });
    };
    nsf(NS_V, [](const ns_neuron_state &r) { return r.voltage_mv; });
    nsf(NS_W, [](const ns_neuron_state &r) { return r.adaptation_pa; });
    nsf(NS_ADP, [](const ns_neuron_state &r) { return r.afterdepol_pa; });
    nsf(NS_INH, [](const ns_neuron_state &r) { return r.inh_ns; });
    nsf(NS_DEPOL, [](const ns_neuron_state &r) { return r.recent_depol_mv; });
    nsf(NS_LAS, [](const ns_neuron_state &r) { return r.log_ampa_scale; });
    for (int d = 0; d < 4; ++d)
        nsf(NS_D0 + d, [d](const ns_neuron_state &r) { return r.dendrite_exc_ns[d]; });
    nsf(NS_RNS, [](const ns_neuron_state &r) { return r.receptor_ns; });
    for (int c = 0; c < kCellFields; ++c)
        field(view.cell, c * N + n0, 4, nc, [&](unsigned char *b) {
            float *f = reinterpret_cast<float *>(b);
            for (int i = 0; i < nc; ++i) f[i] = in_n[i].cell[c];
        });
    field(view.total, n0, 4, nc, [&](unsigned char *b) {
        int32_t *f = reinterpret_cast<int32_t *>(b);
        for (int i = 0; i < nc; ++i) f[i] = in_n[i].total_spikes;
    });
    field(view.spike, n0, 1, nc, [&](unsigned char *b) {
        for (int i = 0; i < nc; ++i) b[i] = in_n[i].spike;
    });
    field(view.rrng, n0, 4, nc, [&](unsigned char *b) {
        uint32_t *f = reinterpret_cast<uint32_t *>(b);
        for (int i = 0; i < nc; ++i) f[i] = in_n[i].receptor_rng;
    });
    field(view.hist, n0, 4, nc, [&](unsigned char *b) { std::memcpy(b, hist, 4u * nc); });
    field(view.modbits, n0, 1, nc, [&](unsigned char *b) { std::memset(b, 0, size_t(nc)); });
    for (int e = 0; e < 6; ++e)
        field(view.stamps, e * N + n0, 8, nc, [&](unsigned char *b) {
            uint64_t *f = reinterpret_cast<uint64_t *>(b);
            for (int i = 0; i < nc; ++i) f[i] = stamps ? stamps[i * 6 + e] : 0;
        });
    // synapses: pool position q holds genome edge order[q]
    auto ssf = [&](int k, auto get) {
        field(view.ss, k * S + s0, 4, sc, [&](unsigned char *b) {
            float *f = reinterpret_cast<float *>(b);
            for (int q = 0; q < sc; ++q) f[q] = get(in_s[order[q]]);
        });
    };
    ssf(SS_WEIGHT, [](const ns_synapse_state &r) { return r.weight_ns; });
    ssf(SS_UTIL, [](const ns_synapse_state &r) { return r.running_utilization; });
    ssf(SS_PTRACE, [](const ns_synapse_state &r) { return r.pre_trace; });
    ssf(SS_QTRACE, [](const ns_synapse_state &r) { return r.post_trace; });
    ssf(SS_ELTP, [](const ns_synapse_state &r) { return r.eligible_ltp; });
    ssf(SS_ELTD, [](const ns_synapse_state &r) { return r.eligible_ltd; });
    ssf(SS_DA_AGE, [](const ns_synapse_state &r) { return r.da_age_ms; });
    ssf(SS_SPINE, [](const ns_synapse_state &r) { return r.spine_volume; });
    ssf(SS_SAGE, [](const ns_synapse_state &r) { return r.structural_age_ms; });
    ssf(SS_PRE_AGE, [](const ns_synapse_state &r) { return r.pre_age_ms; });
    ssf(SS_POST_AGE, [](const ns_synapse_state &r) { return r.post_age_ms; });
    ssf(SS_SLOW, [](const ns_synapse_state &r) { return r.slow_resource; });
    field(view.last, s0, 8, sc, [&](unsigned char *b) {
        int64_t *f = reinterpret_cast<int64_t *>(b);
        for (int q = 0; q < sc; ++q) f[q] = in_s[order[q]].last_arrival_tick;
    });
    field(view.ready, s0, 8, sc, [&](unsigned char *b) {
        uint64_t *f = reinterpret_cast<uint64_t *>(b);
        for (int q = 0; q < sc; ++q) f[q] = in_s[order[q]].ready_mask;
    });
    field(view.srng, s0, 4, sc, [&](unsigned char *b) {
        uint32_t *f = reinterpret_cast<uint32_t *>(b);
        for (int q = 0; q < sc; ++q) f[q] = in_s[order[q]].rng_state;
    });
    field(view.arrived, s0, 1, sc, [&](unsigned char *b) {
        for (int q = 0; q < sc; ++q) b[q] = in_s[order[q]].arrived;
    });
    field(view.released, s0, 4, sc, [&](unsigned char *b) { std::memset(b, 0, 4u * size_t(sc)); });
    field(view.btick, pb, 8, 1, [&](unsigned char *b) { std::memcpy(b, &tick, 8); });
}

// Neuron/synapse state in batch order, genome order within each brain.
void DevicePool::download_state(std::vector<ns_neuron_state> &out_n,
                                std::vector<ns_synapse_state> &out_s, std::vector<uint32_t> &hist,
                                std::vector<int64_t> &ticks, std::vector<uint64_t> *stamps) {
    const int64_t N = view.ncap, S = view.scap;
    std::vector<float> fns(NS_FIELDS * N), cell(kCellFields * N), fss(SS_FIELDS * S);
    std::vector<int32_t> total(N);
    std::vector<uint32_t> rrng(N), h(N), srng(S);
    std::vector<uint8_t> spike(N), arrived(S);
    std::vector<int64_t> last(S), btick(std::max(nb, 1));
    std::vector<uint64_t> ready(S), st(6 * N);
    auto down = [&](auto &dst, const auto *src) {
        NS_HIP(hipMemcpy(dst.data(), src, sizeof(dst[0]) * dst.size(), hipMemcpyDeviceToHost));
    };
    NS_HIP(hipStreamSynchronize(stream));
    down(fns, view.ns); down(cell, view.cell); down(fss, view.ss); down(total, view.total);
    down(rrng, view.rrng); down(h, view.hist); down(srng, view.srng); down(spike, view.spike);
    down(arrived, view.arrived); down(last, view.last); down(btick, view.btick);
    down(ready, view.ready);
    if (stamps) down(st, view.stamps);
    // batch-order offsets
    std::vector<int64_t> noff(nbatch + 1, 0), soff(nbatch + 1, 0);
    for (int i = 0; i < nbatch; ++i) {   // a batch index without a slot holds no brain
        noff[i + 1] = noff[i] + (pool_of[i] >= 0 ? occ_n[pool_of[i]] : 0);
        soff[i + 1] = soff[i] + (pool_of[i] >= 0 ? occ_s[pool_of[i]] : 0);
    }
    out_n.assign(noff[nbatch], ns_neuron_state{});
    out_s.assign(soff[nbatch], ns_synapse_state{});
    hist.assign(noff[nbatch], 0);
    ticks.assign(nbatch, 0);
    if (stamps) stamps->assign(noff[nbatch] * 6, 0);
    for (int i = 0; i < nbatch; ++i) {
        const int pb = pool_of[i];
        if (pb < 0) continue;
        ticks[i] = btick[pb];
        for (int k = 0; k < occ_n[pb]; ++k) {
            const int64_t n = int64_t(bn0[pb]) + k;
            ns_neuron_state &r = out_n[noff[i] + k];
            std::memset(&r, 0, sizeof r);
            r.voltage_mv = fns[NS_V * N + n];
            r.adaptation_pa = fns[NS_W * N + n];
            r.afterdepol_pa = fns[NS_ADP * N + n];
            r.inh_ns = fns[NS_INH * N + n];
            r.recent_depol_mv = fns[NS_DEPOL * N + n];
            r.log_ampa_scale = fns[NS_LAS * N + n];
            for (int d = 0; d < 4; ++d) r.dendrite_exc_ns[d] = fns[(NS_D0 + d) * N + n];
            for (int c = 0; c < kCellFields; ++c) r.cell[c] = cell[c * N + n];
            r.total_spikes = total[n];
            r.spike = spike[n];
            r.receptor_ns = fns[NS_RNS * N + n];
            r.receptor_rng = rrng[n];
            hist[noff[i] + k] = h[n];
            if (stamps)
                for (int e = 0; e < 6; ++e) (*stamps)[(noff[i] + k) * 6 + e] = st[e * N + n];
        }
        for (int k = 0; k < occ_s[pb]; ++k) {
            const int64_t q = int64_t(bs0[pb]) + k;
            ns_synapse_state &r = out_s[soff[i] + syn_perm[q]];
            std::memset(&r, 0, sizeof r);
            r.weight_ns = fss[SS_WEIGHT * S + q];
            r.running_utilization = fss[SS_UTIL * S + q];
            r.pre_trace = fss[SS_PTRACE * S + q];
            r.post_trace = fss[SS_QTRACE * S + q];
            r.eligible_ltp = fss[SS_ELTP * S + q];
            r.eligible_ltd = fss[SS_ELTD * S + q];
            r.da_age_ms = fss[SS_DA_AGE * S + q];
            r.spine_volume = fss[SS_SPINE * S + q];
            r.structural_age_ms = fss[SS_SAGE * S + q];
            r.pre_age_ms = fss[SS_PRE_AGE * S + q];
            r.post_age_ms = fss[SS_POST_AGE * S + q];
            r.last_arrival_tick = last[q];
            r.ready_mask = ready[q];
            r.rng_state = srng[q];
            r.arrived = arrived[q];
            r.slow_resource = fss[SS_SLOW * S + q];
        }
    }
}

void DevicePool::upload_state(const std::vector<ns_neuron_state> &in_n,
                              const std::vector<ns_synapse_state> &in_s,
                              const std::vector<uint32_t> &hist, const std::vector<int64_t> &ticks,
                              const std::vector<uint64_t> *stamps) {
    const int64_t N = view.ncap, S = view.scap;
    std::vector<float> fns(NS_FIELDS * N, 0.0f), cell(kCellFields * N, 0.0f), fss(SS_FIELDS * S, 0.0f);
    std::vector<int32_t> total(N, 0);
    std::vector<uint32_t> rrng(N, 0), h(N, 0), srng(S, 0);
    std::vector<uint8_t> spike(N, 0), arrived(S, 0);
    std::vector<int64_t> last(S, 0), btick(std::max(nb, 1), 0);
    std::vector<uint64_t> ready(S, 0), st(6 * N, 0);
    std::vector<int64_t> noff(nbatch + 1, 0), soff(nbatch + 1, 0);
    for (int i = 0; i < nbatch; ++i) {   // a batch index without a slot holds no brain
        noff[i + 1] = noff[i] + (pool_of[i] >= 0 ? occ_n[pool_of[i]] : 0);
        soff[i + 1] = soff[i] + (pool_of[i] >= 0 ? occ_s[pool_of[i]] : 0);
    }
    if (int64_t(in_n.size()) != noff[nbatch] || int64_t(in_s.size()) != soff[nbatch] ||
        int64_t(hist.size()) != noff[nbatch] || int(ticks.size()) != nbatch)
        throw std::runtime_error("State dimensions differ");
    for (int i = 0; i < nbatch; ++i) {
        const int pb = pool_of[i];
        if (pb < 0) continue;
        btick[pb] = ticks[i];
        for (int k = 0; k < occ_n[pb]; ++k) {
            const int64_t n = int64_t(bn0[pb]) + k;
            const ns_neuron_state &r = in_n[noff[i] + k];
            fns[NS_V * N + n] = r.voltage_mv;
            fns[NS_W * N + n] = r.adaptation_pa;
            fns[NS_ADP * N + n] = r.afterdepol_pa;
            fns[NS_INH * N + n] = r.inh_ns;
            fns[NS_DEPOL * N + n] = r.recent_depol_mv;
            fns[NS_LAS * N + n] = r.log_ampa_scale;
            for (int d = 0; d < 4; ++d) fns[(NS_D0 + d) * N + n] = r.dendrite_exc_ns[d];
            for (int c = 0; c < kCellFields; ++c) cell[c * N + n] = r.cell[c];
            total[n] = r.total_spikes;
            spike[n] = r.spike;
            fns[NS_RNS * N + n] = r.receptor_ns;
            rrng[n] = r.receptor_rng;
            h[n] = hist[noff[i] + k];
            if (stamps)
                for (int e = 0; e < 6; ++e) st[e * N + n] = (*stamps)[(noff[i] + k) * 6 + e];
        }
        for (int k = 0; k < occ_s[pb]; ++k) {
            const int64_t q = int64_t(bs0[pb]) + k;
            const ns_synapse_state &r = in_s[soff[i] + syn_perm[q]];
            fss[SS_WEIGHT * S + q] = r.weight_ns;
            fss[SS_UTIL * S + q] = r.running_utilization;
            fss[SS_PTRACE * S + q] = r.pre_trace;
            fss[SS_QTRACE * S + q] = r.post_trace;
            fss[SS_ELTP * S + q] = r.eligible_ltp;
            fss[SS_ELTD * S + q] = r.eligible_ltd;
            fss[SS_DA_AGE * S + q] = r.da_age_ms;
            fss[SS_SPINE * S + q] = r.spine_volume;
            fss[SS_SAGE * S + q] = r.structural_age_ms;
            fss[SS_PRE_AGE * S + q] = r.pre_age_ms;
            fss[SS_POST_AGE * S + q] = r.post_age_ms;
            last[q] = r.last_arrival_tick;
            ready[q] = r.ready_mask;
            srng[q] = r.rng_state;
            arrived[q] = r.arrived;
            fss[SS_SLOW * S + q] = r.slow_resource;
        }
    }
    auto up = [&](auto *dst, const auto &src) {
        NS_HIP(hipMemcpy(dst, src.data(), sizeof(src[0]) * src.size(), hipMemcpyHostToDevice));
    };
    NS_HIP(hipStreamSynchronize(stream));
    up(view.ns, fns); up(view.cell, cell); up(view.ss, fss); up(view.total, total);
    up(view.rrng, rrng); up(view.hist, h); up(view.srng, srng); up(view.spike, spike);
    up(view.arrived, arrived); up(view.last, last); up(view.ready, ready);
    NS_HIP(hipMemcpy(view.btick, btick.data(), sizeof(int64_t) * nb, hipMemcpyHostToDevice));
    if (stamps) up(view.stamps, st);
    // the modulator bytes are per-tick outputs; cleared, they carry no
    // settled mark over to the new state
    NS_HIP(hipMemset(view.modbits, 0, size_t(N)));
}

// ---------------------------------------------------------------------------
// Batch
// ---------------------------------------------------------------------------
Tile default_tile(const std::vector<const HostBrain *> &brains) {
    // Pick the smallest standard tile whose pack holds the typical (90th
    // percentile) brain; ns_gpu_autotune refines this by measurement.
    std::vector<int> n, s;
    for (const HostBrain *b : brains) {
        n.push_back(int(b->genes.size()));
        s.push_back(int(b->edges.size()));
    }
    std::sort(n.begin(), n.end());
    std::sort(s.begin(), s.end());
    const int n90 = n[n.size() * 9 / 10], s90 = s[s.size() * 9 / 10];
    for (int k = 0; k < kNumTiles; ++k) {
        const TileShape &t = kTiles[k];
        if (t.synapses_per_thread > 0 && t.threads * t.neurons_per_thread >= 4 * n90 &&
            t.threads * t.synapses_per_thread >= 4 * s90)
            return Tile{t.threads, t.neurons_per_thread, t.synapses_per_thread, t.min_waves, 0};
    }
    return Tile{1024, 1, 8, 1, 0};
}

void Batch::create(std::vector<std::shared_ptr<const HostBrain>> &&genomes, int) {
    brains = std::move(genomes);
    count = int(brains.size());
    channels = brains[0]->channels;
    outputs = int(brains[0]->motors.size());
    for (const auto &b : brains)
        if (b->channels != channels || int(b->motors.size()) != outputs)
            throw std::runtime_error("Batch ports differ");
    NS_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    specs.clear();
    for (int i = 0; i < count; ++i)
        specs.push_back(BrainSpec{brains[i].get(), i, i * channels, i * outputs, world});
    std::vector<const HostBrain *> views;
    for (const auto &b : brains) views.push_back(b.get());
    const Tile t0 = default_tile(views);
    pool.growable = true;                    // room for slots appended by replacements
    pool.build(slot_specs(t0), outputs, t0, stream);
    d_input = io.alloc<float>(int64_t(count) * channels);
    d_output = io.alloc<float>(int64_t(count) * outputs);
    last_mask.assign(count, 1);
}

// Capacity rungs ceil(16 * r^k), r = 1 + slack: brains of nearby sizes share
// a rung, so a freed slot fits most brains of its size class.
static int ladder(int n, float slack) {
    if (!(slack > 0) || n <= 0) return n;
    const double r = 1.0 + slack;
    double c = 16.0;
    while (std::ceil(c) < n) c *= c < 1024.0 ? r : r * r;   // coarser rungs for large brains
    return int(std::ceil(c));
}

std::vector<BrainSpec> reserve_slots(const std::vector<BrainSpec> &specs, float slack, float spare,
                                     const Tile &t, bool world) {
    // A brain that fits a pack keeps a slot that fits one: its rung is cut
    // at the tile's pack limits (the smaller LDS limit, as the pool's
    // plasticity is decided at build), so slack never sends it to the slower
    // one-workgroup-per-brain path.
    const int count = int(specs.size());
    const TileShape *shape = find_tile(t);
    const int nmax = t.nmax();
    const int smax = t.ks > 0 ? t.threads * t.ks
                              : shape ? std::min(shape->lds_plastic, shape->lds_static) : 0;
    std::vector<BrainSpec> out = specs;
    for (auto &sp : out) {
        const int n = int(sp.brain->genes.size()), e = int(sp.brain->edges.size());
        sp.cap_n = ladder(n, slack);
        sp.cap_s = ladder(e, slack);
        if (n <= nmax && e <= smax) {
            sp.cap_n = std::min(sp.cap_n, nmax);
            sp.cap_s = std::min(sp.cap_s, smax);
        }
    }
    if (spare > 0 && count > 0) {
        // spare slots at evenly spaced quantiles of the slot sizes, plus two
        // of the largest
        std::vector<int> order(count);
        std::iota(order.begin(), order.end(), 0);
        std::sort(order.begin(), order.end(), [&](int a, int b) {
            return std::make_pair(out[a].cap_n, out[a].cap_s) < std::make_pair(out[b].cap_n, out[b].cap_s);
        });
        const int nsp = std::max(1, int(std::ceil(double(spare) * count)));
        std::vector<BrainSpec> extra;
        for (int j = 0; j < nsp; ++j) {
            const BrainSpec &q = out[order[std::min(count - 1, int((j + 0.5) * count / nsp))]];
            extra.push_back(BrainSpec{nullptr, -1, 0, 0, world, q.cap_n, q.cap_s});
        }
        // "jumbo" spares as large as the largest slot in both dimensions:
        // best fit reaches them only when nothing smaller fits, and they
        // return to the free list when their brain is replaced
        int jn = 0, js = 0;
        for (int i = 0; i < count; ++i) { jn = std::max(jn, out[i].cap_n); js = std::max(js, out[i].cap_s); }
        const int jumbos = std::max(8, std::min(256, count / 256));   // free ones cost no compute
        for (int j = 0; j < jumbos; ++j) extra.push_back(BrainSpec{nullptr, -1, 0, 0, world, jn, js});
        out.insert(out.end(), extra.begin(), extra.end());
    }
    return out;
}

std::vector<BrainSpec> Batch::slot_specs(const Tile &t) const {
    return reserve_slots(specs, slack, spare, t, world);
}

void Batch::replace(int k, const int *idx, std::vector<std::shared_ptr<const HostBrain>> &&fresh,
                    const uint32_t *seeds, hipStream_t st) {
    if (k <= 0) return;
    if (int(fresh.size()) != k) throw std::runtime_error("Replacement count differs");
    std::vector<uint8_t> seen(count, 0);
    for (int j = 0; j < k; ++j) {
        if (idx[j] < 0 || idx[j] >= count || seen[idx[j]]) throw std::runtime_error("Invalid brain index");
        seen[idx[j]] = 1;
        if (fresh[j]->channels != channels || int(fresh[j]->motors.size()) != outputs)
            throw std::runtime_error("Batch ports differ");
    }
    std::vector<int> old_slot(k), new_slot(k, -1);
    for (int j = 0; j < k; ++j) old_slot[j] = pool.pool_of[idx[j]];
    stage.clear();
    stage_keep.clear();
    bool rebuild = false;
    for (int j = 0; j < k && !rebuild; ++j) {
        const HostBrain &f = *fresh[j];
        const bool slow = f.slow_cell >= 0 ? f.slow_cell != 0 : has_slow_cell(f);
        if (slow || (!pool.plastic && has_plastic_pair(f))) rebuild = true;
    }
    // the free set is valid again only once this call has placed its brains
    const bool kept_free = free_gen == pool.generation;
    free_gen = ~uint64_t(0);
    if (!rebuild) {
        // best fit by neuron capacity among free slots (spares, earlier
        // vacancies and the replaced brains' own slots)
        std::set<std::tuple<int, int, int>> &freeset = free_fit;
        if (!kept_free) {
            freeset.clear();
            for (int pb = 0; pb < pool.nb; ++pb)
                if (pool.batch_of[pb] < 0 && pool.slot_kind[pb] != 0)
                    freeset.insert({pool.bnc[pb], pool.bsc[pb], pb});
        }
        for (int j = 0; j < k; ++j)
            if (pool.slot_kind[old_slot[j]] != 0)
                freeset.insert({pool.bnc[old_slot[j]], pool.bsc[old_slot[j]], old_slot[j]});
        std::vector<int> homeless;
        for (int j = 0; j < k; ++j) {
            const int n = int(fresh[j]->genes.size()), e = int(fresh[j]->edges.size());
            auto it = freeset.lower_bound({n, e, -1});
            while (it != freeset.end() && std::get<1>(*it) < e) ++it;
            if (it == freeset.end()) { homeless.push_back(j); continue; }
            new_slot[j] = std::get<2>(*it);
            freeset.erase(it);
        }
        // A newborn that fits no free slot gets a new slot past the last one,
        // in the pool's headroom (rung capacity as the reserve gives), which
        // changes no other slot; the pool is rebuilt only when that room is
        // short.
        if (!homeless.empty()) {
            const int nmax = pool.tile.nmax();
            const int smax = pool.tile.ks > 0 ? pool.tile.threads * pool.tile.ks : pool.smax_lds;
            std::vector<std::pair<int, int>> caps, occ;
            for (int j : homeless) {
                const int n = int(fresh[j]->genes.size()), e = int(fresh[j]->edges.size());
                int cn = ladder(n, slack), cs = ladder(e, slack);
                if (n <= nmax && e <= smax) { cn = std::min(cn, nmax); cs = std::min(cs, smax); }
                caps.push_back({std::max(cn, 1), cs});
                occ.push_back({n, e});
            }
            std::vector<int> added;
            if (pool.append_slots(caps, occ, stage, stage_keep, added)) {
                for (size_t h = 0; h < homeless.size(); ++h) new_slot[homeless[h]] = added[h];
            } else {
                rebuild = true;
            }
        }
    }
    if (rebuild) {
        // survivors keep their state; newborns start empty and are reset below
        NS_HIP(hipStreamSynchronize(st));
        std::vector<ns_neuron_state> on, nn_;
        std::vector<ns_synapse_state> os, ns_;
        std::vector<uint32_t> oh, nh;
        std::vector<int64_t> ot;
        std::vector<uint64_t> ost, nst;
        pool.download_state(on, os, oh, ot, &ost);
        std::vector<int64_t> noff(count + 1, 0), soff(count + 1, 0);
        for (int i = 0; i < count; ++i) {
            noff[i + 1] = noff[i] + int64_t(brains[i]->genes.size());
            soff[i + 1] = soff[i] + int64_t(brains[i]->edges.size());
        }
        for (int j = 0; j < k; ++j) {
            brains[idx[j]] = std::move(fresh[j]);
            specs[idx[j]].brain = brains[idx[j]].get();
        }
        for (int i = 0; i < count; ++i) {
            const int64_t n = int64_t(brains[i]->genes.size()), e = int64_t(brains[i]->edges.size());
            if (seen[i]) {
                nn_.insert(nn_.end(), size_t(n), ns_neuron_state{});
                ns_.insert(ns_.end(), size_t(e), ns_synapse_state{});
                nh.insert(nh.end(), size_t(n), 0u);
                nst.insert(nst.end(), size_t(n) * 6, 0ull);
                ot[i] = 0;
            } else {
                nn_.insert(nn_.end(), on.begin() + noff[i], on.begin() + noff[i + 1]);
                ns_.insert(ns_.end(), os.begin() + soff[i], os.begin() + soff[i + 1]);
                nh.insert(nh.end(), oh.begin() + noff[i], oh.begin() + noff[i + 1]);
                nst.insert(nst.end(), ost.begin() + noff[i] * 6, ost.begin() + noff[i + 1] * 6);
            }
        }
        pool.build(slot_specs(pool.tile), outputs, pool.tile, stream);
        pool.upload_state(nn_, ns_, nh, ot, &nst);
        NS_HIP(hipStreamSynchronize(stream));
        ++relayouts;
    } else {
        // the replaced genomes are released on the reaper's thread
        std::vector<std::shared_ptr<const HostBrain>> retired(static_cast<size_t>(k));
        for (int j = 0; j < k; ++j) {
            retired[size_t(j)] = std::move(brains[idx[j]]);
            brains[idx[j]] = std::move(fresh[j]);
            specs[idx[j]].brain = brains[idx[j]].get();
        }
        GenomeReaper::get().retire(std::move(retired));
        std::vector<std::pair<int, const BrainSpec *>> moves;
        std::vector<uint8_t> reused(pool.nb, 0);
        for (int j = 0; j < k; ++j) {
            moves.push_back({new_slot[j], &specs[idx[j]]});
            reused[new_slot[j]] = 1;
        }
        for (int j = 0; j < k; ++j)
            if (!reused[old_slot[j]]) moves.push_back({old_slot[j], nullptr});
        pool.place(moves, stage);
        placements += k;
        free_gen = pool.generation;   // appended slots, if any, are all taken
    }
    // Fresh state for the newborns' slots (initial states seeded as
    // ns_gpu_reset) and the run flags of the last mask under the new slot
    // mapping, in the same upload as the images.
    std::vector<int32_t> born(2 * size_t(k));
    std::vector<uint8_t> next(pool.nb, 0);
    for (int j = 0; j < k; ++j) {
        born[j] = pool.pool_of[idx[j]];
        born[k + j] = int32_t(seeds ? seeds[j] : 0u);
    }
    for (int pb = 0; pb < pool.nb; ++pb) {
        const int i = pool.batch_of[pb];
        next[pb] = uint8_t((i >= 0 && last_mask[i] ? 1 : 0) | (pool.fallback[pb] ? 2 : 0));
    }
    stage.add(pool.d_reset, born.data(), 4u * born.size());
    stage.add(pool.view.brun, next.data(), size_t(pool.nb));
    pool.upload_stage(stage, st);
    pool.scatter_range(0, int(stage.entries.size()), st);
    // the reset clears the newborns' settled marks, which the slot order reads
    launch_reset_slots(pool.view, pool.d_reset, k, pool.d_norig, pool.d_sorig, st);
    pool.order_touched(st);
    NS_HIP(hipGetLastError());
    h_brun = next;
    brun_valid = true;
}

void Batch::relayout(const Tile &tile) {
    std::vector<ns_neuron_state> n;
    std::vector<ns_synapse_state> s;
    std::vector<uint32_t> h;
    std::vector<int64_t> ticks;
    std::vector<uint64_t> stamps;
    pool.download_state(n, s, h, ticks, &stamps);
    pool.build(slot_specs(tile), outputs, tile, stream);
    pool.upload_state(n, s, h, ticks, &stamps);
    brun_valid = false;
}

// Uploads the per-brain run flags only when they change; the upload waits
// for earlier launches that may still read them.
void Batch::set_run(const uint8_t *mask) {
    last_mask.assign(mask, mask + count);
    std::vector<uint8_t> next(pool.nb);
    for (int pb = 0; pb < pool.nb; ++pb) {
        const int i = pool.batch_of[pb];
        next[pb] = uint8_t((i >= 0 && mask[i] ? 1 : 0) | (pool.fallback[pb] ? 2 : 0));
    }
    if (brun_valid && next == h_brun) return;
    NS_HIP(hipDeviceSynchronize());
    h_brun = next;
    NS_HIP(hipMemcpy(pool.view.brun, h_brun.data(), pool.nb, hipMemcpyHostToDevice));
    brun_valid = true;
}

void Batch::StepGraph::release() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
}

void Batch::run(int ticks, const float *d_in, float gain, float *d_out, int homeostasis,
                hipStream_t st) {
    const int64_t out_count = int64_t(count) * outputs;
    if (!graph_mode) {
        run_pool(pool, form, ticks, d_in, gain, d_out, out_count, homeostasis, st,
                 timing ? &timer : nullptr);
        return;
    }
    StepGraph &g = step_graph;
    if (!g.exec || g.generation != pool.generation || g.form != form || g.ticks != ticks ||
        g.homeostasis != homeostasis || g.in != d_in || g.out != d_out || g.gain != gain ||
        g.stream != st) {
        g.release();
        NS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            run_pool(pool, form, ticks, d_in, gain, d_out, out_count, homeostasis, st, nullptr);
        } catch (...) {
            hipGraph_t broken = nullptr;
            (void)hipStreamEndCapture(st, &broken);
            if (broken) (void)hipGraphDestroy(broken);
            throw;
        }
        NS_HIP(hipStreamEndCapture(st, &g.graph));
        NS_HIP(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
        g.generation = pool.generation;
        g.form = form;
        g.ticks = ticks;
        g.homeostasis = homeostasis;
        g.in = d_in;
        g.out = d_out;
        g.gain = gain;
        g.stream = st;
        ++graph_captures;
    }
    hipEvent_t ev;
    if (timing) timer.begin(st, form == NS_FORM_PER_TICK ? 1 : 0, ev);
    NS_HIP(hipGraphLaunch(g.exec, st));
    if (timing) timer.end(st);
    ++graph_launches;
}

void run_pool(DevicePool &pool, int form, int ticks, const float *d_in, float gain, float *d_out,
              int64_t out_count, int homeostasis, hipStream_t st, KernelTimer *timer) {
    hipEvent_t ev;
    launch_clear_step(d_out, out_count, pool.view.bstatus, pool.nb, st);
    if (form == NS_FORM_PER_TICK) {
        if (timer) timer->begin(st, 1, ev);
        launch_per_tick(pool.view, pool.all_range, ticks, 1, d_in, gain, d_out, homeostasis, st);
        if (timer) timer->end(st);
        return;
    }
    if (!pool.packs.empty() || pool.nbig) {
        const TileShape *shape = find_tile(pool.tile);
        if (!shape) throw std::runtime_error("Unknown tile shape");
        if (timer) timer->begin(st, 0, ev);
        if (shape->synapses_per_thread > 0)
            NS_HIP(shape->launch(pool.view, pool.d_packs, int(pool.packs.size()), pool.d_big,
                                 pool.nbig, pool.d_perm, ticks, d_in, gain, d_out, homeostasis, st));
        else
            NS_HIP((pool.plastic ? shape->launch_plastic : shape->launch_static)(
                pool.view, pool.d_packs, int(pool.packs.size()), pool.d_big, pool.nbig, pool.d_perm,
                ticks, d_in, gain, d_out, homeostasis, pool.smax_lds, st));
        if (timer) timer->end(st);
    }
    if (pool.fallback_brains) {
        if (timer) timer->begin(st, 1, ev);
        launch_per_tick(pool.view, pool.fallback_range, ticks, 3, d_in, gain, d_out, homeostasis, st);
        if (timer) timer->end(st);
    }
}

Batch::~Batch() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (step_graph.stream) (void)hipStreamSynchronize(step_graph.stream);
    step_graph.release();
    pool.arena.release();
    io.release();
    if (stream) (void)hipStreamDestroy(stream);
}

}  // namespace nsa

using namespace nsa;

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char *ns_error(void) { return g_error.c_str(); }

void *ns_create(int count, int edges, int channels, int outputs, const float *ng, const float *sg,
                const int *motors) {
    try {
        return make_handle(parse_gym_brain(count, edges, channels, outputs, ng, sg, motors, 8));
    } catch (const std::exception &e) {
        set_error(e.what());
        return nullptr;
    }
}

void *ns_create_ex(int count, int edges, int channels, int outputs, const float *ng,
                   const float *sg, const int *motors, int max_delay) {
    try {
        return make_handle(parse_gym_brain(count, edges, channels, outputs, ng, sg, motors, max_delay));
    } catch (const std::exception &e) {
        set_error(e.what());
        return nullptr;
    }
}

void ns_original_indices(void *p, const uint32_t *neurons, const uint32_t *synapses) {
    auto &h = *static_cast<GenomeHandle *>(p);
    // a genome a batch shares stays as the batch took it
    if (h.genome.use_count() > 1) h.genome = std::make_shared<HostBrain>(*h.genome);
    HostBrain &b = *h.genome;
    b.orig_n.assign(neurons, neurons + b.genes.size());
    if (!b.edges.empty()) b.orig_s.assign(synapses, synapses + b.edges.size());
}

void ns_destroy(void *p) { delete static_cast<GenomeHandle *>(p); }

static void *make_batch(std::vector<std::shared_ptr<const HostBrain>> &&brains, bool world) {
    auto *batch = new Batch;
    try {
        batch->world = world;
        batch->create(std::move(brains), 0);
    } catch (...) {
        delete batch;
        throw;
    }
    return batch;
}

void *ns_gpu_create(void **handles, int count) {
    try {
        if (count < 1) throw std::runtime_error("Empty GPU batch");
        std::vector<std::shared_ptr<const HostBrain>> brains;
        brains.reserve(count);
        for (int i = 0; i < count; ++i) {
            if (!handles[i]) throw std::runtime_error("Invalid genome handle");
            brains.push_back(handle_genome(handles[i]));
        }
        return make_batch(std::move(brains), false);
    } catch (const std::exception &e) {
        set_error(e.what());
        return nullptr;
    }
}

void *ns_gpu_create_bulk(int count, const int *neuron_counts, const int *edge_counts, int channels,
                         int outputs, const float *ng, const float *sg, const int *motors,
                         int max_delay, int world_rules) {
    try {
        if (count < 1) throw std::runtime_error("Empty GPU batch");
        std::vector<std::shared_ptr<const HostBrain>> brains;
        brains.reserve(count);
        int64_t n0 = 0, s0 = 0;
        for (int i = 0; i < count; ++i) {
            std::unique_ptr<HostBrain> b(parse_gym_brain(neuron_counts[i], edge_counts[i], channels,
                                                         outputs, ng + n0 * 10, sg + s0 * 11,
                                                         motors + int64_t(i) * outputs, max_delay));
            if (world_rules) {
                // src/main.cu:729-737: 230 pA per level, receptor events only on
                // channels below 6 (gene sensory channel = input channel).
                for (size_t k = 0; k < b->genes.size(); ++k) {
                    b->genes[k].sensory_channel = int8_t(b->sensory[k]);
                    b->receptor[k] = b->genes[k].type == 5 && b->sensory[k] >= 0 && b->sensory[k] < 6;
                }
            }
            brains.push_back(std::shared_ptr<const HostBrain>(std::move(b)));
            n0 += neuron_counts[i];
            s0 += edge_counts[i];
        }
        return make_batch(std::move(brains), world_rules != 0);
    } catch (const std::exception &e) {
        set_error(e.what());
        return nullptr;
    }
}

void ns_gpu_destroy(void *p) { delete static_cast<Batch *>(p); }

int ns_gpu_reset(void *p, const uint8_t *mask, const uint32_t *seeds) {
    try {
        auto &b = *static_cast<Batch *>(p);
        std::vector<uint32_t> bseed(b.pool.nb, 0);
        for (int pb = 0; pb < b.pool.nb; ++pb)
            if (b.pool.batch_of[pb] >= 0) bseed[pb] = seeds[b.pool.batch_of[pb]];
        b.set_run(mask);
        NS_HIP(hipMemcpyAsync(b.pool.d_bseed, bseed.data(), sizeof(uint32_t) * b.pool.nb,
                              hipMemcpyHostToDevice, b.stream));
        launch_reset(b.pool.view, b.pool.nn, b.pool.ns, b.pool.nb, b.pool.d_bseed, b.pool.d_norig,
                     b.pool.d_sorig, b.stream);
        NS_HIP(hipGetLastError());
        NS_HIP(hipStreamSynchronize(b.stream));
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

int ns_gpu_step(void *p, const float *input, const uint8_t *mask, int ticks, float gain,
                float *output, int *status) {
    try {
        auto &b = *static_cast<Batch *>(p);
        if (ticks < 1 || ticks > 10000 || !std::isfinite(gain) || gain < 0)
            throw std::runtime_error("Invalid neural integration interval");
        for (int i = 0; i < b.count; ++i)
            if (mask[i])
                for (int j = 0; j < b.channels; ++j) {
                    float v = input[size_t(i) * b.channels + j];
                    if (!std::isfinite(v) || v < 0 || v > 1)
                        throw std::runtime_error("Invalid encoded observation");
                }
        NS_HIP(hipMemcpyAsync(b.d_input, input, sizeof(float) * b.count * b.channels,
                              hipMemcpyHostToDevice, b.stream));
        b.set_run(mask);
        b.run(ticks, b.d_input, gain, b.d_output, b.world ? 0 : 1, b.stream);
        NS_HIP(hipGetLastError());
        NS_HIP(hipMemcpyAsync(output, b.d_output, sizeof(float) * b.count * b.outputs,
                              hipMemcpyDeviceToHost, b.stream));
        b.h_status.resize(b.pool.nb);
        NS_HIP(hipMemcpyAsync(b.h_status.data(), b.pool.view.bstatus, sizeof(int32_t) * b.pool.nb,
                              hipMemcpyDeviceToHost, b.stream));
        NS_HIP(hipStreamSynchronize(b.stream));
        for (int pb = 0; pb < b.pool.nb; ++pb)
            if (b.pool.batch_of[pb] >= 0) status[b.pool.batch_of[pb]] = b.h_status[pb];
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

int ns_gpu_step_device(void *p, const float *input, const uint8_t *mask, int ticks, float gain,
                       float *output, int *status, void *stream) {
    try {
        auto &b = *static_cast<Batch *>(p);
        if (ticks < 1 || ticks > 10000) throw std::runtime_error("Invalid neural integration interval");
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : b.stream;
        if (mask) {
            std::vector<uint8_t> h(b.count);
            NS_HIP(hipMemcpy(h.data(), mask, b.count, hipMemcpyDeviceToHost));
            b.set_run(h.data());
        } else {
            std::vector<uint8_t> ones(b.count, 1);
            b.set_run(ones.data());
        }
        b.run(ticks, input, gain, output, b.world ? 0 : 1, st);
        NS_HIP(hipGetLastError());
        if (status) {
            NS_HIP(hipStreamSynchronize(st));
            b.h_status.resize(b.pool.nb);
            NS_HIP(hipMemcpy(b.h_status.data(), b.pool.view.bstatus, sizeof(int32_t) * b.pool.nb,
                             hipMemcpyDeviceToHost));
            std::vector<int32_t> s(b.count, 0);
            for (int pb = 0; pb < b.pool.nb; ++pb)
                if (b.pool.batch_of[pb] >= 0) s[b.pool.batch_of[pb]] = b.h_status[pb];
            NS_HIP(hipMemcpy(status, s.data(), sizeof(int32_t) * b.count, hipMemcpyHostToDevice));
        }
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

int ns_gpu_set_graph(void *p, int on) {
    try {
        auto &b = *static_cast<Batch *>(p);
        if (on != 0 && on != 1) throw std::runtime_error("Invalid graph mode");
        if (!on && b.step_graph.exec) {
            if (b.step_graph.stream) NS_HIP(hipStreamSynchronize(b.step_graph.stream));
            b.step_graph.release();
        }
        b.graph_mode = on != 0;
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

int ns_gpu_reserve(void *p, float slack, float spare) {
    try {
        auto &b = *static_cast<Batch *>(p);
        if (!(slack >= 0 && slack <= 4) || !(spare >= 0 && spare <= 4))
            throw std::runtime_error("Invalid slot reserve");
        b.slack = slack;
        b.spare = spare;
        b.relayout(b.pool.tile);
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

int ns_gpu_replace(void *p, int count, const int *indices, void **handles, const uint32_t *seeds,
                   void *stream) {
    try {
        auto &b = *static_cast<Batch *>(p);
        if (count < 0 || (count && (!indices || !handles))) throw std::runtime_error("Invalid replacement");
        for (int j = 0; j < count; ++j)
            if (!handles[j]) throw std::runtime_error("Invalid genome handle");
        std::vector<std::shared_ptr<const HostBrain>> fresh(static_cast<size_t>(count));
        for (int j = 0; j < count; ++j) fresh[size_t(j)] = handle_genome(handles[j]);
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : b.stream;
        b.replace(count, indices, std::move(fresh), seeds, st);
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

int ns_gpu_set_timing(void *p, int on) {
    try {
        auto &b = *static_cast<Batch *>(p);
        b.timer.resolve();
        b.timer.ms[0] = b.timer.ms[1] = 0;
        b.timer.count[0] = b.timer.count[1] = 0;
        b.timing = on != 0;
        if (b.timing) b.timer.reserve(512);   // 256 timed launches without creating events
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

int ns_gpu_set_form(void *p, int form) {
    auto &b = *static_cast<Batch *>(p);
    if (form != NS_FORM_RESIDENT && form != NS_FORM_PER_TICK) {
        set_error("Invalid execution form");
        return -1;
    }
    b.form = form;
    return 0;
}

int ns_gpu_set_tiles(void *p, int threads, int kn, int ks, int want_minw, int lds) {
    try {
        auto &b = *static_cast<Batch *>(p);
        const TileShape *found = nullptr;
        for (int k = 0; k < kNumTiles && !found; ++k) {
            const TileShape &t = kTiles[k];
            if (t.threads == threads && t.neurons_per_thread == kn && t.synapses_per_thread == ks &&
                (want_minw == 0 || t.min_waves == want_minw) && (ks > 0 || t.lds_plastic == lds))
                found = &t;
        }
        if (!found) throw std::runtime_error("Tile shape not compiled");
        b.relayout(Tile{threads, kn, ks, found->min_waves, ks > 0 ? 0 : lds});
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

int ns_gpu_autotune(void *p, int ticks, int *threads, int *kn, int *ks, int *minw, int *lds) {
    try {
        auto &b = *static_cast<Batch *>(p);
        std::vector<ns_neuron_state> n;
        std::vector<ns_synapse_state> s;
        std::vector<uint32_t> h;
        std::vector<int64_t> tk;
        std::vector<uint64_t> stamps;
        b.pool.download_state(n, s, h, tk, &stamps);
        unsigned long long events0[3] = {0, 0, 0};
        NS_HIP(hipMemcpy(events0, b.pool.view.events, sizeof events0, hipMemcpyDeviceToHost));
        // A fixed sensory pattern (30% of channels at level 1) at the gym gain,
        // so every tile is timed on the same, driven dynamics.
        std::vector<uint8_t> ones(b.count, 1);
        std::vector<float> levels(size_t(b.count) * b.channels, 0.0f);
        for (size_t i = 0; i < levels.size(); ++i) levels[i] = (i * 7919u) % 10 < 3 ? 1.0f : 0.0f;
        const float gain = 230.0f;
        b.brun_valid = false;
        NS_HIP(hipMemcpy(b.d_input, levels.data(), sizeof(float) * levels.size(), hipMemcpyHostToDevice));
        double best = 1e30;
        Tile best_tile = b.pool.tile;
        hipEvent_t e0, e1;
        NS_HIP(hipEventCreate(&e0));
        NS_HIP(hipEventCreate(&e1));
        for (int k = 0; k < kNumTiles; ++k) {
            Tile t{kTiles[k].threads, kTiles[k].neurons_per_thread, kTiles[k].synapses_per_thread,
                   kTiles[k].min_waves, kTiles[k].synapses_per_thread > 0 ? 0 : kTiles[k].lds_plastic};
            b.pool.build(b.slot_specs(t), b.outputs, t, b.stream);
            b.pool.upload_state(n, s, h, tk, &stamps);
            b.brun_valid = false;
            b.set_run(ones.data());
            b.run(ticks, b.d_input, gain, b.d_output, 1, b.stream);   // warm-up
            b.pool.upload_state(n, s, h, tk, &stamps);
            NS_HIP(hipEventRecord(e0, b.stream));
            b.run(ticks, b.d_input, gain, b.d_output, 1, b.stream);
            b.run(ticks, b.d_input, gain, b.d_output, 1, b.stream);
            NS_HIP(hipEventRecord(e1, b.stream));
            NS_HIP(hipEventSynchronize(e1));
            float ms = 0;
            NS_HIP(hipEventElapsedTime(&ms, e0, e1));
            // a later tile must win by more than 1.5% (stable choice under noise)
            if (ms < best * 0.985) { best = ms; best_tile = t; }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        b.pool.build(b.slot_specs(best_tile), b.outputs, best_tile, b.stream);
        b.pool.upload_state(n, s, h, tk, &stamps);
        b.brun_valid = false;
        NS_HIP(hipMemcpy(b.pool.view.events, events0, sizeof events0, hipMemcpyHostToDevice));
        if (threads) *threads = best_tile.threads;
        if (kn) *kn = best_tile.kn;
        if (ks) *ks = best_tile.ks;
        if (minw) *minw = best_tile.minw;
        if (lds) *lds = best_tile.lds;
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

// A 9-slot ring holds the last 9 ticks, which is the whole delay history only
// for batches whose delays are at most 8 ticks (gym_adapter/core.cuh:18).
static void check_history9(const Batch &b) {
    int md = 0;
    for (const auto &hb : b.brains) md = std::max(md, hb->max_delay);
    if (md > 8)
        throw std::runtime_error("Delay bound " + std::to_string(md) +
                                 " exceeds the 9-slot history; use the spike bit registers");
}

int ns_gpu_export(void *p, ns_neuron_state *neurons, ns_synapse_state *synapses, uint8_t *history9,
                  int64_t *ticks) {
    try {
        auto &b = *static_cast<Batch *>(p);
        if (history9) check_history9(b);
        std::vector<ns_neuron_state> n;
        std::vector<ns_synapse_state> s;
        std::vector<uint32_t> h;
        std::vector<int64_t> tk;
        b.pool.download_state(n, s, h, tk);
        if (neurons) std::memcpy(neurons, n.data(), n.size() * sizeof(n[0]));
        if (synapses && !s.empty()) std::memcpy(synapses, s.data(), s.size() * sizeof(s[0]));
        if (ticks) std::memcpy(ticks, tk.data(), tk.size() * sizeof(tk[0]));
        if (history9) {
            int64_t off = 0;
            for (int i = 0; i < b.count; ++i) {
                const int64_t T = tk[i];
                const int nc = b.pool.occ_n[b.pool.pool_of[i]];
                for (int k = 0; k < nc; ++k) {
                    const uint32_t bits = h[off + k];
                    for (int slot = 0; slot < 9; ++slot) {
                        const int64_t last = T - 1;
                        const int64_t back = ((last - slot) % 9 + 9) % 9;
                        const int64_t tau = last - back;
                        history9[(off + k) * 9 + slot] = tau < 0 ? 0 : uint8_t((bits >> back) & 1u);
                    }
                }
                off += nc;
            }
        }
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

int ns_gpu_import(void *p, const ns_neuron_state *neurons, const ns_synapse_state *synapses,
                  const uint8_t *history9, const int64_t *ticks) {
    try {
        auto &b = *static_cast<Batch *>(p);
        if (history9) check_history9(b);
        std::vector<ns_neuron_state> n;
        std::vector<ns_synapse_state> s;
        std::vector<uint32_t> h;
        std::vector<int64_t> tk;
        b.pool.download_state(n, s, h, tk);
        if (neurons) std::memcpy(n.data(), neurons, n.size() * sizeof(n[0]));
        if (synapses && !s.empty()) std::memcpy(s.data(), synapses, s.size() * sizeof(s[0]));
        if (ticks) std::memcpy(tk.data(), ticks, tk.size() * sizeof(tk[0]));
        if (history9) {
            int64_t off = 0;
            for (int i = 0; i < b.count; ++i) {
                const int64_t T = tk[i];
                const int nc = b.pool.occ_n[b.pool.pool_of[i]];
                for (int k = 0; k < nc; ++k) {
                    uint32_t bits = 0;
                    for (int j = 0; j < 9; ++j) {
                        const int64_t tau = T - 1 - j;
                        if (tau >= 0 && history9[(off + k) * 9 + tau % 9]) bits |= 1u << j;
                    }
                    h[off + k] = bits;
                }
                off += nc;
            }
        }
        b.pool.upload_state(n, s, h, tk);
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

int ns_gpu_export_spike_bits(void *p, uint32_t *bits) {
    try {
        auto &b = *static_cast<Batch *>(p);
        std::vector<ns_neuron_state> n;
        std::vector<ns_synapse_state> s;
        std::vector<uint32_t> h;
        std::vector<int64_t> tk;
        b.pool.download_state(n, s, h, tk);
        std::memcpy(bits, h.data(), h.size() * sizeof(uint32_t));
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

int ns_gpu_import_spike_bits(void *p, const uint32_t *bits) {
    try {
        auto &b = *static_cast<Batch *>(p);
        if (!bits) throw std::runtime_error("No spike registers");
        std::vector<ns_neuron_state> n;
        std::vector<ns_synapse_state> s;
        std::vector<uint32_t> h;
        std::vector<int64_t> tk;
        b.pool.download_state(n, s, h, tk);
        std::memcpy(h.data(), bits, h.size() * sizeof(uint32_t));
        b.pool.upload_state(n, s, h, tk);
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

int ns_gpu_info(void *p, ns_batch_info *info) {
    try {
        auto &b = *static_cast<Batch *>(p);
        unsigned long long events[3] = {0, 0, 0};
        NS_HIP(hipStreamSynchronize(b.stream));
        NS_HIP(hipMemcpy(events, b.pool.view.events, sizeof events, hipMemcpyDeviceToHost));
        info->brains = b.count;
        info->neurons = b.pool.nn;
        info->synapses = b.pool.ns;
        info->resident_packs = int64_t(b.pool.packs.size());
        info->fallback_brains = b.pool.fallback_brains;
        info->events = int64_t(events[0] + events[1]);
        info->resident_events = int64_t(events[0]);
        info->fallback_neurons = b.pool.fallback_range.n_hi - b.pool.fallback_range.n_lo;
        info->fallback_synapses = b.pool.fallback_range.s_hi - b.pool.fallback_range.s_lo;
        info->big_brains = b.pool.nbig;
        info->slots = b.pool.nb;
        int64_t free_slots = 0;
        for (int pb = 0; pb < b.pool.nb; ++pb) free_slots += b.pool.batch_of[pb] < 0;
        info->free_slots = free_slots;
        info->slot_neurons = b.pool.nn;
        info->slot_synapses = b.pool.ns;
        info->placements = b.placements;
        info->relayouts = b.relayouts;
        info->graph_captures = b.graph_captures;
        info->graph_launches = b.graph_launches;
        info->staging_grows = b.pool.staging.grows + b.pool.dstage_grows;
        info->appended_slots = b.pool.appended;
        info->skipped_neuron_ticks = int64_t(events[2]);
        int64_t big_n = 0;
        for (int k = 0; k < b.pool.nbig; ++k) big_n += b.pool.bnc[b.pool.fallback_brains + k];
        info->big_neurons = big_n;
        info->threads = b.pool.tile.threads;
        info->neurons_per_thread = b.pool.tile.kn;
        info->synapses_per_thread = b.pool.tile.ks;
        info->min_waves = b.pool.tile.minw;
        info->lds_synapses = b.pool.smax_lds;
        info->form = b.form;
        info->device_bytes = b.pool.arena.bytes + b.io.bytes;
        b.timer.resolve();
        info->resident_ms = b.timer.ms[0];
        info->per_tick_ms = b.timer.ms[1];
        info->resident_launches = b.timer.count[0];
        info->per_tick_calls = b.timer.count[1];
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

int ns_device_math_digest(int fn, uint64_t lo, uint64_t hi, int64_t chunk, uint64_t *digests) {
    try {
        if ((fn != 0 && fn != 1) || hi <= lo || hi > (uint64_t(1) << 32) || chunk < 1)
            throw std::runtime_error("Invalid digest range");
        const int64_t n = int64_t((hi - lo + uint64_t(chunk) - 1) / uint64_t(chunk));
        unsigned long long *d = nullptr;
        NS_HIP(hipMalloc(&d, sizeof(unsigned long long) * n));
        NS_HIP(hipMemset(d, 0, sizeof(unsigned long long) * n));
        launch_math_digest(fn, lo, hi, chunk, d, nullptr);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e == hipSuccess) e = hipMemcpy(digests, d, sizeof(uint64_t) * n, hipMemcpyDeviceToHost);
        (void)hipFree(d);
        NS_HIP(e);
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

}  // extern "C"
