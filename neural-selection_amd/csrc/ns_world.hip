// World-pool surface of libns_amd.so: the neural part of src/main.cu's step
// loop over checkpointable neuron/synapse records.
//
//   records in    -> ns_pool_import_records  (checkpoint_impl.cuh:272-311 checks)
//   births/deaths -> ns_pool_replace_records (main.cu:364-454 slot claims,
//                                             1302-1338 genome reclaim)
//   16 ticks      -> ns_world_neural_step    (main.cu:2702-2728, rules 659-801)
//   homeostasis   -> ns_world_homeostasis    (main.cu:1287-1300)
//   records out   -> ns_pool_export_records  (main.cu:197-264, ascending slots)
//
// Animal a's records become one brain in the pool slot of batch index a:
// neurons in ascending slot order, synapses sorted by (post, slot).  The
// reference accumulates pending conductances with float atomics (order-
// nondeterministic); here the sum runs in ascending synapse-slot order per
// post, a fixed order.
#include <algorithm>
#include <cstddef>
#include <cstring>
#include <memory>
#include <numeric>
#include <set>
#include <tuple>

#include "ns_host.h"

namespace nsa {

namespace {
constexpr int kSenses = 19, kMotors = 5;   // src/main.cu:40-41 (S, M)
constexpr int kTicksPerStep = 16;          // src/main.cu:47
constexpr int kMaxHistory = 17;            // ring slots of the 1-16 ms delay layout

// Records are held in the 17-slot form (the 9-slot record is the same up to
// its ring, ns_amd.h); a pool's own ring length is `history`.
using NRec = ns_checkpoint_neuron17;
constexpr size_t kRingOffset = offsetof(ns_checkpoint_neuron, spike_history);

NRec widen(const ns_checkpoint_neuron &r) {
    NRec o;
    std::memset(&o, 0, sizeof o);
    std::memcpy(&o, &r, kRingOffset);
    std::memcpy(o.spike_history, r.spike_history, sizeof r.spike_history);
    return o;
}
NRec widen(const NRec &r) { return r; }
void narrow(const NRec &r, ns_checkpoint_neuron *o) {
    std::memset(o, 0, sizeof *o);
    std::memcpy(o, &r, kRingOffset);
    std::memcpy(o->spike_history, r.spike_history, sizeof o->spike_history);
}
void narrow(const NRec &r, NRec *o) { *o = r; }
template <class R> constexpr int history_of() { return sizeof(R::spike_history); }

struct WorldPool {
    int P, ncap, scap;
    int history = kWorldHistory;              // ring slots: 9, or 17 for 1-16 ms delays
    // Per animal: its records in ascending slot order (genes as given; state
    // as imported, refreshed from the device by pull()), its brain and its
    // pool spec (stable addresses: slot placements read them).
    std::vector<std::vector<NRec>> nrec;
    std::vector<std::vector<ns_checkpoint_synapse>> srec;
    std::vector<HostBrain> brains;
    std::vector<BrainSpec> spec;
    std::vector<uint8_t> real;                // animal has neuron records
    std::vector<uint8_t> nused, sused;        // record slots in use
    int nreal = 0;
    int64_t nrecords = 0, srecords = 0;
    // slot reserve (ns_pool_reserve) and turnover counters
    float slack = 0.0f, spare = 0.0f;
    int64_t placements = 0, relayouts = 0;
    bool built = false;                       // pool slot tables describe the animals
    DevicePool pool;
    Stager stage;
    DeviceArena io, age_io;
    float *d_senses = nullptr, *d_out = nullptr;
    int32_t *d_age = nullptr;                 // one age per pool slot
    int age_cap = 0;
    hipStream_t stream = nullptr;
    hipStream_t last_stream = nullptr;        // the caller's stream of the last device step
    int64_t next_tick = -1;                   // tick the device rings are aligned to
    bool rings_pending = true;                // records' rings not yet on the device
    int form = NS_FORM_RESIDENT;

    ~WorldPool() {
        if (stream) (void)hipStreamSynchronize(stream);
        pool.arena.release();
        io.release();
        age_io.release();
        if (stream) (void)hipStreamDestroy(stream);
    }

    // ring slot k (= tick % history) <-> shift-register bit (T-1-tick)
    uint32_t ring_to_bits(const uint8_t *ring, int64_t T) const {
        const int H = history;
        uint32_t bits = 0;
        for (int j = 0; j < H; ++j) {
            const int64_t tau = T - 1 - j;
            if (ring[((tau % H) + H) % H]) bits |= 1u << j;
        }
        return bits;
    }
    void bits_to_ring(uint32_t bits, int64_t T, uint8_t *ring) const {
        const int H = history;
        for (int k = 0; k < H; ++k) {
            const int64_t last = T - 1;
            const int64_t back = ((last - k) % H + H) % H;
            ring[k] = uint8_t((bits >> back) & 1u);
        }
    }

    // checkpoint_impl.cuh:272-311 record checks
    bool valid(const NRec &r) const {
        return r.slot >= 0 && r.slot < ncap && r.owner >= 0 && r.owner < P && r.next >= -1 &&
               r.next < ncap && r.port >= -1 && r.port < kSenses + kMotors && r.gene.type <= 6 &&
               r.gene.modulator <= 3;
    }
    bool valid(const ns_checkpoint_synapse &r) const {
        return r.slot >= 0 && r.slot < scap && r.owner >= 0 && r.owner < P && r.next >= -1 &&
               r.next < scap && r.gene.pre >= 0 && r.gene.pre < ncap && r.gene.post >= 0 &&
               r.gene.post < ncap && r.gene.delay >= 1 && r.gene.delay <= history - 1 &&
               r.gene.branch <= 3;
    }

    // Animal a's brain from its records.  main.cu:702-704: a synapse whose
    // endpoints are not neurons of its own animal never delivers or learns;
    // it keeps its state unchanged (pre = post = 0, inactive).
    HostBrain make_brain(const std::vector<NRec> &nr, const std::vector<ns_checkpoint_synapse> &sr) const {
        HostBrain hb;
        hb.channels = kSenses;
        hb.max_delay = history - 1;               // world_flags: a ring of `history` slots
        hb.motors.assign(kMotors, -1);
        for (size_t k = 0; k < nr.size(); ++k) {
            const auto &r = nr[k];
            hb.genes.push_back(r.gene);
            const int ch = r.gene.sensory_channel;
            const bool sensed = ch >= 0 && ch < kSenses;                   // main.cu:729
            hb.sensory.push_back(sensed ? ch : -1);
            hb.receptor.push_back(sensed && r.gene.type == 5 && ch < 6);   // main.cu:735
            if (r.port >= kSenses && r.port < kSenses + kMotors) {
                // one neuron per motor port, as the animal's port table holds it
                // (main.cu:82 ports[S + M])
                if (hb.motors[r.port - kSenses] >= 0)
                    throw std::runtime_error("Two neurons of one animal on a motor port");
                hb.motors[r.port - kSenses] = int(k);
            }
        }
        auto local = [&](int slot) {
            auto it = std::lower_bound(nr.begin(), nr.end(), slot,
                                       [](const NRec &r, int s) { return r.slot < s; });
            return it != nr.end() && it->slot == slot ? int(it - nr.begin()) : -1;
        };
        for (const auto &r : sr) {
            ns_synapse_gene g = r.gene;
            const int pre = local(g.pre), post = local(g.post);
            if (pre < 0 || post < 0) {
                g.pre = 0; g.post = 0; g.active = 0;
            } else {
                g.pre = pre;
                g.post = post;
            }
            hb.edges.push_back(g);
        }
        return hb;
    }

    void set_animal(int a, std::vector<NRec> &&nr,
                    std::vector<ns_checkpoint_synapse> &&sr, HostBrain &&hb) {
        for (const auto &r : nrec[a]) nused[r.slot] = 0;
        for (const auto &r : srec[a]) sused[r.slot] = 0;
        nrecords += int64_t(nr.size()) - int64_t(nrec[a].size());
        srecords += int64_t(sr.size()) - int64_t(srec[a].size());
        nreal += int(!nr.empty()) - int(real[a]);
        nrec[a] = std::move(nr);
        srec[a] = std::move(sr);
        for (const auto &r : nrec[a]) nused[r.slot] = 1;
        for (const auto &r : srec[a]) sused[r.slot] = 1;
        real[a] = !nrec[a].empty();
        brains[a] = std::move(hb);
        spec[a] = BrainSpec{&brains[a], a, a * kSenses, a * kMotors, true};
    }

    // Device state of every animal into its records (rings only once they
    // are aligned on the device).
    void pull() {
        if (!built || nreal == 0) return;
        std::vector<ns_neuron_state> ns;
        std::vector<ns_synapse_state> ss;
        std::vector<uint32_t> h;
        std::vector<int64_t> ticks;
        std::vector<uint64_t> st;
        pool.download_state(ns, ss, h, ticks, &st);
        int64_t off = 0, soff = 0;
        for (int a = 0; a < P; ++a) {
            if (!real[a]) continue;
            if (pool.pool_of[a] < 0) throw std::runtime_error("Animal without a slot");
            for (auto &r : nrec[a]) {
                r.state = ns[off];
                for (int e = 0; e < 6; ++e) r.mod_event_tick[e] = st[off * 6 + e];
                if (!rings_pending) bits_to_ring(h[off], next_tick, r.spike_history);
                ++off;
            }
            for (auto &r : srec[a]) r.state = ss[soff++];
        }
    }

    // Build the pool over every animal's records (slot reserve applied) and
    // upload the records' state; rings follow at the next step's tick.
    void rebuild() {
        built = false;
        rings_pending = true;
        next_tick = -1;
        std::vector<BrainSpec> specs;
        for (int a = 0; a < P; ++a)
            if (real[a]) specs.push_back(spec[a]);
        if (specs.empty()) return;
        Tile tile{256, 2, 4};
        pool.build(reserve_slots(specs, slack, spare, tile, true), kMotors, tile, stream);
        pool.nbatch = P;                      // batch index = animal
        pool.pool_of.resize(P, -1);
        std::vector<ns_neuron_state> ns;
        std::vector<ns_synapse_state> ss;
        std::vector<uint32_t> h;
        std::vector<uint64_t> st;
        for (int a = 0; a < P; ++a) {
            for (const auto &r : nrec[a]) {
                ns.push_back(r.state);
                h.push_back(0);
                for (int e = 0; e < 6; ++e) st.push_back(r.mod_event_tick[e]);
            }
            for (const auto &r : srec[a]) ss.push_back(r.state);
        }
        std::vector<int64_t> ticks(P, 0);
        pool.upload_state(ns, ss, h, ticks, &st);
        built = true;
    }

    void align_rings(int64_t T) {
        if (rings_pending) {
            std::vector<ns_neuron_state> n;
            std::vector<ns_synapse_state> s;
            std::vector<uint32_t> h;
            std::vector<int64_t> ticks;
            std::vector<uint64_t> st;
            pool.download_state(n, s, h, ticks, &st);
            int64_t off = 0;
            for (int a = 0; a < P; ++a)
                for (const auto &r : nrec[a]) h[off++] = ring_to_bits(r.spike_history, T);
            std::fill(ticks.begin(), ticks.end(), T);
            pool.upload_state(n, s, h, ticks, &st);
            rings_pending = false;
        } else if (T != next_tick) {
            std::vector<ns_neuron_state> n;
            std::vector<ns_synapse_state> s;
            std::vector<uint32_t> h;
            std::vector<int64_t> ticks;
            std::vector<uint64_t> st;
            pool.download_state(n, s, h, ticks, &st);
            uint8_t ring[kMaxHistory];
            for (auto &bits : h) {
                bits_to_ring(bits, next_tick, ring);
                bits = ring_to_bits(ring, T);
            }
            std::fill(ticks.begin(), ticks.end(), T);
            pool.upload_state(n, s, h, ticks, &st);
        }
        next_tick = T;
    }

    // Waits for the pool's stream and the stream of the last device step.
    void sync_all() {
        if (last_stream) NS_HIP(hipStreamSynchronize(last_stream));
        NS_HIP(hipStreamSynchronize(stream));
    }

    // Per-slot run and freeze flags; uploaded (after the work queued on `st`
    // that may still read them) only when they change.
    std::vector<uint8_t> h_run, h_frozen;
    uint64_t flags_generation = 0;
    void set_flags(const uint8_t *frozen, hipStream_t st) {
        std::vector<uint8_t> run(pool.nb), fz(pool.nb);
        for (int pb = 0; pb < pool.nb; ++pb) {
            const int a = pool.batch_of[pb];   // -1: a free slot, which never runs
            run[pb] = uint8_t((a >= 0 ? 1 : 0) | (pool.fallback[pb] ? 2 : 0));
            fz[pb] = frozen && a >= 0 ? frozen[a] : 0;
        }
        if (flags_generation == pool.generation && run == h_run && fz == h_frozen) return;
        NS_HIP(hipStreamSynchronize(st));
        sync_all();
        NS_HIP(hipMemcpy(pool.view.brun, run.data(), run.size(), hipMemcpyHostToDevice));
        NS_HIP(hipMemcpy(pool.view.bfrozen, fz.data(), fz.size(), hipMemcpyHostToDevice));
        h_run = std::move(run);
        h_frozen = std::move(fz);
        flags_generation = pool.generation;
    }

    void replace(int count, const int *animals, const NRec *neurons, int neuron_count,
                 const ns_checkpoint_synapse *synapses, int synapse_count);
    void import_records(const NRec *neurons, int neuron_count, const ns_checkpoint_synapse *synapses,
                        int synapse_count);
};

// Animals animals[0..count) lose their records; the given records (each
// owned by one of them) become their new brains.  Checks everything before
// changing anything.
void WorldPool::replace(int count, const int *animals, const NRec *neurons,
                        int neuron_count, const ns_checkpoint_synapse *synapses, int synapse_count) {
    if (count < 0 || neuron_count < 0 || synapse_count < 0 || (count && !animals) ||
        (neuron_count && !neurons) || (synapse_count && !synapses))
        throw std::runtime_error("Invalid replacement");
    if (count == 0 && (neuron_count || synapse_count))
        throw std::runtime_error("Record owner not replaced");
    if (count == 0) return;
    sync_all();                               // no queued step still reads the slots
    std::vector<int> listed(P, -1);
    for (int j = 0; j < count; ++j) {
        if (animals[j] < 0 || animals[j] >= P || listed[animals[j]] >= 0)
            throw std::runtime_error("Invalid animal index");
        listed[animals[j]] = j;
    }
    int prev = -1;
    for (int i = 0; i < neuron_count; ++i) {
        const auto &r = neurons[i];
        if (!valid(r) || r.slot <= prev) throw std::runtime_error("Invalid checkpoint neuron record");
        if (listed[r.owner] < 0) throw std::runtime_error("Record owner not replaced");
        prev = r.slot;
    }
    prev = -1;
    for (int i = 0; i < synapse_count; ++i) {
        const auto &r = synapses[i];
        if (!valid(r) || r.slot <= prev) throw std::runtime_error("Invalid checkpoint synapse record");
        if (listed[r.owner] < 0) throw std::runtime_error("Record owner not replaced");
        prev = r.slot;
    }
    // new slots must be free once the replaced animals release theirs
    for (int j = 0; j < count; ++j) {
        for (const auto &r : nrec[animals[j]]) nused[r.slot] = 2;
        for (const auto &r : srec[animals[j]]) sused[r.slot] = 2;
    }
    bool clash = false;
    for (int i = 0; i < neuron_count && !clash; ++i) clash = nused[neurons[i].slot] == 1;
    for (int i = 0; i < synapse_count && !clash; ++i) clash = sused[synapses[i].slot] == 1;
    for (int j = 0; j < count; ++j) {
        for (const auto &r : nrec[animals[j]]) nused[r.slot] = 1;
        for (const auto &r : srec[animals[j]]) sused[r.slot] = 1;
    }
    if (clash) throw std::runtime_error("Record slot in use");
    std::vector<std::vector<NRec>> nr(count);
    std::vector<std::vector<ns_checkpoint_synapse>> sr(count);
    for (int i = 0; i < neuron_count; ++i) nr[listed[neurons[i].owner]].push_back(neurons[i]);
    for (int i = 0; i < synapse_count; ++i) {
        const int j = listed[synapses[i].owner];
        if (nr[j].empty()) throw std::runtime_error("Synapse owner has no neurons");
        sr[j].push_back(synapses[i]);
    }
    std::vector<HostBrain> fresh(count);
    for (int j = 0; j < count; ++j) fresh[j] = make_brain(nr[j], sr[j]);

    // Placement into free slots (spares, vacancies and the replaced animals'
    // own slots), best fit by neuron capacity; the pool is rebuilt when a
    // newborn fits none or needs what the pool was not built with.
    bool rebuild_pool = !built;
    for (int j = 0; j < count && !rebuild_pool; ++j)
        if (!nr[j].empty() && (has_slow_cell(fresh[j]) || (!pool.plastic && has_plastic_pair(fresh[j]))))
            rebuild_pool = true;
    std::vector<int> old_slot(count, -1), new_slot(count, -1);
    if (!rebuild_pool) {
        std::set<std::tuple<int, int, int>> freeset;
        for (int pb = 0; pb < pool.nb; ++pb)
            if (pool.batch_of[pb] < 0 && pool.slot_kind[pb] != 0)
                freeset.insert({pool.bnc[pb], pool.bsc[pb], pb});
        for (int j = 0; j < count; ++j) {
            old_slot[j] = pool.pool_of[animals[j]];
            if (old_slot[j] >= 0 && pool.slot_kind[old_slot[j]] != 0)
                freeset.insert({pool.bnc[old_slot[j]], pool.bsc[old_slot[j]], old_slot[j]});
        }
        for (int j = 0; j < count && !rebuild_pool; ++j) {
            if (nr[j].empty()) continue;
            const int n = int(fresh[j].genes.size()), e = int(fresh[j].edges.size());
            auto it = freeset.lower_bound({n, e, -1});
            while (it != freeset.end() && std::get<1>(*it) < e) ++it;
            if (it == freeset.end()) { rebuild_pool = true; break; }
            new_slot[j] = std::get<2>(*it);
            freeset.erase(it);
        }
    }
    if (rebuild_pool) {
        // survivors keep their device state through their records
        sync_all();
        pull();
        for (int j = 0; j < count; ++j)
            set_animal(animals[j], std::move(nr[j]), std::move(sr[j]), std::move(fresh[j]));
        rebuild();
        ++relayouts;
        return;
    }
    for (int j = 0; j < count; ++j)
        set_animal(animals[j], std::move(nr[j]), std::move(sr[j]), std::move(fresh[j]));
    std::vector<std::pair<int, const BrainSpec *>> moves;
    std::vector<uint8_t> reused(pool.nb, 0);
    for (int j = 0; j < count; ++j)
        if (new_slot[j] >= 0) {
            moves.push_back({new_slot[j], &spec[animals[j]]});
            reused[new_slot[j]] = 1;
        }
    for (int j = 0; j < count; ++j)
        if (old_slot[j] >= 0 && !reused[old_slot[j]]) moves.push_back({old_slot[j], nullptr});
    stage.clear();
    pool.place(moves, stage);
    // The newborns' state from their records: rings as shift registers at
    // the tick the device is aligned to (or, before the first step, left to
    // the first step's alignment, which reads every animal's records).
    std::vector<std::vector<unsigned char>> keep;
    for (int j = 0; j < count; ++j) {
        const int a = animals[j];
        if (new_slot[j] < 0) continue;
        const auto &r = nrec[a];
        const int n = int(r.size());
        std::vector<ns_neuron_state> ns(n);
        std::vector<uint32_t> h(n, 0);
        std::vector<uint64_t> st(size_t(n) * 6);
        for (int k = 0; k < n; ++k) {
            ns[k] = r[k].state;
            if (!rings_pending) h[k] = ring_to_bits(r[k].spike_history, next_tick);
            for (int e = 0; e < 6; ++e) st[size_t(k) * 6 + e] = r[k].mod_event_tick[e];
        }
        std::vector<ns_synapse_state> ss(srec[a].size());
        for (size_t k = 0; k < ss.size(); ++k) ss[k] = srec[a][k].state;
        pool.stage_state(new_slot[j], brains[a], ns.data(), h.data(), st.data(), ss.data(),
                         rings_pending ? 0 : next_tick, stage, keep);
    }
    pool.upload_stage(stage, stream);
    pool.scatter_range(0, int(stage.entries.size()), stream);
    pool.order_touched(stream);
    NS_HIP(hipGetLastError());
    sync_all();
    for (int j = 0; j < count; ++j) placements += new_slot[j] >= 0;
}

// Every animal's records at once (checks first, then the pool is rebuilt).
void WorldPool::import_records(const NRec *neurons, int neuron_count, const ns_checkpoint_synapse *synapses,
                               int synapse_count) {
    if (neuron_count < 0 || synapse_count < 0 || (neuron_count && !neurons) || (synapse_count && !synapses))
        throw std::runtime_error("Invalid record count");
    int prev = -1;
    for (int i = 0; i < neuron_count; ++i) {
        if (!valid(neurons[i]) || neurons[i].slot <= prev)
            throw std::runtime_error("Invalid checkpoint neuron record");
        prev = neurons[i].slot;
    }
    prev = -1;
    for (int i = 0; i < synapse_count; ++i) {
        if (!valid(synapses[i]) || synapses[i].slot <= prev)
            throw std::runtime_error("Invalid checkpoint synapse record");
        prev = synapses[i].slot;
    }
    // group by owner (records stay in ascending slot order per animal)
    std::vector<std::vector<NRec>> nr(P);
    std::vector<std::vector<ns_checkpoint_synapse>> sr(P);
    for (int i = 0; i < neuron_count; ++i) nr[neurons[i].owner].push_back(neurons[i]);
    for (int i = 0; i < synapse_count; ++i) {
        if (nr[synapses[i].owner].empty()) throw std::runtime_error("Synapse owner has no neurons");
        sr[synapses[i].owner].push_back(synapses[i]);
    }
    sync_all();
    for (int a = 0; a < P; ++a) {
        HostBrain hb = make_brain(nr[a], sr[a]);
        set_animal(a, std::move(nr[a]), std::move(sr[a]), std::move(hb));
    }
    rebuild();
}

// The record calls for either record layout R, which must be the pool's.
template <class R> WorldPool &pool_for(void *p) {
    if (!p) throw std::runtime_error("Invalid pool handle");
    auto &w = *static_cast<WorldPool *>(p);
    if (w.history != history_of<R>())
        throw std::runtime_error(w.history == 17 ? "Pool holds 17-slot neuron records"
                                                 : "Pool holds 9-slot neuron records");
    return w;
}

template <class R> std::vector<NRec> widen_all(const R *neurons, int count) {
    std::vector<NRec> out;
    if (count > 0 && neurons) {
        out.reserve(size_t(count));
        for (int i = 0; i < count; ++i) out.push_back(widen(neurons[i]));
    }
    return out;
}

template <class R>
int import_records(void *p, const R *neurons, int neuron_count, const ns_checkpoint_synapse *synapses,
                   int synapse_count) {
    try {
        auto &w = pool_for<R>(p);
        if (neuron_count < 0 || (neuron_count && !neurons)) throw std::runtime_error("Invalid record count");
        const std::vector<NRec> n = widen_all(neurons, neuron_count);
        w.import_records(n.data(), neuron_count, synapses, synapse_count);
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

template <class R>
int replace_records(void *p, int animal_count, const int *animals, const R *neurons, int neuron_count,
                    const ns_checkpoint_synapse *synapses, int synapse_count) {
    try {
        auto &w = pool_for<R>(p);
        if (neuron_count < 0 || (neuron_count && !neurons)) throw std::runtime_error("Invalid replacement");
        const std::vector<NRec> n = widen_all(neurons, neuron_count);
        w.replace(animal_count, animals, n.data(), neuron_count, synapses, synapse_count);
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

template <class R>
int export_records(void *p, R *neurons, int neuron_capacity, ns_checkpoint_synapse *synapses,
                   int synapse_capacity, int *neuron_count, int *synapse_count) {
    try {
        auto &w = pool_for<R>(p);
        if (!neuron_count || !synapse_count) throw std::runtime_error("Invalid count pointer");
        if (neuron_capacity < w.nrecords || synapse_capacity < w.srecords ||
            (w.nrecords && !neurons) || (w.srecords && !synapses))
            throw std::runtime_error("Record buffer too small");
        w.sync_all();
        w.pull();
        // ascending slot order over every animal's records
        std::vector<std::pair<int, const NRec *>> on;
        std::vector<std::pair<int, const ns_checkpoint_synapse *>> os;
        on.reserve(size_t(w.nrecords));
        os.reserve(size_t(w.srecords));
        for (int a = 0; a < w.P; ++a) {
            for (const auto &r : w.nrec[a]) on.push_back({r.slot, &r});
            for (const auto &r : w.srec[a]) os.push_back({r.slot, &r});
        }
        std::sort(on.begin(), on.end());
        std::sort(os.begin(), os.end());
        for (size_t i = 0; i < on.size(); ++i) narrow(*on[i].second, &neurons[i]);
        for (size_t i = 0; i < os.size(); ++i) std::memcpy(&synapses[i], os[i].second, sizeof(synapses[i]));
        *neuron_count = int(on.size());
        *synapse_count = int(os.size());
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}
}  // namespace

}  // namespace nsa

using namespace nsa;

extern "C" {

void *ns_pool_create_ex(int animal_capacity, int neuron_capacity, int synapse_capacity, int history) {
    try {
        if (animal_capacity < 1 || neuron_capacity < 1 || synapse_capacity < 0)
            throw std::runtime_error("Invalid pool capacity");
        if (history != kWorldHistory && history != kMaxHistory)
            throw std::runtime_error("Invalid ring length (9 or 17 slots)");
        auto *w = new WorldPool;
        w->history = history;
        w->P = animal_capacity;
        w->ncap = neuron_capacity;
        w->scap = synapse_capacity;
        try {
            w->nrec.resize(w->P);
            w->srec.resize(w->P);
            w->brains.resize(w->P);
            w->spec.assign(w->P, BrainSpec{nullptr, -1, 0, 0, true});
            w->real.assign(w->P, 0);
            w->nused.assign(w->ncap, 0);
            w->sused.assign(std::max(1, w->scap), 0);
            NS_HIP(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking));
            w->d_senses = w->io.alloc<float>(int64_t(w->P) * kSenses);
            w->d_out = w->io.alloc<float>(int64_t(w->P) * kMotors);
        } catch (...) {
            delete w;
            throw;
        }
        return w;
    } catch (const std::exception &e) {
        set_error(e.what());
        return nullptr;
    }
}

void *ns_pool_create(int animal_capacity, int neuron_capacity, int synapse_capacity) {
    return ns_pool_create_ex(animal_capacity, neuron_capacity, synapse_capacity, kWorldHistory);
}

int ns_pool_history(void *p) { return p ? static_cast<WorldPool *>(p)->history : -1; }

void ns_pool_destroy(void *p) { delete static_cast<WorldPool *>(p); }

int ns_pool_import_records(void *p, const ns_checkpoint_neuron *neurons, int neuron_count,
                           const ns_checkpoint_synapse *synapses, int synapse_count) {
    return import_records(p, neurons, neuron_count, synapses, synapse_count);
}

int ns_pool_import_records17(void *p, const ns_checkpoint_neuron17 *neurons, int neuron_count,
                             const ns_checkpoint_synapse *synapses, int synapse_count) {
    return import_records(p, neurons, neuron_count, synapses, synapse_count);
}

int ns_pool_reserve(void *p, float slack, float spare) {
    try {
        auto &w = *static_cast<WorldPool *>(p);
        if (!(slack >= 0 && slack <= 4) || !(spare >= 0 && spare <= 4))
            throw std::runtime_error("Invalid slot reserve");
        w.sync_all();
        w.pull();
        w.slack = slack;
        w.spare = spare;
        if (w.nreal > 0) w.rebuild();
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

int ns_pool_replace_records(void *p, int animal_count, const int *animals,
                            const ns_checkpoint_neuron *neurons, int neuron_count,
                            const ns_checkpoint_synapse *synapses, int synapse_count) {
    return replace_records(p, animal_count, animals, neurons, neuron_count, synapses, synapse_count);
}

int ns_pool_replace_records17(void *p, int animal_count, const int *animals,
                              const ns_checkpoint_neuron17 *neurons, int neuron_count,
                              const ns_checkpoint_synapse *synapses, int synapse_count) {
    return replace_records(p, animal_count, animals, neurons, neuron_count, synapses, synapse_count);
}

int ns_pool_info(void *p, ns_pool_stats *info) {
    try {
        auto &w = *static_cast<WorldPool *>(p);
        if (!info) throw std::runtime_error("Invalid info pointer");
        ns_pool_stats s{};
        s.animals = w.nreal;
        s.neuron_records = w.nrecords;
        s.synapse_records = w.srecords;
        s.slots = w.built ? w.pool.nb : 0;
        for (int pb = 0; pb < s.slots; ++pb) s.free_slots += w.pool.batch_of[pb] < 0;
        s.placements = w.placements;
        s.relayouts = w.relayouts;
        *info = s;
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

int ns_pool_export_records(void *p, ns_checkpoint_neuron *neurons, int neuron_capacity,
                           ns_checkpoint_synapse *synapses, int synapse_capacity, int *neuron_count,
                           int *synapse_count) {
    return export_records(p, neurons, neuron_capacity, synapses, synapse_capacity, neuron_count,
                          synapse_count);
}

int ns_pool_export_records17(void *p, ns_checkpoint_neuron17 *neurons, int neuron_capacity,
                             ns_checkpoint_synapse *synapses, int synapse_capacity, int *neuron_count,
                             int *synapse_count) {
    return export_records(p, neurons, neuron_capacity, synapses, synapse_capacity, neuron_count,
                          synapse_count);
}

int ns_world_neural_step(void *p, int64_t first_tick, const float *senses, int *motors,
                         const uint8_t *frozen) {
    try {
        auto &w = *static_cast<WorldPool *>(p);
        if (first_tick < 0) throw std::runtime_error("Invalid world tick");
        if (w.nreal == 0) return 0;
        w.sync_all();
        w.align_rings(first_tick);
        w.set_flags(frozen, w.stream);
        NS_HIP(hipMemcpy(w.d_senses, senses, sizeof(float) * w.P * kSenses, hipMemcpyHostToDevice));
        run_pool(w.pool, w.form, kTicksPerStep, w.d_senses, 230.0f, w.d_out, int64_t(w.P) * kMotors,
                 0, w.stream);
        NS_HIP(hipGetLastError());
        std::vector<float> out(size_t(w.P) * kMotors);
        NS_HIP(hipMemcpyAsync(out.data(), w.d_out, sizeof(float) * out.size(), hipMemcpyDeviceToHost,
                              w.stream));
        w.sync_all();
        for (int a = 0; a < w.P; ++a)
            if (w.real[a])
                for (int m = 0; m < kMotors; ++m) motors[a * kMotors + m] += int(out[a * kMotors + m]);
        w.next_tick = first_tick + kTicksPerStep;
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

int ns_world_neural_step_device(void *p, int64_t first_tick, const float *senses, int *motors,
                                const uint8_t *frozen, void *stream) {
    try {
        auto &w = *static_cast<WorldPool *>(p);
        if (first_tick < 0 || !senses || !motors) throw std::runtime_error("Invalid world step");
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : w.stream;
        if (w.nreal == 0) return 0;
        // a previous step queued on another stream finishes first
        if (w.last_stream && w.last_stream != st) NS_HIP(hipStreamSynchronize(w.last_stream));
        if (w.rings_pending || first_tick != w.next_tick) {
            NS_HIP(hipStreamSynchronize(st));
            w.sync_all();
            w.align_rings(first_tick);
        }
        w.set_flags(frozen, st);
        run_pool(w.pool, w.form, kTicksPerStep, senses, 230.0f, w.d_out, int64_t(w.P) * kMotors, 0, st);
        launch_add_counts(motors, w.d_out, int64_t(w.P) * kMotors, st);
        NS_HIP(hipGetLastError());
        w.next_tick = first_tick + kTicksPerStep;
        w.last_stream = st;
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

int ns_world_homeostasis(void *p, const int *age, const uint8_t *frozen) {
    try {
        auto &w = *static_cast<WorldPool *>(p);
        if (w.nreal == 0) return 0;
        w.sync_all();
        w.set_flags(frozen, w.stream);
        std::vector<int32_t> a(w.pool.nb);
        for (int pb = 0; pb < w.pool.nb; ++pb) a[pb] = w.pool.batch_of[pb] >= 0 ? age[w.pool.batch_of[pb]] : 0;
        if (w.age_cap < w.pool.nb) {          // slots (spares included) can outnumber animals
            w.age_io.release();
            w.d_age = w.age_io.alloc<int32_t>(w.pool.nb);
            w.age_cap = w.pool.nb;
        }
        NS_HIP(hipMemcpy(w.d_age, a.data(), sizeof(int32_t) * a.size(), hipMemcpyHostToDevice));
        launch_world_homeostasis(w.pool.view, w.pool.nn, w.d_age, w.stream);
        NS_HIP(hipGetLastError());
        w.sync_all();
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

}  // extern "C"
