// World-pool surface of libns_amd.so: the neural part of src/main.cu's step
// loop over checkpointable neuron/synapse records.
//
//   records in  -> ns_pool_import_records  (checkpoint_impl.cuh:272-311 checks)
//   16 ticks    -> ns_world_neural_step    (main.cu:2702-2728, rules 659-801)
//   homeostasis -> ns_world_homeostasis    (main.cu:1287-1300)
//   records out -> ns_pool_export_records  (main.cu:197-264, ascending slots)
//
// Each animal's records become one brain segment: neurons in ascending slot
// order, synapses sorted by (post, slot).  The reference accumulates pending
// conductances with float atomics (order-nondeterministic); here the sum runs
// in ascending synapse-slot order per post, a fixed order.
#include <algorithm>
#include <cstring>
#include <map>
#include <unordered_map>
#include <memory>

#include "ns_host.h"

namespace nsa {

namespace {
constexpr int kSenses = 19, kMotors = 5;   // src/main.cu:40-41 (S, M)
constexpr int kTicksPerStep = 16;          // src/main.cu:47

struct WorldPool {
    int P, ncap, scap;
    std::vector<ns_checkpoint_neuron> nrec;   // imported records (static parts)
    std::vector<ns_checkpoint_synapse> srec;
    std::vector<HostBrain> brains;            // one per animal with records
    std::vector<int> animal_of;               // brain -> animal
    std::vector<std::vector<int>> nidx, sidx; // brain -> record indices (genome order)
    std::vector<BrainSpec> specs;
    DevicePool pool;
    DeviceArena io;
    float *d_senses = nullptr, *d_out = nullptr;
    int32_t *d_age = nullptr;
    hipStream_t stream = nullptr;
    int64_t next_tick = -1;                   // tick the device rings are aligned to
    bool rings_pending = true;                // records' rings not yet on the device
    int form = NS_FORM_RESIDENT;

    ~WorldPool() {
        if (stream) (void)hipStreamSynchronize(stream);
        pool.arena.release();
        io.release();
        if (stream) (void)hipStreamDestroy(stream);
    }

    // ring slot k (= tick % 9) <-> shift-register bit (T-1-tick)
    static uint32_t ring_to_bits(const uint8_t *ring, int64_t T) {
        uint32_t bits = 0;
        for (int j = 0; j < kWorldHistory; ++j) {
            const int64_t tau = T - 1 - j;
            if (ring[((tau % kWorldHistory) + kWorldHistory) % kWorldHistory]) bits |= 1u << j;
        }
        return bits;
    }
    static void bits_to_ring(uint32_t bits, int64_t T, uint8_t *ring) {
        for (int k = 0; k < kWorldHistory; ++k) {
            const int64_t last = T - 1;
            const int64_t back = ((last - k) % kWorldHistory + kWorldHistory) % kWorldHistory;
            ring[k] = uint8_t((bits >> back) & 1u);
        }
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
            for (size_t b = 0; b < brains.size(); ++b)
                for (size_t k = 0; k < nidx[b].size(); ++k, ++off)
                    h[off] = ring_to_bits(nrec[nidx[b][k]].spike_history, T);
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
            uint8_t ring[kWorldHistory];
            for (auto &bits : h) {
                bits_to_ring(bits, next_tick, ring);
                bits = ring_to_bits(ring, T);
            }
            std::fill(ticks.begin(), ticks.end(), T);
            pool.upload_state(n, s, h, ticks, &st);
        }
        next_tick = T;
    }

    void set_flags(const uint8_t *frozen) {
        std::vector<uint8_t> run(pool.nb), fz(pool.nb);
        for (int pb = 0; pb < pool.nb; ++pb) {
            run[pb] = uint8_t(1 | (pool.fallback[pb] ? 2 : 0));
            fz[pb] = frozen ? frozen[animal_of[pool.batch_of[pb]]] : 0;
        }
        NS_HIP(hipStreamSynchronize(stream));
        NS_HIP(hipMemcpy(pool.view.brun, run.data(), run.size(), hipMemcpyHostToDevice));
        NS_HIP(hipMemcpy(pool.view.bfrozen, fz.data(), fz.size(), hipMemcpyHostToDevice));
    }
};
}  // namespace

}  // namespace nsa

using namespace nsa;

extern "C" {

void *ns_pool_create(int animal_capacity, int neuron_capacity, int synapse_capacity) {
    try {
        if (animal_capacity < 1 || neuron_capacity < 1 || synapse_capacity < 0)
            throw std::runtime_error("Invalid pool capacity");
        auto *w = new WorldPool;
        w->P = animal_capacity;
        w->ncap = neuron_capacity;
        w->scap = synapse_capacity;
        try {
            NS_HIP(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking));
            w->d_senses = w->io.alloc<float>(int64_t(w->P) * kSenses);
            w->d_out = w->io.alloc<float>(int64_t(w->P) * kMotors);
            w->d_age = w->io.alloc<int32_t>(w->P);
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

void ns_pool_destroy(void *p) { delete static_cast<WorldPool *>(p); }

int ns_pool_import_records(void *p, const ns_checkpoint_neuron *neurons, int neuron_count,
                           const ns_checkpoint_synapse *synapses, int synapse_count) {
    try {
        auto &w = *static_cast<WorldPool *>(p);
        if (neuron_count < 0 || synapse_count < 0) throw std::runtime_error("Invalid record count");
        // checkpoint_impl.cuh:272-311 record checks
        int prev = -1;
        for (int i = 0; i < neuron_count; ++i) {
            const auto &r = neurons[i];
            if (r.slot <= prev || r.slot >= w.ncap || r.owner < 0 || r.owner >= w.P ||
                r.next < -1 || r.next >= w.ncap || r.port < -1 || r.port >= kSenses + kMotors ||
                r.gene.type > 6 || r.gene.modulator > 3)
                throw std::runtime_error("Invalid checkpoint neuron record");
            prev = r.slot;
        }
        prev = -1;
        for (int i = 0; i < synapse_count; ++i) {
            const auto &r = synapses[i];
            if (r.slot <= prev || r.slot >= w.scap || r.owner < 0 || r.owner >= w.P ||
                r.next < -1 || r.next >= w.scap || r.gene.pre < 0 || r.gene.pre >= w.ncap ||
                r.gene.post < 0 || r.gene.post >= w.ncap || r.gene.delay < 1 ||
                r.gene.delay > kWorldHistory - 1 || r.gene.branch > 3)
                throw std::runtime_error("Invalid checkpoint synapse record");
            prev = r.slot;
        }
        w.nrec.assign(neurons, neurons + neuron_count);
        w.srec.assign(synapses, synapses + synapse_count);
        // group by owner; slot -> (animal, local index)
        std::map<int, int> brain_of_animal;
        std::vector<std::pair<int, int>> slot_local;   // (slot, local) for neurons, sorted by slot
        w.brains.clear(); w.animal_of.clear(); w.nidx.clear(); w.sidx.clear();
        for (int i = 0; i < neuron_count; ++i) {
            const int a = neurons[i].owner;
            auto it = brain_of_animal.find(a);
            if (it == brain_of_animal.end()) {
                it = brain_of_animal.emplace(a, int(w.brains.size())).first;
                w.brains.emplace_back();
                w.animal_of.push_back(a);
                w.nidx.emplace_back();
                w.sidx.emplace_back();
            }
            w.nidx[it->second].push_back(i);
        }
        std::unordered_map<int, std::pair<int, int>> where;   // neuron slot -> (brain, local)
        where.reserve(size_t(neuron_count) * 2);
        for (size_t b = 0; b < w.brains.size(); ++b) {
            HostBrain &hb = w.brains[b];
            hb.channels = kSenses;
            hb.max_delay = kWorldHistory - 1;
            hb.motors.assign(kMotors, -1);
            for (size_t k = 0; k < w.nidx[b].size(); ++k) {
                const auto &r = neurons[w.nidx[b][k]];
                where[r.slot] = {int(b), int(k)};
                hb.genes.push_back(r.gene);
                const int ch = r.gene.sensory_channel;
                const bool sensed = ch >= 0 && ch < kSenses;          // main.cu:729
                hb.sensory.push_back(sensed ? ch : -1);
                hb.receptor.push_back(sensed && r.gene.type == 5 && ch < 6);   // main.cu:735
                if (r.port >= kSenses && r.port < kSenses + kMotors) hb.motors[r.port - kSenses] = int(k);
            }
        }
        for (int i = 0; i < synapse_count; ++i) {
            const auto &r = synapses[i];
            auto it = brain_of_animal.find(r.owner);
            if (it == brain_of_animal.end()) throw std::runtime_error("Synapse owner has no neurons");
            const int b = it->second;
            ns_synapse_gene g = r.gene;
            auto pre = where.find(g.pre), post = where.find(g.post);
            // main.cu:702-704: a synapse whose endpoints are not owned by its
            // animal never delivers or learns; it keeps its state unchanged.
            if (pre == where.end() || post == where.end() || pre->second.first != b ||
                post->second.first != b) {
                g.pre = 0; g.post = 0; g.active = 0;
            } else {
                g.pre = pre->second.second;
                g.post = post->second.second;
            }
            w.brains[b].edges.push_back(g);
            w.sidx[b].push_back(i);
        }
        w.specs.clear();
        for (size_t b = 0; b < w.brains.size(); ++b) {
            const int a = w.animal_of[b];
            w.specs.push_back(BrainSpec{&w.brains[b], int(b), a * kSenses, a * kMotors, true});
        }
        Tile tile{256, 2, 4};
        w.pool.build(w.specs, kMotors, tile, w.stream);
        // upload record state
        std::vector<ns_neuron_state> ns;
        std::vector<ns_synapse_state> ss;
        std::vector<uint32_t> h;
        std::vector<uint64_t> st;
        for (size_t b = 0; b < w.brains.size(); ++b) {
            for (int i : w.nidx[b]) {
                ns.push_back(neurons[i].state);
                h.push_back(0);
                for (int e = 0; e < 6; ++e) st.push_back(neurons[i].mod_event_tick[e]);
            }
            for (int i : w.sidx[b]) ss.push_back(synapses[i].state);
        }
        std::vector<int64_t> ticks(w.brains.size(), 0);
        w.pool.upload_state(ns, ss, h, ticks, &st);
        w.rings_pending = true;
        w.next_tick = -1;
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

int ns_pool_export_records(void *p, ns_checkpoint_neuron *neurons, int neuron_capacity,
                           ns_checkpoint_synapse *synapses, int synapse_capacity, int *neuron_count,
                           int *synapse_count) {
    try {
        auto &w = *static_cast<WorldPool *>(p);
        if (neuron_capacity < int(w.nrec.size()) || synapse_capacity < int(w.srec.size()))
            throw std::runtime_error("Record buffer too small");
        std::vector<ns_neuron_state> ns;
        std::vector<ns_synapse_state> ss;
        std::vector<uint32_t> h;
        std::vector<int64_t> ticks;
        std::vector<uint64_t> st;
        if (!w.brains.empty()) w.pool.download_state(ns, ss, h, ticks, &st);
        std::vector<ns_checkpoint_neuron> nout = w.nrec;
        std::vector<ns_checkpoint_synapse> sout = w.srec;
        int64_t off = 0, soff = 0;
        for (size_t b = 0; b < w.brains.size(); ++b) {
            for (size_t k = 0; k < w.nidx[b].size(); ++k, ++off) {
                auto &r = nout[w.nidx[b][k]];
                r.state = ns[off];
                for (int e = 0; e < 6; ++e) r.mod_event_tick[e] = st[off * 6 + e];
                if (!w.rings_pending) WorldPool::bits_to_ring(h[off], w.next_tick, r.spike_history);
            }
            for (size_t k = 0; k < w.sidx[b].size(); ++k, ++soff) sout[w.sidx[b][k]].state = ss[soff];
        }
        std::memcpy(neurons, nout.data(), nout.size() * sizeof(nout[0]));
        if (!sout.empty()) std::memcpy(synapses, sout.data(), sout.size() * sizeof(sout[0]));
        *neuron_count = int(nout.size());
        *synapse_count = int(sout.size());
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

int ns_world_neural_step(void *p, int64_t first_tick, const float *senses, int *motors,
                         const uint8_t *frozen) {
    try {
        auto &w = *static_cast<WorldPool *>(p);
        if (first_tick < 0) throw std::runtime_error("Invalid world tick");
        if (w.brains.empty()) return 0;
        w.align_rings(first_tick);
        w.set_flags(frozen);
        NS_HIP(hipMemcpy(w.d_senses, senses, sizeof(float) * w.P * kSenses, hipMemcpyHostToDevice));
        run_pool(w.pool, w.form, kTicksPerStep, w.d_senses, 230.0f, w.d_out, int64_t(w.P) * kMotors,
                 0, w.stream);
        NS_HIP(hipGetLastError());
        std::vector<float> out(size_t(w.P) * kMotors);
        NS_HIP(hipMemcpyAsync(out.data(), w.d_out, sizeof(float) * out.size(), hipMemcpyDeviceToHost,
                              w.stream));
        NS_HIP(hipStreamSynchronize(w.stream));
        for (int a : w.animal_of)
            for (int m = 0; m < kMotors; ++m) motors[a * kMotors + m] += int(out[a * kMotors + m]);
        w.next_tick = first_tick + kTicksPerStep;
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

int ns_world_homeostasis(void *p, const int *age, const uint8_t *frozen) {
    try {
        auto &w = *static_cast<WorldPool *>(p);
        if (w.brains.empty()) return 0;
        w.set_flags(frozen);
        std::vector<int32_t> a(w.pool.nb);
        for (int pb = 0; pb < w.pool.nb; ++pb) a[pb] = age[w.animal_of[w.pool.batch_of[pb]]];
        NS_HIP(hipMemcpy(w.d_age, a.data(), sizeof(int32_t) * a.size(), hipMemcpyHostToDevice));
        launch_world_homeostasis(w.pool.view, w.pool.nn, w.d_age, w.stream);
        NS_HIP(hipGetLastError());
        NS_HIP(hipStreamSynchronize(w.stream));
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

}  // extern "C"
