// Serving surface of libns_amd.so: continuous batching of genome-evaluation
// episodes (no reference counterpart; gym_adapter/batch_evaluate.py runs
// lock-step batches with masks instead).
//
// A request is one episode: a brain genome (ns_create encoding, 8 sensory
// channels, 2 motors) driving a seeded 4-d linear system
//     x' = A x + B u + sigma * n,   u_j = motor_j spikes / ticks per step,
// observed through an opponent rate code on 8 channels (2 per dimension).
// Return = -sum over steps of |x|^2.
//
// One device pool of slots serves every episode.  Slot capacities follow a
// ladder of size classes (ratio sqrt 2) with the same number of slots per
// class, within the configured neuron and synapse capacity.  At every launch
// boundary the serving thread admits queued requests into the best-fitting
// free slots (their images and fresh state written on the device, like a
// brain replacement), then advances every live episode by `steps_per_launch`
// environment steps: per step one encode, one resident tick launch over the
// whole pool and one environment update, all for every live slot at once.
// Finished episodes free their slots for the next boundary.  Each episode's
// arithmetic is independent of what else is in the pool (per-brain
// sequential sums, seeds per episode), so results are bit-identical alone
// and in any mix.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <set>
#include <thread>
#include <tuple>

#include "ns_host.h"

namespace nsa {

namespace {
constexpr int kChannels = 8, kOutputs = 2, kDims = 4;
// task record per slot (floats): x[4], A[16], B[8], return, motor totals[2], steps
constexpr int kTaskFloats = 32;
enum { T_X = 0, T_A = 4, T_B = 20, T_RET = 28, T_M0 = 29, T_M1 = 30, T_STEPS = 31 };

__global__ void k_encode(const float *task, const uint8_t *run, int slots, float *input) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= slots || !(run[e] & 1)) return;
    const float *x = task + size_t(e) * kTaskFloats + T_X;
    for (int d = 0; d < kDims; ++d) {
        input[e * kChannels + 2 * d] = fminf(1.0f, fmaxf(0.0f, x[d]));
        input[e * kChannels + 2 * d + 1] = fminf(1.0f, fmaxf(0.0f, -x[d]));
    }
}

__device__ inline float uniform01(uint32_t &s) {
    s ^= s << 13; s ^= s >> 17; s ^= s << 5;
    return (s >> 8) * (1.0f / 16777216.0f);
}

// x' = A x + B u + sigma * (sum of 12 uniforms - 6); fixed evaluation order.
// A slot stops updating once its episode has `steps` steps.
__global__ void k_env(float *task, uint32_t *rng, const float *out, const uint8_t *run, int slots,
                      float ticks, float sigma, float steps) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= slots || !(run[e] & 1)) return;
    float *t = task + size_t(e) * kTaskFloats;
    if (!(t[T_STEPS] < steps)) return;
    const float u0 = out[e * kOutputs] / ticks, u1 = out[e * kOutputs + 1] / ticks;
    uint32_t s = rng[e];
    float nx[kDims];
    for (int i = 0; i < kDims; ++i) {
        float acc = 0.0f;
        for (int j = 0; j < kDims; ++j) acc += t[T_A + i * kDims + j] * t[T_X + j];
        acc += t[T_B + i * 2] * u0;
        acc += t[T_B + i * 2 + 1] * u1;
        float n = -6.0f;
        for (int k = 0; k < 12; ++k) n += uniform01(s);
        nx[i] = acc + sigma * n;
    }
    float cost = 0.0f;
    for (int i = 0; i < kDims; ++i) {
        t[T_X + i] = nx[i];
        cost += nx[i] * nx[i];
    }
    t[T_RET] -= cost;
    t[T_M0] += out[e * kOutputs];
    t[T_M1] += out[e * kOutputs + 1];
    t[T_STEPS] += 1.0f;
    rng[e] = s;
}

struct Request {
    int64_t id;
    HostBrain brain;
    uint32_t seed, task_seed;
    double submit_s, admit_s = 0;
};

// Seeded stable task: rows of A scaled so |A|_inf = 0.9 (spectral radius
// < 1); B entries in [-0.5, 0.5]; x0 in [-1, 1].
void make_task(uint32_t seed, float *t) {
    uint64_t z = 0x9e3779b97f4a7c15ull * (uint64_t(seed) + 1);
    auto next = [&]() {   // SplitMix64
        z += 0x9e3779b97f4a7c15ull;
        uint64_t r = z;
        r = (r ^ (r >> 30)) * 0xbf58476d1ce4e5b9ull;
        r = (r ^ (r >> 27)) * 0x94d049bb133111ebull;
        r ^= r >> 31;
        return float(double(r >> 11) * (1.0 / 9007199254740992.0));
    };
    std::memset(t, 0, sizeof(float) * kTaskFloats);
    for (int i = 0; i < kDims; ++i) t[T_X + i] = 2.0f * next() - 1.0f;
    for (int i = 0; i < kDims; ++i) {
        float row[kDims], norm = 0;
        for (int j = 0; j < kDims; ++j) { row[j] = 2.0f * next() - 1.0f; norm += std::fabs(row[j]); }
        for (int j = 0; j < kDims; ++j) t[T_A + i * kDims + j] = 0.9f * row[j] / norm;
    }
    for (int k = 0; k < 8; ++k) t[T_B + k] = next() - 0.5f;
}

// An admitted episode in slot `pb`.
struct Episode {
    Request req;
    BrainSpec spec;
    int pb = 0;
    int steps_done = 0;
};

struct Server {
    ns_serve_config cfg;
    std::chrono::steady_clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Request> queue;
    std::vector<ns_serve_result> done;
    std::thread worker;
    bool stop = false;
    std::string failure;
    int64_t next_id = 0;
    ns_serve_stats stats{};
    // the slot pool (serving thread only)
    DevicePool pool;
    DeviceArena io;
    float *d_input = nullptr, *d_out = nullptr, *d_task = nullptr;
    uint32_t *d_rng = nullptr;
    hipStream_t stream = nullptr;
    std::vector<std::unique_ptr<Episode>> live;          // by slot
    std::set<std::tuple<int, int, int>> freeset;         // (neurons, synapses, slot)
    int max_n = 0, max_s = 0;                            // largest slot shape
    Stager stage;
    std::vector<uint8_t> h_run;                          // run flags uploaded last

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    // Slots: classes ceil(16 * sqrt2^k) up to the first class that holds
    // max_episode_neurons; an equal number of slots per class, at least one.
    void build_pool() {
        std::vector<int> cls;
        const double r = std::sqrt(2.0);
        const int top = cfg.max_episode_neurons > 0 ? cfg.max_episode_neurons : 4096;
        for (double c = 16.0;; c *= r) {
            cls.push_back(int(std::ceil(c)));
            if (cls.back() >= top) break;
        }
        const double syn_ratio = std::min(32.0, std::max(1.0, double(cfg.synapse_capacity) /
                                                                 cfg.neuron_capacity));
        int64_t per_class_n = 0;
        for (int c : cls) per_class_n += c;
        int per = std::max(1, std::min(int(cfg.episode_capacity / int(cls.size())),
                                       int(cfg.neuron_capacity / std::max<int64_t>(per_class_n, 1))));
        std::vector<BrainSpec> specs;
        for (int c : cls)
            for (int j = 0; j < per; ++j)
                specs.push_back(BrainSpec{nullptr, -1, 0, 0, false, c, int(std::ceil(c * syn_ratio))});
        max_n = cls.back();
        max_s = int(std::ceil(cls.back() * syn_ratio));
        pool.force_plastic = true;
        pool.build(specs, kOutputs, Tile{1024, 1, 0, 1, 3584}, stream);
        pool.nbatch = pool.nb;                 // rows are slots
        pool.pool_of.assign(pool.nb, -1);
        const int rows = std::max(pool.nb, 1);
        d_input = io.alloc<float>(int64_t(rows) * kChannels);
        d_out = io.alloc<float>(int64_t(rows) * kOutputs);
        d_task = io.alloc<float>(int64_t(rows) * kTaskFloats);
        d_rng = io.alloc<uint32_t>(rows);
        NS_HIP(hipMemset(d_task, 0, sizeof(float) * size_t(rows) * kTaskFloats));
        NS_HIP(hipMemset(d_input, 0, sizeof(float) * size_t(rows) * kChannels));
        live.clear();
        live.resize(pool.nb);
        for (int pb = 0; pb < pool.nb; ++pb)
            if (pool.slot_kind[pb] != 0) freeset.insert({pool.bnc[pb], pool.bsc[pb], pb});
        h_run.assign(pool.nb, 0);
    }

    // Best free slot for (n, s), or -1.
    int fit(int n, int s) {
        auto it = freeset.lower_bound({n, s, -1});
        while (it != freeset.end() && std::get<1>(*it) < s) ++it;
        if (it == freeset.end()) return -1;
        const int pb = std::get<2>(*it);
        freeset.erase(it);
        return pb;
    }

    // Admit queued requests into free slots (queue order; a request that
    // does not fit yet waits without blocking smaller ones behind it).
    std::vector<Episode *> admit() {
        std::vector<Episode *> fresh;
        std::lock_guard<std::mutex> lock(mu);
        for (auto it = queue.begin(); it != queue.end();) {
            const int n = int(it->brain.genes.size()), s = int(it->brain.edges.size());
            const int pb = fit(n, s);
            if (pb < 0) { ++it; continue; }
            auto ep = std::make_unique<Episode>();
            ep->req = std::move(*it);
            ep->req.admit_s = now();
            ep->pb = pb;
            ep->spec = BrainSpec{&ep->req.brain, pb, pb * kChannels, pb * kOutputs, false};
            fresh.push_back(ep.get());
            live[pb] = std::move(ep);
            it = queue.erase(it);
        }
        stats.admitted += int64_t(fresh.size());
        return fresh;
    }

    // Write the newcomers' images, fresh state, tasks and the run flags of
    // every occupied slot, in one staged upload.
    void start(const std::vector<Episode *> &fresh, bool flags_changed) {
        if (fresh.empty() && !flags_changed) return;
        stage.clear();
        std::vector<std::pair<int, const BrainSpec *>> moves;
        for (Episode *e : fresh) moves.push_back({e->pb, &e->spec});
        pool.place(moves, stage);
        const int k = int(fresh.size());
        std::vector<uint8_t> next(pool.nb, 0);
        std::vector<int32_t> born(2 * size_t(k));   // slots to reset, then their seeds
        std::vector<float> task(size_t(std::max(k, 1)) * kTaskFloats);
        std::vector<uint32_t> rng(std::max(k, 1));
        for (int j = 0; j < k; ++j) {
            const Episode &e = *fresh[j];
            born[j] = e.pb;
            born[k + j] = int32_t(e.req.seed);
            make_task(e.req.task_seed, &task[j * kTaskFloats]);
            rng[j] = e.req.task_seed * 2654435761u + 12345u;
            if (!rng[j]) rng[j] = 1;
        }
        for (int pb = 0; pb < pool.nb; ++pb) next[pb] = live[pb] ? 1 : 0;
        if (k) stage.add(pool.d_reset, born.data(), 4u * born.size());
        for (int j = 0; j < k; ++j) {
            const int pb = fresh[j]->pb;
            stage.add(d_task + size_t(pb) * kTaskFloats, &task[j * kTaskFloats], 4u * kTaskFloats);
            stage.add(d_rng + pb, &rng[j], 4);
        }
        stage.add(pool.view.brun, next.data(), size_t(pool.nb));
        // the staged vectors are copied into pinned memory by upload_stage
        pool.upload_stage(stage, stream);
        pool.scatter_range(0, int(stage.entries.size()), stream);
        pool.order_touched(stream);
        launch_reset_slots(pool.view, pool.d_reset, k, pool.d_norig, pool.d_sorig, stream);
        NS_HIP(hipGetLastError());
        h_run = next;
    }

    void steps(int n) {
        const unsigned blocks = unsigned((pool.nb + 127) / 128);
        for (int k = 0; k < n; ++k) {
            hipLaunchKernelGGL(k_encode, dim3(blocks), dim3(128), 0, stream, d_task, pool.view.brun,
                               pool.nb, d_input);
            run_pool(pool, NS_FORM_RESIDENT, cfg.ticks_per_step, d_input, cfg.input_gain, d_out,
                     int64_t(pool.nb) * kOutputs, 1, stream);
            hipLaunchKernelGGL(k_env, dim3(blocks), dim3(128), 0, stream, d_task, d_rng, d_out,
                               pool.view.brun, pool.nb, float(cfg.ticks_per_step), cfg.noise_sigma,
                               float(cfg.steps_per_episode));
        }
        NS_HIP(hipGetLastError());
    }

    // Collect finished episodes and free their slots.  Returns whether any
    // slot was freed (its run flag must be cleared).
    bool finish() {
        std::vector<int> ended;
        for (int pb = 0; pb < pool.nb; ++pb)
            if (live[pb] && live[pb]->steps_done >= cfg.steps_per_episode) ended.push_back(pb);
        if (ended.empty()) return false;
        std::vector<float> task(size_t(pool.nb) * kTaskFloats);
        NS_HIP(hipMemcpy(task.data(), d_task, sizeof(float) * task.size(), hipMemcpyDeviceToHost));
        const double t = now();
        std::vector<std::pair<int, const BrainSpec *>> vacate;
        for (int pb : ended) vacate.push_back({pb, nullptr});
        stage.clear();
        pool.place(vacate, stage);             // slot tables only
        std::lock_guard<std::mutex> lock(mu);
        for (int pb : ended) {
            const Episode &e = *live[pb];
            const float *tk = &task[size_t(pb) * kTaskFloats];
            ns_serve_result r{};
            r.id = e.req.id;
            r.submit_s = e.req.submit_s;
            r.admit_s = e.req.admit_s;
            r.finish_s = t;
            r.episode_return = tk[T_RET];
            for (int d = 0; d < kDims; ++d) r.final_state[d] = tk[T_X + d];
            r.motor_spikes[0] = int32_t(tk[T_M0]);
            r.motor_spikes[1] = int32_t(tk[T_M1]);
            r.status = 0;
            done.push_back(r);
            freeset.insert({pool.bnc[pb], pool.bsc[pb], pb});
            live[pb].reset();
        }
        stats.finished += int64_t(ended.size());
        cv.notify_all();
        return true;
    }

    void loop() {
        try {
            bool freed = false;
            while (true) {
                {
                    std::unique_lock<std::mutex> lock(mu);
                    bool any_live = false;
                    for (const auto &e : live) any_live = any_live || bool(e);
                    if (!any_live)
                        cv.wait(lock, [&] { return stop || !queue.empty(); });
                    if (stop) break;
                }
                const auto busy0 = std::chrono::steady_clock::now();
                std::vector<Episode *> fresh = admit();
                start(fresh, freed);
                int resident = 0;
                for (const auto &e : live) resident += bool(e);
                if (resident > stats.resident_peak) stats.resident_peak = resident;
                if (resident) {
                    const int n = cfg.steps_per_launch;
                    steps(n);
                    NS_HIP(hipStreamSynchronize(stream));
                    for (auto &e : live)
                        if (e) e->steps_done += n;
                    stats.launches += 1;
                }
                stats.busy_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - busy0).count();
                freed = finish();
            }
        } catch (const std::exception &e) {
            std::lock_guard<std::mutex> lock(mu);
            failure = e.what();
            stop = true;
            cv.notify_all();
        }
        if (stream) (void)hipStreamSynchronize(stream);
    }
};
}  // namespace

}  // namespace nsa

using namespace nsa;

extern "C" {

void *ns_serve_create(const ns_serve_config *config) {
    try {
        if (!config || config->neuron_capacity < 1 || config->synapse_capacity < 0 ||
            config->episode_capacity < 1 || config->steps_per_episode < 1 ||
            config->ticks_per_step < 1 || config->ticks_per_step > 10000 ||
            config->steps_per_launch < 1 || !(config->noise_sigma >= 0) || !(config->input_gain >= 0) ||
            config->max_episode_neurons < 0 || config->max_episode_neurons > (1 << 20))
            throw std::runtime_error("Invalid serving configuration");
        auto *s = new Server;
        s->cfg = *config;
        s->t0 = std::chrono::steady_clock::now();
        try {
            NS_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
            s->build_pool();
            s->worker = std::thread([s] { s->loop(); });
        } catch (...) {
            if (s->stream) (void)hipStreamDestroy(s->stream);
            delete s;
            throw;
        }
        return s;
    } catch (const std::exception &e) {
        set_error(e.what());
        return nullptr;
    }
}

void ns_serve_destroy(void *p) {
    auto *s = static_cast<Server *>(p);
    {
        std::lock_guard<std::mutex> lock(s->mu);
        s->stop = true;
        s->queue.clear();
    }
    s->cv.notify_all();
    if (s->worker.joinable()) s->worker.join();
    s->live.clear();
    hipStream_t st = s->stream;
    delete s;
    if (st) (void)hipStreamDestroy(st);
}

int64_t ns_serve_submit(void *p, int count, int edges, const float *ng, const float *sg,
                        const int *motors, uint32_t seed, uint32_t task_seed) {
    try {
        auto &s = *static_cast<Server *>(p);
        std::unique_ptr<HostBrain> b(parse_gym_brain(count, edges, kChannels, kOutputs, ng, sg,
                                                     motors, 8));
        if (count > s.max_n || edges > s.max_s)
            throw std::runtime_error("Episode exceeds the largest slot");
        for (const auto &g : b->genes)
            if (g.type == 2 || g.type == 3)
                throw std::runtime_error("LP/Prinz cells are not served (per-tick cells)");
        std::lock_guard<std::mutex> lock(s.mu);
        if (!s.failure.empty()) throw std::runtime_error("Server failed: " + s.failure);
        Request r;
        r.id = s.next_id++;
        r.brain = std::move(*b);
        r.seed = seed;
        r.task_seed = task_seed;
        r.submit_s = s.now();
        s.queue.push_back(std::move(r));
        s.stats.submitted += 1;
        s.cv.notify_all();
        return s.next_id - 1;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

int ns_serve_poll(void *p, ns_serve_result *out, int capacity) {
    auto &s = *static_cast<Server *>(p);
    std::lock_guard<std::mutex> lock(s.mu);
    if (!s.failure.empty()) {
        set_error("Server failed: " + s.failure);
        return -1;
    }
    const int n = std::min<int>(capacity, int(s.done.size()));
    std::copy(s.done.begin(), s.done.begin() + n, out);
    s.done.erase(s.done.begin(), s.done.begin() + n);
    return n;
}

int ns_serve_drain(void *p) {
    auto &s = *static_cast<Server *>(p);
    std::unique_lock<std::mutex> lock(s.mu);
    s.cv.wait(lock, [&] { return !s.failure.empty() || s.stats.finished == s.stats.submitted; });
    if (!s.failure.empty()) {
        set_error("Server failed: " + s.failure);
        return -1;
    }
    return 0;
}

int ns_serve_stats_get(void *p, ns_serve_stats *stats) {
    auto &s = *static_cast<Server *>(p);
    std::lock_guard<std::mutex> lock(s.mu);
    *stats = s.stats;
    return 0;
}

}  // extern "C"
