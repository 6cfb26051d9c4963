// Serving surface of libns_amd.so: continuous batching of genome-evaluation
// episodes (no reference counterpart; gym_adapter/batch_evaluate.py runs
// lock-step batches with masks instead).
//
// A request is one evaluation of a brain genome (ns_create encoding) on the
// server's task:
//  * linear (task 0): a seeded 4-d linear system
//        x' = A x + B u + sigma * n,   u_j = motor_j spikes / ticks per step,
//    observed through an opponent rate code on 8 channels (2 per dimension),
//    2 motors.  Return = -sum over steps of |x|^2.
//  * mazes (task 1): `maze_episodes` seeded 31 x 31 depth-first mazes in a
//    row, the brain's state kept from one to the next.  9 channels: walls
//    north/east/south/west of the agent, goal east/west/south/north of it,
//    and reward; 4 motors move north/east/south/west (most spikes wins, ties
//    to the lower motor, none = stay).  A maze ends at the goal or after
//    `maze_horizon` steps; reaching the goal after k steps pays
//    1 - 0.9 k / horizon (gym_adapter/maze_gpu.cu:46) and sets the reward
//    channel for the next step, so a dopamine modulator neuron listening on
//    it delivers the reward as a modulatory event.
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
constexpr int kChannels = 8, kOutputs = 2, kDims = 4;            // linear task
constexpr int kMazeChannels = 9, kMazeOutputs = 4;               // maze task
constexpr int kTaskWords = 64;   // task record per slot (4-byte words)
// linear record (floats): x[4], A[16], B[8], return, motor totals[2], steps
enum { T_X = 0, T_A = 4, T_B = 20, T_RET = 28, T_M0 = 29, T_M1 = 30, T_STEPS = 31 };
// maze record (u32 words; return and motor totals hold float bits): open
// cells per row, agent, goal, maze index, steps in this maze, steps in all,
// mazes solved, reward pending, return, motor totals[4], request seed
enum {
    M_ROWS = 0, M_AX = 31, M_AY, M_GX, M_GY, M_INDEX, M_STEPS, M_TOTAL, M_SOLVED, M_PENDING,
    M_RET, M_MOT, M_SEED = M_MOT + 4
};
constexpr int kMazeSide = 31, kMazeRooms = 15;

__global__ void k_encode(const float *task, const uint8_t *run, int slots, float *input) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= slots || !(run[e] & 1)) return;
    const float *x = task + size_t(e) * kTaskWords + T_X;
    for (int d = 0; d < kDims; ++d) {
        input[e * kChannels + 2 * d] = fminf(1.0f, fmaxf(0.0f, x[d]));
        input[e * kChannels + 2 * d + 1] = fminf(1.0f, fmaxf(0.0f, -x[d]));
    }
}

__host__ __device__ inline uint64_t splitmix64(uint64_t &z) {
    z += 0x9e3779b97f4a7c15ull;
    uint64_t r = z;
    r = (r ^ (r >> 30)) * 0xbf58476d1ce4e5b9ull;
    r = (r ^ (r >> 27)) * 0x94d049bb133111ebull;
    return r ^ (r >> 31);
}

// Maze `index` of a request seeded `seed`: a depth-first (recursive
// backtracker) carve of the 15 x 15 rooms of a 31 x 31 grid, from room 0 at
// (1, 1), neighbours tried in the order north, east, south, west and one
// drawn uniformly (SplitMix64 mod count); then the goal room, drawn from the
// other 224.  rows[y] bit x = cell (x, y) open.
__host__ __device__ inline void make_maze(uint32_t seed, uint32_t index, uint32_t *rows,
                                          uint32_t &gx, uint32_t &gy) {
    uint64_t z = (0x9e3779b97f4a7c15ull * (uint64_t(seed) + 1)) ^ (uint64_t(index) * 0xd1b54a32d192ed03ull);
    for (int y = 0; y < kMazeSide; ++y) rows[y] = 0;
    uint32_t visited[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint8_t stack[kMazeRooms * kMazeRooms];
    int top = 0;
    rows[1] |= 1u << 1;
    visited[0] = 1;
    stack[top++] = 0;
    while (top > 0) {
        const int r = stack[top - 1], rx = r % kMazeRooms, ry = r / kMazeRooms;
        int cand[4], nc = 0;
        auto fresh = [&](int q) { return !((visited[q >> 5] >> (q & 31)) & 1u); };
        if (ry > 0 && fresh(r - kMazeRooms)) cand[nc++] = r - kMazeRooms;
        if (rx < kMazeRooms - 1 && fresh(r + 1)) cand[nc++] = r + 1;
        if (ry < kMazeRooms - 1 && fresh(r + kMazeRooms)) cand[nc++] = r + kMazeRooms;
        if (rx > 0 && fresh(r - 1)) cand[nc++] = r - 1;
        if (!nc) { --top; continue; }
        const int q = cand[splitmix64(z) % uint64_t(nc)];
        const int ax = 1 + 2 * rx, ay = 1 + 2 * ry;
        const int bx = 1 + 2 * (q % kMazeRooms), by = 1 + 2 * (q / kMazeRooms);
        rows[(ay + by) / 2] |= 1u << ((ax + bx) / 2);
        rows[by] |= 1u << bx;
        visited[q >> 5] |= 1u << (q & 31);
        stack[top++] = uint8_t(q);
    }
    const int g = 1 + int(splitmix64(z) % uint64_t(kMazeRooms * kMazeRooms - 1));
    gx = uint32_t(1 + 2 * (g % kMazeRooms));
    gy = uint32_t(1 + 2 * (g / kMazeRooms));
}

__host__ __device__ inline bool maze_open(const uint32_t *rows, int x, int y) {
    return x >= 0 && x < kMazeSide && y >= 0 && y < kMazeSide && ((rows[y] >> x) & 1u);
}

__global__ void k_encode_maze(const uint32_t *task, const uint8_t *run, int slots, float *input) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= slots || !(run[e] & 1)) return;
    const uint32_t *t = task + size_t(e) * kTaskWords;
    const int x = int(t[M_AX]), y = int(t[M_AY]), gx = int(t[M_GX]), gy = int(t[M_GY]);
    float *in = input + size_t(e) * kMazeChannels;
    in[0] = maze_open(t + M_ROWS, x, y - 1) ? 0.0f : 1.0f;
    in[1] = maze_open(t + M_ROWS, x + 1, y) ? 0.0f : 1.0f;
    in[2] = maze_open(t + M_ROWS, x, y + 1) ? 0.0f : 1.0f;
    in[3] = maze_open(t + M_ROWS, x - 1, y) ? 0.0f : 1.0f;
    in[4] = gx > x ? 1.0f : 0.0f;
    in[5] = gx < x ? 1.0f : 0.0f;
    in[6] = gy > y ? 1.0f : 0.0f;
    in[7] = gy < y ? 1.0f : 0.0f;
    in[8] = t[M_PENDING] ? 1.0f : 0.0f;
}

// One maze step per live slot (see the header for the rules); a request
// whose last maze ended sets its done flag and stops changing.
__global__ void k_env_maze(uint32_t *task, const float *out, const uint8_t *run, int slots,
                           uint32_t mazes, uint32_t horizon, uint8_t *done) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= slots || !(run[e] & 1)) return;
    uint32_t *t = task + size_t(e) * kTaskWords;
    if (t[M_INDEX] >= mazes) return;
    t[M_PENDING] = 0;
    const float *o = out + size_t(e) * kMazeOutputs;
    int best = -1;
    float top = 0.0f;
    for (int m = 0; m < kMazeOutputs; ++m) {
        if (o[m] > top) { top = o[m]; best = m; }
        t[M_MOT + m] = __float_as_uint(__uint_as_float(t[M_MOT + m]) + o[m]);
    }
    int x = int(t[M_AX]), y = int(t[M_AY]);
    if (best >= 0) {
        const int nx = x + (best == 1) - (best == 3), ny = y + (best == 2) - (best == 0);
        if (maze_open(t + M_ROWS, nx, ny)) { x = nx; y = ny; }
    }
    t[M_AX] = uint32_t(x);
    t[M_AY] = uint32_t(y);
    t[M_STEPS] += 1;
    t[M_TOTAL] += 1;
    bool next = false;
    if (x == int(t[M_GX]) && y == int(t[M_GY])) {
        const float reward = float(1.0 - 0.9 * (double(t[M_STEPS]) / double(horizon)));
        t[M_RET] = __float_as_uint(__uint_as_float(t[M_RET]) + reward);
        t[M_SOLVED] += 1;
        t[M_PENDING] = 1;
        next = true;
    } else if (t[M_STEPS] >= horizon) {
        next = true;
    }
    if (!next) return;
    t[M_INDEX] += 1;
    if (t[M_INDEX] >= mazes) {
        done[e] = 1;
        return;
    }
    uint32_t gx, gy;
    make_maze(t[M_SEED], t[M_INDEX], t + M_ROWS, gx, gy);
    t[M_GX] = gx;
    t[M_GY] = gy;
    t[M_AX] = 1;
    t[M_AY] = 1;
    t[M_STEPS] = 0;
}

__device__ inline float uniform01(uint32_t &s) {
    s ^= s << 13; s ^= s >> 17; s ^= s << 5;
    return (s >> 8) * (1.0f / 16777216.0f);
}

// x' = A x + B u + sigma * (sum of 12 uniforms - 6); fixed evaluation order.
// A slot stops updating once its episode has `steps` steps (done flag set).
__global__ void k_env(float *task, uint32_t *rng, const float *out, const uint8_t *run, int slots,
                      float ticks, float sigma, float steps, uint8_t *done) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= slots || !(run[e] & 1)) return;
    float *t = task + size_t(e) * kTaskWords;
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
    if (!(t[T_STEPS] < steps)) done[e] = 1;
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
    std::memset(t, 0, sizeof(float) * kTaskWords);
    for (int i = 0; i < kDims; ++i) t[T_X + i] = 2.0f * next() - 1.0f;
    for (int i = 0; i < kDims; ++i) {
        float row[kDims], norm = 0;
        for (int j = 0; j < kDims; ++j) { row[j] = 2.0f * next() - 1.0f; norm += std::fabs(row[j]); }
        for (int j = 0; j < kDims; ++j) t[T_A + i * kDims + j] = 0.9f * row[j] / norm;
    }
    for (int k = 0; k < 8; ++k) t[T_B + k] = next() - 0.5f;
}

// First maze of a maze request (later ones are carved on the device).
void make_maze_task(uint32_t seed, uint32_t *t) {
    std::memset(t, 0, sizeof(uint32_t) * kTaskWords);
    uint32_t gx, gy;
    make_maze(seed, 0, t + M_ROWS, gx, gy);
    t[M_GX] = gx;
    t[M_GY] = gy;
    t[M_AX] = 1;
    t[M_AY] = 1;
    t[M_SEED] = seed;
}

// An admitted episode in slot `pb`.
struct Episode {
    Request req;
    BrainSpec spec;
    int pb = 0;
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
    uint8_t *d_done = nullptr;                           // per slot: request finished
    // A launch round's outcome (done flags and task rows of every slot),
    // copied to pinned host memory behind the round on the serving stream;
    // double buffered, so that round r's copies are read while round r + 1
    // runs.  begin/end bracket the round's device work.
    struct RoundOut {
        uint8_t *done = nullptr;
        float *task = nullptr;
        hipEvent_t begin = nullptr, end = nullptr;
        int64_t round = -1;
    };
    RoundOut outs[2];
    std::vector<int64_t> admit_round;                    // by slot: the round its occupant started in
    int channels = kChannels, outputs = kOutputs;        // of the task
    bool maze = false;
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
        pool.build(specs, outputs, Tile{1024, 1, 0, 1, 4448}, stream);
        pool.nbatch = pool.nb;                 // rows are slots
        pool.pool_of.assign(pool.nb, -1);
        const int rows = std::max(pool.nb, 1);
        d_input = io.alloc<float>(int64_t(rows) * channels);
        d_out = io.alloc<float>(int64_t(rows) * outputs);
        d_task = io.alloc<float>(int64_t(rows) * kTaskWords);
        d_rng = io.alloc<uint32_t>(rows);
        d_done = io.alloc<uint8_t>(rows);
        NS_HIP(hipMemset(d_task, 0, sizeof(float) * size_t(rows) * kTaskWords));
        NS_HIP(hipMemset(d_input, 0, sizeof(float) * size_t(rows) * channels));
        NS_HIP(hipMemset(d_done, 0, size_t(rows)));
        for (RoundOut &o : outs) {
            NS_HIP(hipHostMalloc(reinterpret_cast<void **>(&o.done), size_t(rows)));
            NS_HIP(hipHostMalloc(reinterpret_cast<void **>(&o.task), sizeof(float) * size_t(rows) * kTaskWords));
            NS_HIP(hipEventCreate(&o.begin));
            NS_HIP(hipEventCreate(&o.end));
        }
        admit_round.assign(pool.nb, 0);
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
            ep->spec = BrainSpec{&ep->req.brain, pb, pb * channels, pb * outputs, false};
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
        std::vector<float> task(size_t(std::max(k, 1)) * kTaskWords);
        std::vector<uint32_t> rng(std::max(k, 1));
        for (int j = 0; j < k; ++j) {
            const Episode &e = *fresh[j];
            born[j] = e.pb;
            born[k + j] = int32_t(e.req.seed);
            if (maze) make_maze_task(e.req.task_seed, reinterpret_cast<uint32_t *>(&task[j * kTaskWords]));
            else make_task(e.req.task_seed, &task[j * kTaskWords]);
            rng[j] = e.req.task_seed * 2654435761u + 12345u;
            if (!rng[j]) rng[j] = 1;
        }
        for (int pb = 0; pb < pool.nb; ++pb) next[pb] = live[pb] ? 1 : 0;
        if (k) stage.add(pool.d_reset, born.data(), 4u * born.size());
        static const uint8_t kNotDone = 0;
        for (int j = 0; j < k; ++j) {
            const int pb = fresh[j]->pb;
            stage.add(d_task + size_t(pb) * kTaskWords, &task[j * kTaskWords], 4u * kTaskWords);
            stage.add(d_rng + pb, &rng[j], 4);
            stage.add(d_done + pb, &kNotDone, 1);
        }
        stage.add(pool.view.brun, next.data(), size_t(pool.nb));
        // the staged vectors are copied into pinned memory by upload_stage
        pool.upload_stage(stage, stream);
        pool.scatter_range(0, int(stage.entries.size()), stream);
        // the reset clears the newborns' settled marks, which the slot order reads
        launch_reset_slots(pool.view, pool.d_reset, k, pool.d_norig, pool.d_sorig, stream);
        pool.order_touched(stream);
        NS_HIP(hipGetLastError());
        h_run = next;
    }

    void steps(int n) {
        const unsigned blocks = unsigned((pool.nb + 127) / 128);
        uint32_t *task_words = reinterpret_cast<uint32_t *>(d_task);
        const uint32_t mazes = uint32_t(cfg.maze_episodes > 0 ? cfg.maze_episodes : 10);
        const uint32_t horizon = uint32_t(cfg.maze_horizon > 0 ? cfg.maze_horizon : 100);
        for (int k = 0; k < n; ++k) {
            if (maze)
                hipLaunchKernelGGL(k_encode_maze, dim3(blocks), dim3(128), 0, stream, task_words,
                                   pool.view.brun, pool.nb, d_input);
            else
                hipLaunchKernelGGL(k_encode, dim3(blocks), dim3(128), 0, stream, d_task,
                                   pool.view.brun, pool.nb, d_input);
            run_pool(pool, NS_FORM_RESIDENT, cfg.ticks_per_step, d_input, cfg.input_gain, d_out,
                     int64_t(pool.nb) * outputs, 1, stream);
            if (maze)
                hipLaunchKernelGGL(k_env_maze, dim3(blocks), dim3(128), 0, stream, task_words, d_out,
                                   pool.view.brun, pool.nb, mazes, horizon, d_done);
            else
                hipLaunchKernelGGL(k_env, dim3(blocks), dim3(128), 0, stream, d_task, d_rng, d_out,
                                   pool.view.brun, pool.nb, float(cfg.ticks_per_step),
                                   cfg.noise_sigma, float(cfg.steps_per_episode), d_done);
        }
        NS_HIP(hipGetLastError());
    }

    // Collect the episodes a completed round finished (slots whose occupant
    // started in that round or before: a slot admitted since holds a new
    // episode, whose flag in the round's copy is its predecessor's) and free
    // their slots.  A finished episode's task row no longer changes, so the
    // round's copy holds its result even while later rounds still step its
    // brain.  Returns whether any slot was freed (its run flag must be
    // cleared).
    bool finish(const RoundOut &o) {
        std::vector<int> ended;
        for (int pb = 0; pb < pool.nb; ++pb)
            if (live[pb] && admit_round[pb] <= o.round && o.done[pb]) ended.push_back(pb);
        if (ended.empty()) return false;
        const float *task = o.task;
        const double t = now();
        std::vector<std::pair<int, const BrainSpec *>> vacate;
        for (int pb : ended) vacate.push_back({pb, nullptr});
        stage.clear();
        pool.place(vacate, stage);             // slot tables only
        std::lock_guard<std::mutex> lock(mu);
        for (int pb : ended) {
            const Episode &e = *live[pb];
            const float *tk = &task[size_t(pb) * kTaskWords];
            ns_serve_result r{};
            r.id = e.req.id;
            r.submit_s = e.req.submit_s;
            r.admit_s = e.req.admit_s;
            r.finish_s = t;
            if (maze) {
                const uint32_t *w = reinterpret_cast<const uint32_t *>(tk);
                float mot[2];
                std::memcpy(&r.episode_return, &w[M_RET], 4);
                std::memcpy(mot, &w[M_MOT], 8);
                r.final_state[0] = float(w[M_SOLVED]);
                r.final_state[1] = float(w[M_TOTAL]);
                r.final_state[2] = float(w[M_INDEX]);
                r.final_state[3] = 0.0f;
                r.motor_spikes[0] = int32_t(mot[0]);
                r.motor_spikes[1] = int32_t(mot[1]);
            } else {
                r.episode_return = tk[T_RET];
                for (int d = 0; d < kDims; ++d) r.final_state[d] = tk[T_X + d];
                r.motor_spikes[0] = int32_t(tk[T_M0]);
                r.motor_spikes[1] = int32_t(tk[T_M1]);
            }
            r.status = 0;
            done.push_back(r);
            freeset.insert({pool.bnc[pb], pool.bsc[pb], pb});
            live[pb].reset();
        }
        stats.finished += int64_t(ended.size());
        cv.notify_all();
        return true;
    }

    // Launch rounds back to back: while round r runs on the device, the host
    // collects round r - 1's finished episodes, admits queued requests into
    // the slots they freed and stages round r + 1's newcomers, so admission
    // and collection overlap device work.  A finished episode's brain may run
    // one round more (its run flag clears at the next upload); its result is
    // already fixed.
    void loop() {
        try {
            bool freed = false;
            int64_t round = 0;
            int inflight = -1;   // output buffer of the round in flight, or -1
            while (true) {
                {
                    std::unique_lock<std::mutex> lock(mu);
                    bool any_live = false;
                    for (const auto &e : live) any_live = any_live || bool(e);
                    if (!any_live && inflight < 0)
                        cv.wait(lock, [&] { return stop || !queue.empty(); });
                    if (stop) break;
                }
                std::vector<Episode *> fresh = admit();
                for (Episode *e : fresh) admit_round[e->pb] = round;
                start(fresh, freed);
                freed = false;
                int resident = 0;
                for (const auto &e : live) resident += bool(e);
                if (resident > stats.resident_peak) stats.resident_peak = resident;
                int launched = -1;
                if (resident) {
                    RoundOut &o = outs[round & 1];
                    NS_HIP(hipEventRecord(o.begin, stream));
                    steps(cfg.steps_per_launch);
                    NS_HIP(hipMemcpyAsync(o.done, d_done, size_t(pool.nb), hipMemcpyDeviceToHost, stream));
                    NS_HIP(hipMemcpyAsync(o.task, d_task, sizeof(float) * size_t(pool.nb) * kTaskWords,
                                          hipMemcpyDeviceToHost, stream));
                    NS_HIP(hipEventRecord(o.end, stream));
                    o.round = round;
                    stats.launches += 1;
                    launched = int(round & 1);
                    ++round;
                }
                if (inflight >= 0) {
                    const RoundOut &o = outs[inflight];
                    NS_HIP(hipEventSynchronize(o.end));
                    float ms = 0;
                    NS_HIP(hipEventElapsedTime(&ms, o.begin, o.end));
                    stats.busy_s += 1e-3 * ms;
                    freed = finish(o);
                }
                inflight = launched;
            }
        } catch (const std::exception &e) {
            std::lock_guard<std::mutex> lock(mu);
            failure = e.what();
            stop = true;
            cv.notify_all();
        }
        if (stream) (void)hipStreamSynchronize(stream);
    }

    ~Server() {
        for (RoundOut &o : outs) {
            if (o.done) (void)hipHostFree(o.done);
            if (o.task) (void)hipHostFree(o.task);
            if (o.begin) (void)hipEventDestroy(o.begin);
            if (o.end) (void)hipEventDestroy(o.end);
        }
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
            config->max_episode_neurons < 0 || config->max_episode_neurons > (1 << 20) ||
            config->task < 0 || config->task > 1 || config->maze_episodes < 0 ||
            config->maze_horizon < 0)
            throw std::runtime_error("Invalid serving configuration");
        auto *s = new Server;
        s->cfg = *config;
        s->maze = config->task == 1;
        s->channels = s->maze ? kMazeChannels : kChannels;
        s->outputs = s->maze ? kMazeOutputs : kOutputs;
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
        std::unique_ptr<HostBrain> b(parse_gym_brain(count, edges, s.channels, s.outputs, ng, sg,
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

void ns_maze_layout(uint32_t seed, uint32_t index, uint32_t *rows, uint32_t *goal) {
    make_maze(seed, index, rows, goal[0], goal[1]);
}

int ns_serve_stats_get(void *p, ns_serve_stats *stats) {
    auto &s = *static_cast<Server *>(p);
    std::lock_guard<std::mutex> lock(s.mu);
    *stats = s.stats;
    return 0;
}

}  // extern "C"
