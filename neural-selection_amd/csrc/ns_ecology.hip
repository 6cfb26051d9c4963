// The world step around the neural ticks: senses in, motors out
// (src/main.cu update_senses 619-657, body_step 1146-1249, bite_animals /
// finish_predation 1251-1285, regrow 1138-1144, index_animals 612-617), for
// the C2 world: a toroidal grid (or the reference's clamped one), food kinds
// A/B/C at a density each with per-cell regrowth, animals whose brains are a
// world-rules batch.
//
// Device state is structure-of-arrays per animal plus the food and kind grids.
// The reference resolves contention with atomics in arrival order (who eats a
// cell's single food item, which prey an attacker bites first, whose bite
// kills); here every such choice is made in ascending animal index, so a step
// is deterministic and a CPU restatement reproduces it bit for bit:
//   * cells keep the ascending list of their living animals (count, scan,
//     place, sort), which the other-animal senses and the bites walk;
//   * eating is a claim (lowest index wins the cell's item), then the feed;
//   * bites run per cell, attackers in index order, each on the first living
//     other animal of its cell at its turn.
// The observers of body_step (behavior statistics, founder counters, bite
// event log) and reproduction are not part of this step.
#include <climits>
#include <cmath>

#include "ns_host.h"
#include "ns_trig.h"

namespace nsa {

void set_error(const std::string &msg);

namespace {

constexpr int kSenses = 19, kMotors = 5, kContactFood = 16;   // main.cu:40-42
constexpr int kTicks = 16;                                    // main.cu:47
constexpr float kBodyUpkeep = 0.003f, kNeuronUpkeep = 0.00002f;   // main.cu:53-57
constexpr float kBiteDamage = 2.0f, kBiteEnergyCost = 0.02f, kMeatFraction = 0.45f;
constexpr float kMuscleDecay = 0.92311635f, kMuscleTwitch = 0.30f;   // main.cu:60-61
constexpr float kPi = 3.14159265358979323846f;

__host__ __device__ inline float clampf(float x, float lo, float hi) { return fminf(hi, fmaxf(lo, x)); }

// main.cu:331-337 random64 / uniform, and seed_for (341-346)
__host__ __device__ inline unsigned long long random64(unsigned long long &s) {
    s ^= s >> 12; s ^= s << 25; s ^= s >> 27;
    return s * 2685821657736338717ULL;
}
__host__ __device__ inline float uniform01(unsigned long long &s) {
    return (float)((random64(s) >> 40) * (1.0 / 16777216.0));
}
__host__ __device__ inline unsigned long long seed_for(unsigned long long seed, unsigned long long i) {
    unsigned long long x = seed + 0x9e3779b97f4a7c15ULL * (i + 1);
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return (x ^ (x >> 31)) | 1ULL;
}

struct Eco {
    int P, W, H, torus, instant;
    float regrowth;
    unsigned long long seed;
    // animals
    float *x, *y, *heading, *energy, *health, *fullness, *pain, *taste, *muscle;   // taste 4P, muscle 3P
    int32_t *age, *alive, *active_neurons, *food_cell;
    uint8_t *eat, *bite;
    // grid
    int32_t *food, *eat_claim, *cell_count, *cell_start, *cell_fill, *cell_list;
    uint8_t *kind;
    float *senses, *motors;   // P x 19, P x 5
};

// main.cu:347-351 cell(), or the torus's wrapped cell
__device__ inline int cell_of(const Eco &e, float x, float y) {
    if (e.torus) {
        int ix = int(floorf(x)) % e.W, iy = int(floorf(y)) % e.H;
        if (ix < 0) ix += e.W;
        if (iy < 0) iy += e.H;
        return iy * e.W + ix;
    }
    const int ix = int(floorf(clampf(x, 0.0f, float(e.W) - 0.001f)));
    const int iy = int(floorf(clampf(y, 0.0f, float(e.H) - 0.001f)));
    return iy * e.W + ix;
}

__device__ inline float wrap_or_clamp(float v, int extent, int torus) {
    if (!torus) return clampf(v, 0.0f, float(extent) - 0.001f);
    if (v < 0.0f) v += float(extent);
    if (v >= float(extent)) v -= float(extent);
    return v;
}

// body_step (main.cu:1146-1213) up to the feed: motor forces, heading and
// position, upkeep and decays; an animal that eats claims its cell's item.
__global__ void k_eco_move(Eco e) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= e.P || e.alive[a] != 1) return;
    const float old_x = e.x[a], old_y = e.y[a];
    float force[3];
    for (int m = 0; m < 3; ++m) {
        const float spikes = e.motors[a * kMotors + m];
        if (e.instant) {
            force[m] = spikes / float(kTicks);
        } else {
            const float mu = fminf(1.0f, e.muscle[3 * a + m] * kMuscleDecay + spikes * kMuscleTwitch);
            e.muscle[3 * a + m] = mu;
            force[m] = mu;
        }
    }
    const float thrust = force[0], left = force[1], right = force[2];
    const bool eat = e.motors[a * kMotors + 3] > 0.0f;
    const bool bite = e.motors[a * kMotors + 4] > 0.0f;
    float heading = e.heading[a] + (right - left) * 1.2f;
    float s, c;
    world_sincos(heading, &s, &c);
    e.heading[a] = heading;
    e.x[a] = wrap_or_clamp(old_x + c * thrust * 1.5f, e.W, e.torus);
    e.y[a] = wrap_or_clamp(old_y + s * thrust * 1.5f, e.H, e.torus);
    e.energy[a] -= kBodyUpkeep + float(e.active_neurons[a]) * kNeuronUpkeep + thrust * 0.02f +
                   (bite ? kBiteEnergyCost : 0.0f);
    e.fullness[a] = fmaxf(0.0f, e.fullness[a] - 0.05f);
    e.pain[a] *= 0.8f;
    for (int k = 0; k < 4; ++k) e.taste[4 * a + k] *= 0.7f;
    // main.cu:1216: the animal feeds where it was before this step's move
    const int fc = cell_of(e, old_x, old_y);
    e.food_cell[a] = fc;
    e.eat[a] = eat;
    e.bite[a] = bite;
    if (eat) atomicMin(&e.eat_claim[fc], a);
}

// body_step's feed, starvation, ageing and death (main.cu:1226-1248): the
// lowest-index animal eating at a cell takes its item.
__global__ void k_eco_feed(Eco e) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= e.P || e.alive[a] != 1) return;
    float energy = e.energy[a], health = e.health[a], fullness = e.fullness[a], pain = e.pain[a];
    const int fc = e.food_cell[a];
    if (e.eat[a] && e.eat_claim[fc] == a && e.food[fc] > 0) {
        e.food[fc] -= 1;
        const int kind = e.kind[fc];
        e.taste[4 * a + kind] = 1.0f;
        fullness += 1.0f;
        if (kind == 0) energy = fminf(30.0f, energy + 4.0f);
        if (kind == 1) { energy = fminf(30.0f, energy + 2.0f); health -= 1.4f; pain += 1.4f; }
        if (fullness > 8.0f) { health -= 0.8f; pain += 0.8f; }
    }
    if (energy <= 0.0f) { energy = 0.0f; health -= 0.15f; pain += 0.15f; }
    const int age = e.age[a] + 1;
    e.age[a] = age;
    if (health <= 0.0f || age >= 10000) e.alive[a] = 0;
    e.energy[a] = energy;
    e.health[a] = health;
    e.fullness[a] = fullness;
    e.pain[a] = pain;
}

__global__ void k_eco_fill(int32_t *x, int n, int32_t v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

// Cell lists of the living animals (index_animals, main.cu:612-617), in
// ascending animal index: count, exclusive scan, place, sort each cell.
__global__ void k_eco_count(Eco e) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= e.P || e.alive[a] != 1) return;
    atomicAdd(&e.cell_count[cell_of(e, e.x[a], e.y[a])], 1);
}

constexpr int kScanThreads = 1024;
__global__ __launch_bounds__(kScanThreads) void k_eco_scan(Eco e) {
    __shared__ int part[kScanThreads];
    const int n = e.W * e.H, tid = threadIdx.x;
    const int per = (n + kScanThreads - 1) / kScanThreads;
    const int lo = tid * per, hi = min(n, lo + per);
    int sum = 0;
    for (int c = lo; c < hi; ++c) sum += e.cell_count[c];
    part[tid] = sum;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        const int v = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int run = part[tid] - sum;
    for (int c = lo; c < hi; ++c) {
        e.cell_start[c] = run;
        e.cell_fill[c] = 0;
        run += e.cell_count[c];
    }
}

__global__ void k_eco_place(Eco e) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= e.P || e.alive[a] != 1) return;
    const int c = cell_of(e, e.x[a], e.y[a]);
    e.cell_list[e.cell_start[c] + atomicAdd(&e.cell_fill[c], 1)] = a;
}

__global__ void k_eco_sort(Eco e) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= e.W * e.H) return;
    const int n = e.cell_count[c];
    if (n < 2) return;
    int32_t *l = e.cell_list + e.cell_start[c];
    for (int i = 1; i < n; ++i) {
        const int v = l[i];
        int j = i - 1;
        while (j >= 0 && l[j] > v) { l[j + 1] = l[j]; --j; }
        l[j + 1] = v;
    }
}

// bite_animals and finish_predation (main.cu:1251-1285), one cell at a time,
// attackers in index order, each biting the first living other animal of the
// cell at its turn; a bite that finds the prey's health at most the damage
// kills it and feeds the attacker.
__global__ void k_eco_bite(Eco e) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= e.W * e.H) return;
    const int n = e.cell_count[c];
    if (n < 2) return;
    const int32_t *l = e.cell_list + e.cell_start[c];
    for (int i = 0; i < n; ++i) {
        const int a = l[i];
        if (e.alive[a] != 1 || !e.bite[a]) continue;
        for (int j = 0; j < n; ++j) {
            const int prey = l[j];
            if (prey == a || e.alive[prey] != 1) continue;
            const float previous = e.health[prey];
            e.health[prey] = previous - kBiteDamage;
            e.pain[prey] += kBiteDamage;
            if (previous <= kBiteDamage) {
                e.alive[prey] = -2;
                e.energy[a] += fmaxf(0.0f, e.energy[prey]) * kMeatFraction;
                e.taste[4 * a + 3] = 1.0f;
                e.fullness[a] += 1.0f;
            }
            break;
        }
    }
}

__global__ void k_eco_finish(Eco e, uint8_t *run) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= e.P) return;
    if (e.alive[a] == -2) e.alive[a] = 0;
    run[a] = e.alive[a] == 1;
}

// regrow (main.cu:1138-1144): an empty cell gets food with the renewal
// probability, drawn from the cell's own stream for this step.
__global__ void k_eco_regrow(Eco e, int64_t step) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= e.W * e.H || e.food[c] != 0) return;
    unsigned long long rng = seed_for(e.seed + (unsigned long long)step, (unsigned long long)c);
    if (uniform01(rng) < e.regrowth) e.food[c] = 1;
}

// update_senses (main.cu:619-657) over the cell lists.
__global__ void k_eco_senses(Eco e) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= e.P || e.alive[a] != 1) return;
    const float x = e.x[a], y = e.y[a], heading = e.heading[a];
    float *out = e.senses + size_t(a) * kSenses;
    float rs[3], rc[3];
    for (int d = 0; d < 3; ++d) world_sincos(heading + float(d - 1) * 0.7f, &rs[d], &rc[d]);
    for (int s = 0; s < 6; ++s) {
        const int d = s % 3;
        float level = 0;
        for (int range = 1; range <= 4; ++range) {
            const int c = cell_of(e, x + rc[d] * float(range), y + rs[d] * float(range));
            if (e.kind[c] == s / 3 && e.food[c] > 0) level += 0.25f;
        }
        out[s] = clampf(level, 0.0f, 1.0f);
    }
    out[6] = clampf((16.0f - e.energy[a]) / 16.0f, 0.0f, 1.0f);
    out[7] = clampf(e.fullness[a] / 8.0f, 0.0f, 1.0f);
    out[8] = clampf(e.pain[a] / 3.0f, 0.0f, 1.0f);
    for (int k = 0; k < 3; ++k) out[9 + k] = clampf(e.taste[4 * a + k], 0.0f, 1.0f);
    for (int d = 0; d < 3; ++d) {
        float level = 0;
        for (int range = 1; range <= 4; ++range) {
            const int c = cell_of(e, x + rc[d] * float(range), y + rs[d] * float(range));
            const int32_t *l = e.cell_list + e.cell_start[c];
            for (int j = 0; j < e.cell_count[c]; ++j)
                if (l[j] != a && e.alive[l[j]] == 1) level += 0.25f;
        }
        out[12 + d] = clampf(level, 0.0f, 1.0f);
    }
    out[15] = clampf(e.taste[4 * a + 3], 0.0f, 1.0f);
    const int here = cell_of(e, x, y);
    for (int k = 0; k < 3; ++k)
        out[kContactFood + k] = e.food[here] > 0 && e.kind[here] == k ? 1.0f : 0.0f;
}

// World homeostasis input: each pool brain's animal age (-1 for a slot
// without an animal).
__global__ void k_eco_slot_age(const int32_t *age, const int32_t *batch_of, int nb, int32_t *slot_age) {
    const int pb = blockIdx.x * blockDim.x + threadIdx.x;
    if (pb >= nb) return;
    slot_age[pb] = batch_of[pb] >= 0 ? age[batch_of[pb]] : 0;
}

struct EcologyHost {
    Batch *batch = nullptr;
    Eco e{};
    DeviceArena arena;
    uint8_t *d_run = nullptr;
    int32_t *d_batch_of = nullptr, *d_slot_age = nullptr;
    uint64_t built_generation = 0;
    std::vector<uint8_t> h_run;
    int64_t steps = 0;
};

inline unsigned blocks(int64_t n, int t) { return unsigned((n + t - 1) / t); }

void rebuild_lists(const Eco &e, hipStream_t st) {
    const int cells = e.W * e.H;
    hipLaunchKernelGGL(k_eco_fill, dim3(blocks(cells, 256)), dim3(256), 0, st, e.cell_count, cells, 0);
    hipLaunchKernelGGL(k_eco_count, dim3(blocks(e.P, 256)), dim3(256), 0, st, e);
    hipLaunchKernelGGL(k_eco_scan, dim3(1), dim3(kScanThreads), 0, st, e);
    hipLaunchKernelGGL(k_eco_place, dim3(blocks(e.P, 256)), dim3(256), 0, st, e);
    hipLaunchKernelGGL(k_eco_sort, dim3(blocks(cells, 256)), dim3(256), 0, st, e);
}

}  // namespace
}  // namespace nsa

using namespace nsa;

extern "C" {

void *ns_ecology_create(void *batch, const ns_ecology_config *cfg) {
    auto *h = new EcologyHost;
    try {
        Batch &b = *static_cast<Batch *>(batch);
        if (!b.world) throw std::runtime_error("The ecology needs a world-rules batch");
        if (b.channels != kSenses || b.outputs != kMotors)
            throw std::runtime_error("The ecology needs 19 sensory channels and 5 motors");
        if (cfg->width < 1 || cfg->height < 1 || int64_t(cfg->width) * cfg->height > (1 << 26) ||
            !(cfg->food_density >= 0 && cfg->food_density <= 1.0f / 3) ||
            !(cfg->regrowth >= 0 && cfg->regrowth <= 1))
            throw std::runtime_error("Invalid ecology configuration");
        h->batch = &b;
        Eco &e = h->e;
        e.P = b.count; e.W = cfg->width; e.H = cfg->height; e.torus = cfg->torus != 0;
        e.instant = cfg->instant_motors != 0; e.regrowth = cfg->regrowth; e.seed = cfg->seed;
        const int P = e.P, cells = e.W * e.H;
        // founders (main.cu:530-545): position, heading, energy 14, health 10
        std::vector<float> x(P), y(P), hd(P), en(P, 14.0f), he(P, 10.0f);
        std::vector<int32_t> active(P), alive(P, 1), zero(P, 0), food(cells);
        std::vector<uint8_t> kind(cells);
        for (int a = 0; a < P; ++a) {
            unsigned long long rng = seed_for(e.seed ^ 0x5eedf00du, (unsigned long long)a);
            const float span_x = e.torus ? float(e.W) : float(e.W - 1);
            const float span_y = e.torus ? float(e.H) : float(e.H - 1);
            x[a] = uniform01(rng) * span_x;
            y[a] = uniform01(rng) * span_y;
            hd[a] = uniform01(rng) * 2.0f * kPi;
            int n = 0;
            for (const auto &g : b.brains[a]->genes) n += g.active != 0;
            active[a] = n;
        }
        // main.cu:458-462 with the configured density per kind
        for (int c = 0; c < cells; ++c) {
            unsigned long long rng = seed_for(e.seed, (unsigned long long)c);
            kind[c] = uint8_t(int(uniform01(rng) * 3));
            food[c] = uniform01(rng) < 3.0f * cfg->food_density ? 1 : 0;
        }
        auto up = [&](auto *dst, const auto &src) {
            NS_HIP(hipMemcpy(dst, src.data(), sizeof(src[0]) * src.size(), hipMemcpyHostToDevice));
        };
        DeviceArena &A = h->arena;
        e.x = A.alloc<float>(P); up(e.x, x);
        e.y = A.alloc<float>(P); up(e.y, y);
        e.heading = A.alloc<float>(P); up(e.heading, hd);
        e.energy = A.alloc<float>(P); up(e.energy, en);
        e.health = A.alloc<float>(P); up(e.health, he);
        std::vector<float> zf(size_t(P) * 4, 0.0f);
        e.fullness = A.alloc<float>(P); up(e.fullness, std::vector<float>(P, 0.0f));
        e.pain = A.alloc<float>(P); up(e.pain, std::vector<float>(P, 0.0f));
        e.taste = A.alloc<float>(4 * int64_t(P)); up(e.taste, zf);
        e.muscle = A.alloc<float>(3 * int64_t(P)); up(e.muscle, std::vector<float>(size_t(P) * 3, 0.0f));
        e.age = A.alloc<int32_t>(P); up(e.age, zero);
        e.alive = A.alloc<int32_t>(P); up(e.alive, alive);
        e.active_neurons = A.alloc<int32_t>(P); up(e.active_neurons, active);
        e.food_cell = A.alloc<int32_t>(P); up(e.food_cell, zero);
        e.eat = A.alloc<uint8_t>(P);
        e.bite = A.alloc<uint8_t>(P);
        e.food = A.alloc<int32_t>(cells); up(e.food, food);
        e.kind = A.alloc<uint8_t>(cells); up(e.kind, kind);
        e.eat_claim = A.alloc<int32_t>(cells);
        e.cell_count = A.alloc<int32_t>(cells);
        e.cell_start = A.alloc<int32_t>(cells);
        e.cell_fill = A.alloc<int32_t>(cells);
        e.cell_list = A.alloc<int32_t>(P);
        e.senses = A.alloc<float>(int64_t(P) * kSenses);
        e.motors = A.alloc<float>(int64_t(P) * kMotors);
        h->d_run = A.alloc<uint8_t>(P);
        h->d_slot_age = A.alloc<int32_t>(std::max(b.pool.nb, 1));
        h->d_batch_of = A.alloc<int32_t>(std::max(b.pool.nb, 1));
        NS_HIP(hipMemset(e.eat, 0, P));
        NS_HIP(hipMemset(e.bite, 0, P));
        NS_HIP(hipMemset(e.senses, 0, sizeof(float) * P * kSenses));
        NS_HIP(hipMemset(e.motors, 0, sizeof(float) * P * kMotors));
        h->h_run.assign(P, 1);
        // the founders' first senses (main.cu: update_senses ends every step)
        hipStream_t st = b.stream;
        rebuild_lists(e, st);
        hipLaunchKernelGGL(k_eco_senses, dim3(blocks(P, 128)), dim3(128), 0, st, e);
        NS_HIP(hipGetLastError());
        NS_HIP(hipStreamSynchronize(st));
        return h;
    } catch (const std::exception &ex) {
        delete h;
        set_error(ex.what());
        return nullptr;
    }
}

void ns_ecology_destroy(void *p) {
    auto *h = static_cast<EcologyHost *>(p);
    if (h && h->batch && h->batch->stream) (void)hipStreamSynchronize(h->batch->stream);
    delete h;
}

// One world step (main.cu:2634-2768 without reproduction): the 16 neural ticks
// of every living animal on its current senses, then body, bites, world
// homeostasis, regrowth and the next senses.  Synchronous.
int ns_ecology_step(void *p, int64_t step) {
    try {
        auto &h = *static_cast<EcologyHost *>(p);
        Batch &b = *h.batch;
        Eco &e = h.e;
        hipStream_t st = b.stream;
        const int P = e.P, cells = e.W * e.H;
        b.set_run(h.h_run.data());
        b.run(kTicks, e.senses, 230.0f, e.motors, 0, st);
        hipLaunchKernelGGL(k_eco_fill, dim3(blocks(cells, 256)), dim3(256), 0, st, e.eat_claim, cells, INT_MAX);
        hipLaunchKernelGGL(k_eco_move, dim3(blocks(P, 128)), dim3(128), 0, st, e);
        hipLaunchKernelGGL(k_eco_feed, dim3(blocks(P, 128)), dim3(128), 0, st, e);
        rebuild_lists(e, st);
        hipLaunchKernelGGL(k_eco_bite, dim3(blocks(cells, 256)), dim3(256), 0, st, e);
        hipLaunchKernelGGL(k_eco_finish, dim3(blocks(P, 256)), dim3(256), 0, st, e, h.d_run);
        NS_HIP(hipMemcpyAsync(h.h_run.data(), h.d_run, P, hipMemcpyDeviceToHost, st));
        NS_HIP(hipStreamSynchronize(st));
        // homeostasis_step (main.cu:1287-1300) of the animals still alive
        b.set_run(h.h_run.data());
        if (h.built_generation != b.pool.generation) {
            NS_HIP(hipMemcpy(h.d_batch_of, b.pool.batch_of.data(), sizeof(int32_t) * b.pool.nb,
                             hipMemcpyHostToDevice));
            h.built_generation = b.pool.generation;
        }
        hipLaunchKernelGGL(k_eco_slot_age, dim3(blocks(b.pool.nb, 256)), dim3(256), 0, st, e.age,
                           h.d_batch_of, b.pool.nb, h.d_slot_age);
        launch_world_homeostasis(b.pool.view, b.pool.nn, h.d_slot_age, st);
        hipLaunchKernelGGL(k_eco_regrow, dim3(blocks(cells, 256)), dim3(256), 0, st, e, step);
        rebuild_lists(e, st);
        hipLaunchKernelGGL(k_eco_senses, dim3(blocks(P, 128)), dim3(128), 0, st, e);
        NS_HIP(hipGetLastError());
        NS_HIP(hipStreamSynchronize(st));
        ++h.steps;
        return 0;
    } catch (const std::exception &ex) {
        set_error(ex.what());
        return -1;
    }
}

int ns_ecology_export(void *p, ns_world_animal *animals, int32_t *food, uint8_t *kind, float *senses,
                      float *motors) {
    try {
        auto &h = *static_cast<EcologyHost *>(p);
        const Eco &e = h.e;
        const int P = e.P, cells = e.W * e.H;
        NS_HIP(hipStreamSynchronize(h.batch->stream));
        auto down = [](void *dst, const void *src, size_t n) {
            NS_HIP(hipMemcpy(dst, src, n, hipMemcpyDeviceToHost));
        };
        if (animals) {
            std::vector<float> x(P), y(P), hd(P), en(P), he(P), fu(P), pa(P), ta(4 * size_t(P)),
                mu(3 * size_t(P));
            std::vector<int32_t> age(P), alive(P);
            down(x.data(), e.x, 4 * P); down(y.data(), e.y, 4 * P); down(hd.data(), e.heading, 4 * P);
            down(en.data(), e.energy, 4 * P); down(he.data(), e.health, 4 * P);
            down(fu.data(), e.fullness, 4 * P); down(pa.data(), e.pain, 4 * P);
            down(ta.data(), e.taste, 16 * size_t(P)); down(mu.data(), e.muscle, 12 * size_t(P));
            down(age.data(), e.age, 4 * P); down(alive.data(), e.alive, 4 * P);
            for (int a = 0; a < P; ++a) {
                ns_world_animal &r = animals[a];
                r.x = x[a]; r.y = y[a]; r.heading = hd[a]; r.energy = en[a]; r.health = he[a];
                r.fullness = fu[a]; r.pain = pa[a];
                for (int k = 0; k < 4; ++k) r.taste[k] = ta[4 * a + k];
                for (int k = 0; k < 3; ++k) r.muscle[k] = mu[3 * a + k];
                r.age = age[a]; r.alive = alive[a];
            }
        }
        if (food) down(food, e.food, sizeof(int32_t) * cells);
        if (kind) down(kind, e.kind, cells);
        if (senses) down(senses, e.senses, sizeof(float) * P * kSenses);
        if (motors) down(motors, e.motors, sizeof(float) * P * kMotors);
        return 0;
    } catch (const std::exception &ex) {
        set_error(ex.what());
        return -1;
    }
}

}  // extern "C"
