// Host-side structures of libns_amd.so: genome handles, the device pool and
// its builder.  Shared by the batch, world and serving surfaces.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <memory>
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <exception>
#include <functional>
#include <mutex>
#include <set>
#include <stdexcept>
#include <thread>
#include <tuple>
#include <string>
#include <vector>

#include "ns_amd.h"
#include "ns_kernels.h"

namespace nsa {

void set_error(const std::string &msg);

#define NS_HIP(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_));    \
    } while (0)

// Genome handle made by ns_create (gym_adapter/native.cpp:23-37, minus state).
struct HostBrain {
    std::vector<ns_neuron_gene> genes;
    std::vector<ns_synapse_gene> edges;   // pre/post brain-local
    std::vector<int32_t> sensory;         // input channel or -1
    std::vector<uint8_t> receptor;        // receptor-event sensory path
    std::vector<int32_t> motors;
    std::vector<uint32_t> orig_n, orig_s;
    int channels = 1;
    int max_delay = 8;
    int8_t slow_cell = -1;   // has_slow_cell, once known (-1: not yet)
};

ns_neuron_gene profile_gene(int type);
// A valid excitatory-to-excitatory synapse with a plasticity profile.
bool has_plastic_pair(const HostBrain &b);
// LP or Prinz cells (type 2 or 3), which run on the per-tick kernels.
bool has_slow_cell(const HostBrain &b);
HostBrain *parse_gym_brain(int count, int edges, int channels, int outputs, const float *ng,
                           const float *sg, const int *motors, int max_delay);
inline int site_count(const ns_synapse_gene &g) {   // physiology.cuh:376-378
    return g.release_sites ? (g.release_sites > 64 ? 64 : g.release_sites) : 14;
}

// A carved device range ran out (see DeviceArena::carve).
struct CarveFull : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// Owning list of device allocations, or, after carve(), consecutive 256-byte
// aligned ranges of one caller-owned device block (no allocation; CarveFull
// when the block is exhausted).
struct DeviceArena {
    std::vector<void *> blocks;
    int64_t bytes = 0;
    unsigned char *carve_base = nullptr;
    size_t carve_cap = 0, carve_used = 0;
    template <class T> T *alloc(int64_t count) {
        if (count <= 0) count = 1;
        const size_t n = sizeof(T) * size_t(count);
        if (carve_base) {
            const size_t off = (carve_used + 255) & ~size_t(255);
            if (off > carve_cap || n > carve_cap - off)
                throw CarveFull("static device block of " + std::to_string(carve_cap) +
                                " bytes exhausted");
            carve_used = off + n;
            bytes += int64_t(n);
            return reinterpret_cast<T *>(carve_base + off);
        }
        void *p = nullptr;
        NS_HIP(hipMalloc(&p, n));
        blocks.push_back(p);
        bytes += int64_t(n);
        return static_cast<T *>(p);
    }
    void carve(void *base, size_t cap) {
        release();
        carve_base = static_cast<unsigned char *>(base);
        carve_cap = cap;
    }
    void release() {
        for (void *p : blocks) (void)hipFree(p);
        blocks.clear();
        bytes = 0;
        carve_used = 0;
    }
    ~DeviceArena() { release(); }
};

// One pool slot as the builder sees it: a brain, or no brain (a spare slot),
// with an optional capacity larger than the brain.
struct BrainSpec {
    const HostBrain *brain;   // nullptr: an empty slot
    int batch_index;          // input / output row (-1 for an empty slot)
    int input_base;           // absolute input index of channel 0
    int output_base;          // absolute output index of motor 0
    bool world;               // src/main.cu neuron_step rules
    int cap_n = 0, cap_s = 0; // slot capacity; at least the brain's size
};

// Plans one staged upload: each entry is a device range and an offset into
// the staged data.  Data is written when the buffer exists: `add` records a
// copy from host memory that must stay valid until write(); `alloc` only
// reserves room that a filler writes later.  Chunks are 16-byte aligned.
struct Stager {
    struct Copy { size_t off; const void *src; size_t bytes; };
    std::vector<ScatterEntry> entries;
    std::vector<Copy> copies;
    size_t size = 0;
    void clear() { entries.clear(); copies.clear(); size = 0; }
    size_t alloc(const void *dst, size_t bytes) {
        const size_t off = (size + 15) & ~size_t(15);
        size = off + bytes;
        entries.push_back(ScatterEntry{uint64_t(reinterpret_cast<uintptr_t>(dst)), uint64_t(off),
                                       uint64_t(bytes)});
        return off;
    }
    void add(const void *dst, const void *src, size_t bytes) {
        if (!bytes) return;
        copies.push_back(Copy{alloc(dst, bytes), src, bytes});
    }
    size_t head() const { return (entries.size() * sizeof(ScatterEntry) + 15) & ~size_t(15); }
    size_t bytes() const { return head() + size; }
    // entries (offsets rebased past them) and the recorded copies; returns
    // the start of the data region for fillers
    unsigned char *write(unsigned char *out) const {
        const size_t h = head();
        ScatterEntry *e = reinterpret_cast<ScatterEntry *>(out);
        for (size_t i = 0; i < entries.size(); ++i) {
            e[i] = entries[i];
            e[i].src += h;
        }
        for (const Copy &c : copies) std::memcpy(out + h + c.off, c.src, c.bytes);
        return out + h;
    }
};

// Offsets (into the staged data) of one newborn's slot-image chunks.
struct PlaceChunks {
    int pb = 0;
    const BrainSpec *spec = nullptr;
    size_t ng[NG_FIELDS], fl, inp, no, ni, sg[SG_FIELDS], sf, pre, post, so, mo;
    int ninx = 0;
};

// Staging buffers start at this size and double when outgrown (a regrowth
// waits for the buffer's last use).  A growth maps new memory while the
// previous launch still runs, which has stretched that launch by several ms,
// so the floor covers a 1% turnover of 32,768 brains (about 7-14 MB of
// images per step) twice over.
constexpr size_t kStageMinBytes = size_t(64) << 20;

// Pinned host staging for slot placements, double buffered so one buffer can
// be filled while the other's copy is still in flight.
// Host worker threads for the placement work between launches (genome
// copies, slot images), created once per process and kept waiting, so a
// placement of a few hundred newborns does not pay for thread creation.
// run(n, fn) calls fn(i) for i in [0, n) on the workers and the caller and
// returns when every call has returned; calls from several host threads
// take turns.
class HostWorkers {
public:
    static HostWorkers &get();
    void run(int n, const std::function<void(int)> &fn);
    ~HostWorkers();

private:
    HostWorkers();
    void loop();
    std::vector<std::thread> threads_;
    std::mutex run_lock_, m_;
    std::condition_variable wake_, done_;
    const std::function<void(int)> *fn_ = nullptr;
    std::atomic<int> next_{0};
    int n_ = 0, busy_ = 0;
    uint64_t round_ = 0;
    bool stop_ = false;
    std::exception_ptr error_;   // the first exception of the round, rethrown by run
};

// Releases retired genomes on a host thread of its own: freeing a replaced
// brain's genome (a few heap blocks per brain, cold in cache) stays off the
// thread that places the newborns.
class GenomeReaper {
public:
    static GenomeReaper &get();
    void retire(std::vector<std::shared_ptr<const HostBrain>> &&genomes);
    ~GenomeReaper();

private:
    GenomeReaper();
    void loop();
    std::thread thread_;
    std::mutex m_;
    std::condition_variable wake_;
    std::vector<std::shared_ptr<const HostBrain>> pending_;
    bool stop_ = false;
};

struct PinnedStaging {
    struct Buf { unsigned char *p = nullptr; size_t cap = 0; hipEvent_t done = nullptr; };
    Buf bufs[2];
    int next = 0;
    int64_t grows = 0;
    Buf &take(size_t bytes) {
        Buf &b = bufs[next];
        next ^= 1;
        if (b.done) NS_HIP(hipEventSynchronize(b.done));
        if (b.cap < bytes) {
            ++grows;
            if (b.p) (void)hipHostFree(b.p);
            b.p = nullptr;
            b.cap = 0;
            const size_t cap = std::max(2 * bytes, kStageMinBytes);
            NS_HIP(hipHostMalloc(&b.p, cap));
            b.cap = cap;
        }
        if (!b.done) NS_HIP(hipEventCreateWithFlags(&b.done, hipEventDisableTiming));
        return b;
    }
    ~PinnedStaging() {
        for (auto &b : bufs) {
            if (b.done) { (void)hipEventSynchronize(b.done); (void)hipEventDestroy(b.done); }
            if (b.p) (void)hipHostFree(b.p);
        }
    }
};

// Tile of the resident launch: ks > 0 selects the register-slot kernel, ks == 0
// the LDS kernel with lds synapses per pack.
struct Tile {
    int threads = 256, kn = 2, ks = 4, minw = 1, lds = 0;
    int nmax() const { return threads * kn; }
    int smax() const { return ks > 0 ? threads * ks : lds; }
};
const TileShape *find_tile(const Tile &t);

// Pool slots for `specs` under a slot reserve: capacities rounded up a
// geometric ladder of ratio 1 + slack (cut at the tile's pack limits for
// brains that fit a pack), plus empty slots for a fraction `spare` of the
// brains at quantiles of the slot sizes and a few "jumbo" slots as large as
// the largest.
std::vector<BrainSpec> reserve_slots(const std::vector<BrainSpec> &specs, float slack, float spare,
                                     const Tile &tile, bool world);

// Structure-of-arrays device pool over a set of brains, laid out in packs.
class DevicePool {
public:
    ~DevicePool();
    void build(const std::vector<BrainSpec> &specs, int outputs, const Tile &tile,
               hipStream_t stream);

    // state transfer in genome order for pool brain pb
    void download_state(std::vector<ns_neuron_state> &ns, std::vector<ns_synapse_state> &ss,
                        std::vector<uint32_t> &hist, std::vector<int64_t> &ticks,
                        std::vector<uint64_t> *stamps = nullptr);
    void upload_state(const std::vector<ns_neuron_state> &ns,
                      const std::vector<ns_synapse_state> &ss, const std::vector<uint32_t> &hist,
                      const std::vector<int64_t> &ticks, const std::vector<uint64_t> *stamps = nullptr);

    Pool view{};
    DeviceArena arena;
    int64_t nn = 0, ns = 0;          // totals
    int nb = 0, outputs = 0;
    Tile tile;
    // per pool brain
    std::vector<int32_t> bn0, bnc, bs0, bsc, batch_of;   // bnc/bsc: slot capacities
    std::vector<uint8_t> fallback;   // brun bit 1
    std::vector<int32_t> pool_of;    // batch index -> pool brain
    // genome order -> pool position, per pool brain (offsets bs0)
    std::vector<int32_t> syn_perm;   // for pool synapse position q: genome edge index
    std::vector<Pack> packs;
    Pack *d_packs = nullptr;
    PlainRange fallback_range{}, all_range{};
    int32_t *d_slow = nullptr, *d_big = nullptr;
    uint16_t *d_perm = nullptr;
    bool plastic = false;            // any valid plastic excitatory pair
    bool force_plastic = false;      // build the plastic layout even without one (serving)
    int smax_lds = 0;                // LDS synapse capacity per pack (LDS tiles)
    int slow_count = 0, fallback_brains = 0, nbig = 0;
    // Slots.  Every pool brain index is a slot; its capacity (bnc, bsc) may
    // exceed its occupant (occ_n, occ_s; 0 = free, batch_of = -1).  Neurons
    // past the occupant are inactive and synapses past it invalid.
    int nbatch = 0;                  // batch brains (pool_of's size)
    std::vector<int32_t> occ_n, occ_s, pack_of, h_bout0;
    std::vector<uint8_t> slot_kind;  // 0 per-tick class, 1 own workgroup, 2 pack
    std::vector<uint32_t> h_nflags;  // host copies behind pack slot orders
    std::vector<int32_t> touched;   // packs place() changed, awaiting order_touched
    int32_t *d_touched = nullptr;
    PinnedStaging staging;
    // Device copies of staged placements, double buffered: one is uploaded
    // on copy_stream (overlapping work already queued on the caller's
    // stream) while scatters of the other may still be queued.
    struct DeviceStage {
        DeviceArena arena;
        unsigned char *p = nullptr;
        size_t cap = 0;
        hipEvent_t uploaded = nullptr;   // on copy_stream, after its upload
        hipEvent_t read = nullptr;       // after the last scatter that read it
    };
    DeviceStage dstage[2];
    int stage_cur = 0, stage_next = 0;
    int64_t dstage_grows = 0;
    hipStream_t copy_stream = nullptr;   // default priority
    // Copy the staged buffer to the device (one upload on copy_stream, which
    // `st` then waits for); entries [lo, hi) are then scattered on `st` by
    // scatter_range.
    void upload_stage(const struct Stager &sg, hipStream_t st);
    std::vector<PlaceChunks> fill_list;   // images place() reserved, filled by upload_stage
    void scatter_range(int lo, int hi, hipStream_t st);
    // Recompute on the device the slot orders of the packs the last place()
    // filled; call after the staged images have been scattered.
    void order_touched(hipStream_t st);
    // Stage the state of the occupant of slot pb (placed by place() in the
    // same staging): neurons in genome order with their shift registers and
    // modulator stamps (6 per neuron), synapses in genome edge order, and the
    // slot's tick.  The staged field buffers are kept in `keep` until the
    // upload.
    void stage_state(int pb, const HostBrain &brain, const ns_neuron_state *neurons,
                     const uint32_t *hist, const uint64_t *stamps,
                     const ns_synapse_state *synapses, int64_t tick, struct Stager &sg,
                     std::vector<std::vector<unsigned char>> &keep);
    // Stage the images of `moves` (slot, occupant or nullptr = vacate) into
    // `sg`: genes, flags, routes, endpoints, motors and the affected packs'
    // slot orders, and update the host-side slot tables.  State is left
    // alone (the caller resets newborns).
    void place(const std::vector<std::pair<int, const BrainSpec *>> &moves, struct Stager &sg);
    uint32_t *d_norig = nullptr, *d_sorig = nullptr, *d_bseed = nullptr;
    int32_t *d_reset = nullptr;      // 2 x nb: slots to reset, then their seeds
    hipStream_t stream = nullptr;
    uint64_t generation = 0;         // counts layouts (builds, appended slots)
    // Headroom for slots appended after a build: a growable pool reserves
    // room past its last slot for neurons, synapses, slots, packs and
    // own-workgroup entries; append_slots fills it and a build starts over.
    bool growable = false;
    int nb_cap = 0, packs_cap = 0, big_cap = 0;
    int64_t appended = 0;            // slots appended since the pool was created
    int open_pack = -1;              // the pack appended slots extend while it has room
    // Append empty slots of capacities caps[j] = (neurons, synapses), each
    // sized for the newborn of occ[j] = (neurons, synapses) that place() puts
    // there next: a place in the last appended pack while it has room (its
    // slots and neurons end the pool), else a new pack, or an own-workgroup
    // entry when larger than a pack, and the parts of the slot image place()
    // leaves alone (brain indexes, the tails of the runs and endpoints)
    // staged into `sg` (buffers kept in `keep`).  The new slot indices go to
    // `slots`.  False, with nothing changed, when the headroom is short.
    bool append_slots(const std::vector<std::pair<int, int>> &caps,
                      const std::vector<std::pair<int, int>> &occ, struct Stager &sg,
                      std::vector<std::vector<unsigned char>> &keep, std::vector<int> &slots);
};

// One call of `ticks` ticks over every brain whose run flag is set: the
// resident kernel over packs plus the per-tick kernels over the rest (or the
// per-tick kernels over everything when form == NS_FORM_PER_TICK).
struct KernelTimer;
void run_pool(DevicePool &pool, int form, int ticks, const float *d_in, float gain, float *d_out,
              int64_t out_count, int homeostasis, hipStream_t st, KernelTimer *timer = nullptr);

// HIP-event timing of tick-kernel launches (pairs resolved lazily).  Events
// come from a pool created when timing is switched on and are reused once
// resolved, so timed launches create no events while kernels run.
struct KernelTimer {
    struct Pair { hipEvent_t a, b; int kind; };
    std::vector<Pair> pending;
    std::vector<hipEvent_t> spare;
    double ms[2] = {0, 0};
    int64_t count[2] = {0, 0};
    hipEvent_t take() {
        hipEvent_t e = nullptr;
        if (spare.empty()) {
            NS_HIP(hipEventCreate(&e));
        } else {
            e = spare.back();
            spare.pop_back();
        }
        return e;
    }
    void reserve(size_t n) {
        while (spare.size() < n) {
            hipEvent_t e = nullptr;
            NS_HIP(hipEventCreate(&e));
            spare.push_back(e);
        }
    }
    void begin(hipStream_t st, int kind, hipEvent_t &a) {
        a = take();
        NS_HIP(hipEventRecord(a, st));
        pending.push_back({a, nullptr, kind});
    }
    void end(hipStream_t st) {
        pending.back().b = take();
        NS_HIP(hipEventRecord(pending.back().b, st));
    }
    void resolve() {
        for (auto &p : pending) {
            NS_HIP(hipEventSynchronize(p.b));
            float t = 0;
            NS_HIP(hipEventElapsedTime(&t, p.a, p.b));
            ms[p.kind] += t;
            count[p.kind] += 1;
            spare.push_back(p.a);
            spare.push_back(p.b);
        }
        pending.clear();
    }
    ~KernelTimer() {
        for (auto &p : pending) {
            (void)hipEventDestroy(p.a);
            if (p.b) (void)hipEventDestroy(p.b);
        }
        for (hipEvent_t e : spare) (void)hipEventDestroy(e);
    }
};

// Batched-episode surface (gym_adapter/gpu.cu GpuBatch).
struct Batch {
    DevicePool pool;
    std::vector<BrainSpec> specs;
    // Genomes of the brains in batch order, shared with the genome handles they
    // came from: a genome is immutable once a batch holds it (a handle that
    // changes its genome while shared takes a copy first).
    std::vector<std::shared_ptr<const HostBrain>> brains;
    int count = 0, channels = 0, outputs = 0;
    int form = NS_FORM_RESIDENT;
    bool world = false;
    float *d_input = nullptr, *d_output = nullptr;
    uint8_t *d_run_host_order = nullptr;
    DeviceArena io;
    hipStream_t stream = nullptr;
    std::vector<uint8_t> h_brun;                       // per slot
    bool brun_valid = false;
    std::vector<int32_t> h_status;
    bool timing = false;
    KernelTimer timer;
    // Graph mode: the captured step and the arguments it was captured for.
    struct StepGraph {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        uint64_t generation = 0;
        int form = 0, ticks = 0, homeostasis = 0;
        const float *in = nullptr;
        float *out = nullptr;
        float gain = 0.0f;
        hipStream_t stream = nullptr;
        void release();
    };
    bool graph_mode = false;
    StepGraph step_graph;
    int64_t graph_captures = 0, graph_launches = 0;

    // Slot reserve for brain replacement: capacities rounded up a geometric
    // ladder of ratio 1 + slack, plus spare empty slots (a fraction `spare`
    // of the batch) spread over the brain-size distribution.
    float slack = 0.0f, spare = 0.0f;
    int64_t relayouts = 0, placements = 0;
    std::vector<uint8_t> last_mask;                    // batch order, last run mask
    Stager stage;                                      // reused across replacements
    // Free slots by (neuron capacity, synapse capacity, slot), kept across
    // replacements while the layout generation is free_gen.
    std::set<std::tuple<int, int, int>> free_fit;
    uint64_t free_gen = ~uint64_t(0);
    std::vector<std::vector<unsigned char>> stage_keep;   // buffers `stage` copies from

    void create(std::vector<std::shared_ptr<const HostBrain>> &&genomes, int max_delay_unused);
    void relayout(const Tile &tile);
    void set_run(const uint8_t *mask);                 // host mask, batch order
    // Replace brains idx[0..k) by new genomes with fresh state (seeds as
    // ns_gpu_reset): each goes to the best-fitting free slot and the pool is
    // rebuilt only when none fits.  Stream-ordered on `st`.
    void replace(int k, const int *idx, std::vector<std::shared_ptr<const HostBrain>> &&fresh,
                 const uint32_t *seeds, hipStream_t st);
    // Pool slots for the current brains under the reserve policy.
    std::vector<BrainSpec> slot_specs(const Tile &tile) const;
    void run(int ticks, const float *d_in, float gain, float *d_out, int homeostasis,
             hipStream_t st);
    ~Batch();
};

// Default tile for a set of brains (the pool builder's first layout).
Tile default_tile(const std::vector<const HostBrain *> &brains);

// One-brain device pool a genome handle steps on (ns_step): pool arrays plus
// the call's input and output rows.
struct HandleDevice {
    DevicePool pool;
    float *d_in = nullptr, *d_out = nullptr;
};

// Genome handle of ns_create / ns_create_ex: the genome and the state the
// reference's CPU handle keeps with it (gym_adapter/native.cpp:23-34):
// neuron and synapse states, the delay ring as spike shift registers, and the
// tick.  ns_step's one-brain device pool is made on its first call.
struct GenomeHandle {
    std::shared_ptr<HostBrain> genome;   // shared with the batches made from the handle
    const HostBrain &brain() const { return *genome; }
    std::vector<ns_neuron_state> neurons;
    std::vector<ns_synapse_state> synapses;
    std::vector<uint32_t> hist;
    int64_t tick = 0;
    std::unique_ptr<HandleDevice> device;
};
inline const HostBrain &handle_brain(void *h) { return static_cast<GenomeHandle *>(h)->brain(); }
inline std::shared_ptr<const HostBrain> handle_genome(void *h) {
    return static_cast<GenomeHandle *>(h)->genome;
}
GenomeHandle *make_handle(HostBrain *parsed);
// ns_step's argument checks (native.cpp:97-98), the one-brain pool layout and
// one step on it, shared by the allocating and the static-pool entry points.
void check_step(const HostBrain &b, const float *input, int ticks, float gain);
void build_one_brain(DevicePool &pool, const HostBrain &b, float *&d_in, float *&d_out);
void step_one_brain(DevicePool &pool, GenomeHandle &h, float *d_in, float *d_out,
                    const float *input, int ticks, float gain, float *output);

}  // namespace nsa
