// NSCHKP01 checkpoint files of the world simulator (src/checkpoint_impl.cuh,
// src/main.cu:1986-2000, 2575-2612) for the world-pool surface.  Host code
// only: gzip container, header and layout checks, record checks, atomic
// save.  The world sections (observer histories, animals, food grid, senses,
// motors, counters, traces, the plasticity-cohort extension) are carried as
// one opaque block in file order; the neural records go to and come from a
// world pool.
#include <fcntl.h>
#include <unistd.h>
#include <zlib.h>

#include <cmath>
#include <cstddef>
#include <cstring>
#include <filesystem>
#include <memory>
#include <string>
#include <vector>

#include "ns_host.h"

namespace nsa {
namespace {

// The world build's constants (src/main.cu:33-47, physiology.cuh:17) and
// struct sizes (main.cu:66-104, 164-177, 1845-1849).
constexpr uint64_t kP = 32768, kNeuronCap = 4194304, kSynapseCap = 8388608;
constexpr uint64_t kS = 19, kM = 5, kW = 128, kH = 128, kDendrites = 4, kTicks = 16;
constexpr uint64_t kAnimalBytes = 472;          // Animal, Behavior 276 B at offset 196
constexpr uint64_t kTotalsBytes = 528;          // BehaviorTotals
constexpr uint64_t kLifeBytes = 504;            // LifeRecord
constexpr uint64_t kBiteBytes = 24;             // BiteEvent
constexpr uint64_t kCohortBytes = 80;           // PlasticityCohort
constexpr size_t kAliveOffset = 80, kNeuronCountOffset = 92, kSynapseCountOffset = 96;
constexpr size_t kChunk = size_t(1) << 20;      // checkpoint_impl.cuh:3
constexpr uint64_t kCohortTag = 0x4e53504c41535431ull;   // "NSPLAST1"

static_assert(sizeof(ns_checkpoint_header) == 608, "CheckpointHeader layout");
static_assert(sizeof(ns_checkpoint_neuron) == 344 && sizeof(ns_checkpoint_synapse) == 136,
              "checkpoint record layout");
static_assert(sizeof(ns_checkpoint_neuron17) == 352 &&
                  offsetof(ns_checkpoint_neuron17, spike_history) ==
                      offsetof(ns_checkpoint_neuron, spike_history),
              "17-slot checkpoint record layout");

// The ring length of a neuron record type (main.cu:46 HISTORY).
template <class R> constexpr uint64_t history_of();
template <> constexpr uint64_t history_of<ns_checkpoint_neuron>() { return 9; }
template <> constexpr uint64_t history_of<ns_checkpoint_neuron17>() { return 17; }

// Bytes of the fixed world arrays after the observer section, in file order
// (checkpoint_impl.cuh:173-190): animals, food, kind, senses, motors, five
// counters, founder motor steps, founder observed steps, six trace arrays.
constexpr uint64_t kArrayBytes = kP * kAnimalBytes + 2 * kW * kH * 4 + kP * kS * 4 + kP * kM * 4 +
                                 5 * 4 + kM * 4 + 4 + 6 * kM * 4;
constexpr uint64_t kExtensionBytes = 4 * 8 + kP * kCohortBytes;

// checkpoint_impl.cuh:21-32 over a build with `history` ring slots (9 or 17)
uint64_t layout_hash(uint64_t history) {
    const uint64_t record = history == 17 ? sizeof(ns_checkpoint_neuron17) : sizeof(ns_checkpoint_neuron);
    const uint64_t values[] = {kP, kNeuronCap, kSynapseCap, kS, kM, kW, kH, kDendrites, history,
                               kTicks, kAnimalBytes, sizeof(ns_neuron_gene), sizeof(ns_neuron_state),
                               sizeof(ns_synapse_gene), sizeof(ns_synapse_state),
                               record, sizeof(ns_checkpoint_synapse),
                               kTotalsBytes, kLifeBytes, kBiteBytes};
    uint64_t h = 1469598103934665603ull;            // FNV-1a over whole values
    for (uint64_t v : values) {
        h ^= v;
        h *= 1099511628211ull;
    }
    return h;
}

// The layout version a header names: 9 or 17 ring slots, 0 for neither.
uint64_t header_history(const ns_checkpoint_header &h) {
    return h.layout == layout_hash(9) ? 9 : h.layout == layout_hash(17) ? 17 : 0;
}

// checkpoint_impl.cuh:368-383, the layout being that of `history` slots
void check_header(const ns_checkpoint_header &h, uint64_t history) {
    if (std::memcmp(h.magic, "NSCHKP01", 8) != 0 || h.version != 1 || h.endian != 0x01020304u ||
        h.layout != layout_hash(history) || h.step < 0 || h.step > INT64_MAX / int64_t(kTicks) ||
        h.checkpoint_every <= 0 || h.initial < 1 || uint64_t(h.initial) > kP ||
        h.trace_animal < -1 || h.trace_animal >= h.initial || (h.flags & ~127u) != 0 ||
        !std::isfinite(h.food_renewal) || h.food_renewal < 0 || h.food_renewal > 1 ||
        h.neuron_count > kNeuronCap || h.synapse_count > kSynapseCap ||
        !std::memchr(h.output, 0, sizeof(h.output)))
        throw std::runtime_error("Incompatible or invalid checkpoint header");
}

struct GzFile {
    gzFile f = nullptr;
    ~GzFile() { if (f) gzclose(f); }
    void read(void *data, size_t bytes) {   // checkpoint_impl.cuh:48-59
        auto *p = static_cast<unsigned char *>(data);
        while (bytes) {
            const unsigned n = unsigned(std::min(bytes, kChunk));
            if (gzread(f, p, n) != int(n))
                throw std::runtime_error(std::string("Checkpoint is incomplete or corrupt: ") +
                                         gzerror(f, nullptr));
            p += n;
            bytes -= n;
        }
    }
    void write(const void *data, size_t bytes) {   // checkpoint_impl.cuh:35-46
        auto *p = static_cast<const unsigned char *>(data);
        while (bytes) {
            const unsigned n = unsigned(std::min(bytes, kChunk));
            if (gzwrite(f, p, n) != int(n))
                throw std::runtime_error(std::string("Checkpoint write: ") + gzerror(f, nullptr));
            p += n;
            bytes -= n;
        }
    }
};

template <class T> T load_at(const unsigned char *p) {
    T v;
    std::memcpy(&v, p, sizeof(T));
    return v;
}

// Size of the observer section at the front of a world block
// (checkpoint_impl.cuh:85-107), or 0 when the block is too short or its
// history counts exceed the reader's limits.
uint64_t observer_bytes(const unsigned char *w, uint64_t size) {
    uint64_t off = kTotalsBytes;
    if (size < off + 4) return 0;
    const uint32_t behaviour = load_at<uint32_t>(w + off);
    if (behaviour > 10002) return 0;
    off += 4 + uint64_t(behaviour) * (8 + kTotalsBytes);
    if (size < off + 4) return 0;
    const uint32_t bites = load_at<uint32_t>(w + off);
    if (bites > 1024) return 0;
    off += 4 + uint64_t(bites) * (8 + kBiteBytes) + 8;
    return size < off ? 0 : off;
}

// Living animals' brain sizes must add up to the record counts
// (checkpoint_impl.cuh:155-166, 221-233).
void check_animals(const unsigned char *animals, uint64_t neurons, uint64_t synapses,
                   std::vector<uint8_t> *alive, const char *message) {
    uint64_t n = 0, s = 0;
    if (alive) alive->assign(kP, 0);
    for (uint64_t a = 0; a < kP; ++a) {
        const unsigned char *r = animals + a * kAnimalBytes;
        if (load_at<int32_t>(r + kAliveOffset) != 1) continue;
        const int32_t nc = load_at<int32_t>(r + kNeuronCountOffset);
        const int32_t sc = load_at<int32_t>(r + kSynapseCountOffset);
        if (nc < 0 || sc < 0) throw std::runtime_error("Invalid checkpoint animal counts");
        n += uint64_t(nc);
        s += uint64_t(sc);
        if (alive) (*alive)[a] = 1;
    }
    if (n != neurons || s != synapses) throw std::runtime_error(message);
}

// checkpoint_impl.cuh:247-256
template <class R>
void check_neurons(const R *neurons, uint64_t nn, const std::vector<uint8_t> &alive) {
    int prev = -1;
    for (uint64_t i = 0; i < nn; ++i) {
        const auto &r = neurons[i];
        if (r.slot <= prev || uint64_t(r.slot) >= kNeuronCap || r.owner < 0 || uint64_t(r.owner) >= kP ||
            !alive[r.owner] || r.next < -1 || r.next >= int64_t(kNeuronCap) || r.port < -1 ||
            r.port >= int(kS + kM))
            throw std::runtime_error("Invalid checkpoint neuron record");
        prev = r.slot;
    }
}

// checkpoint_impl.cuh:280-291
void check_synapses(const ns_checkpoint_synapse *synapses, uint64_t ns, const std::vector<uint8_t> &alive) {
    int prev = -1;
    for (uint64_t i = 0; i < ns; ++i) {
        const auto &r = synapses[i];
        if (r.slot <= prev || uint64_t(r.slot) >= kSynapseCap || r.owner < 0 || uint64_t(r.owner) >= kP ||
            !alive[r.owner] || r.next < -1 || r.next >= int64_t(kSynapseCap) || r.gene.pre < 0 ||
            uint64_t(r.gene.pre) >= kNeuronCap || r.gene.post < 0 || uint64_t(r.gene.post) >= kNeuronCap)
            throw std::runtime_error("Invalid checkpoint synapse record");
        prev = r.slot;
    }
}

struct Checkpoint {
    ns_checkpoint_header header{};
    uint64_t history = 9;
    std::vector<unsigned char> world;
    std::vector<ns_checkpoint_neuron> neurons;      // 9-slot files
    std::vector<ns_checkpoint_neuron17> neurons17;  // 17-slot files
    std::vector<ns_checkpoint_synapse> synapses;
};

// The whole file, checked in the reference's read order
// (main.cu:1986-2000, checkpoint_impl.cuh:201-366).
Checkpoint *read_checkpoint(const char *path) {
    if (!path) throw std::runtime_error("Invalid checkpoint path");
    GzFile in;
    in.f = gzopen(path, "rb");
    if (!in.f) throw std::runtime_error(std::string("Cannot open checkpoint: ") + path);
    gzbuffer(in.f, unsigned(kChunk));
    auto ck = std::make_unique<Checkpoint>();
    in.read(&ck->header, sizeof(ck->header));
    ck->history = header_history(ck->header);
    check_header(ck->header, ck->history ? ck->history : 9);
    // observer section, read field by field for its limits
    std::vector<unsigned char> &w = ck->world;
    auto take = [&](uint64_t bytes) {
        const size_t at = w.size();
        w.resize(at + bytes);
        in.read(w.data() + at, bytes);
        return at;
    };
    take(kTotalsBytes);
    size_t at = take(4);                  // take() may move the block: offsets, not pointers
    const uint32_t behaviour = load_at<uint32_t>(w.data() + at);
    if (behaviour > 10002) throw std::runtime_error("Invalid checkpoint behavior history");
    take(uint64_t(behaviour) * (8 + kTotalsBytes));
    at = take(4);
    const uint32_t bites = load_at<uint32_t>(w.data() + at);
    if (bites > 1024) throw std::runtime_error("Invalid checkpoint bite history");
    take(uint64_t(bites) * (8 + kBiteBytes));
    take(8);
    const size_t arrays = take(kArrayBytes);
    std::vector<uint8_t> alive;
    check_animals(w.data() + arrays, ck->header.neuron_count, ck->header.synapse_count, &alive,
                  "Checkpoint brain allocation counts disagree");
    if (ck->history == 17) {
        ck->neurons17.resize(ck->header.neuron_count);
        in.read(ck->neurons17.data(), ck->neurons17.size() * sizeof(ns_checkpoint_neuron17));
        check_neurons(ck->neurons17.data(), ck->neurons17.size(), alive);
    } else {
        ck->neurons.resize(ck->header.neuron_count);
        in.read(ck->neurons.data(), ck->neurons.size() * sizeof(ns_checkpoint_neuron));
        check_neurons(ck->neurons.data(), ck->neurons.size(), alive);
    }
    ck->synapses.resize(ck->header.synapse_count);
    in.read(ck->synapses.data(), ck->synapses.size() * sizeof(ns_checkpoint_synapse));
    check_synapses(ck->synapses.data(), ck->synapses.size(), alive);
    if (ck->header.flags & 64u) {   // plasticity-cohort extension
        at = take(32);
        const uint64_t tag = load_at<uint64_t>(w.data() + at), version = load_at<uint64_t>(w.data() + at + 8),
                       size = load_at<uint64_t>(w.data() + at + 16), count = load_at<uint64_t>(w.data() + at + 24);
        if (tag != kCohortTag || version != 1 || size != kCohortBytes || count != kP)
            throw std::runtime_error("Incompatible plasticity intervention extension");
        take(kP * kCohortBytes);
    }
    unsigned char extra;
    if (gzread(in.f, &extra, 1) != 0) throw std::runtime_error("Unexpected trailing checkpoint data");
    return ck.release();
}

// checkpoint_impl.cuh:135-199 (file order) and 385-413 (temporary file,
// gzip level 1, fsync, rename, directory fsync).
template <class R>
void write_checkpoint(const char *path, const ns_checkpoint_header &given, const unsigned char *world,
                      uint64_t world_bytes, const R *neurons, uint64_t nn,
                      const ns_checkpoint_synapse *synapses, uint64_t ns) {
    if (!path || (world_bytes && !world) || (nn && !neurons) || (ns && !synapses))
        throw std::runtime_error("Invalid checkpoint arguments");
    ns_checkpoint_header h = given;
    h.neuron_count = uint32_t(nn);
    h.synapse_count = uint32_t(ns);
    if (nn > kNeuronCap || ns > kSynapseCap) throw std::runtime_error("Incompatible or invalid checkpoint header");
    check_header(h, history_of<R>());
    const uint64_t obs = observer_bytes(world, world_bytes);
    const uint64_t want = obs + kArrayBytes + ((h.flags & 64u) ? kExtensionBytes : 0);
    if (!obs || world_bytes != want) throw std::runtime_error("Invalid checkpoint world sections");
    std::vector<uint8_t> alive;
    check_animals(world + obs, nn, ns, &alive, "Checkpoint pool ownership disagrees with living brains");
    check_neurons(neurons, nn, alive);
    check_synapses(synapses, ns, alive);
    if (h.flags & 64u) {
        const unsigned char *e = world + obs + kArrayBytes;
        if (load_at<uint64_t>(e) != kCohortTag || load_at<uint64_t>(e + 8) != 1 ||
            load_at<uint64_t>(e + 16) != kCohortBytes || load_at<uint64_t>(e + 24) != kP)
            throw std::runtime_error("Incompatible plasticity intervention extension");
    }
    const std::string temporary = std::string(path) + ".tmp";
    try {
        {
            GzFile out;
            out.f = gzopen(temporary.c_str(), "wb1");
            if (!out.f) throw std::runtime_error("Cannot open checkpoint for writing: " + temporary);
            gzbuffer(out.f, unsigned(kChunk));
            out.write(&h, sizeof(h));
            out.write(world, obs + kArrayBytes);
            out.write(neurons, nn * sizeof(R));
            out.write(synapses, ns * sizeof(ns_checkpoint_synapse));
            if (h.flags & 64u) out.write(world + obs + kArrayBytes, kExtensionBytes);
            const int rc = gzclose(out.f);
            out.f = nullptr;
            if (rc != Z_OK) throw std::runtime_error("Cannot finish checkpoint: " + temporary);
        }
        int fd = open(temporary.c_str(), O_RDONLY);
        if (fd < 0 || fsync(fd) != 0) {
            if (fd >= 0) close(fd);
            throw std::runtime_error("Cannot sync checkpoint: " + temporary);
        }
        close(fd);
        std::filesystem::rename(temporary, path);
    } catch (...) {
        std::error_code ec;                   // no partial file is left behind
        std::filesystem::remove(temporary, ec);
        throw;
    }
    std::filesystem::path parent = std::filesystem::path(path).parent_path();
    if (parent.empty()) parent = ".";
    const int fd = open(parent.c_str(), O_RDONLY | O_DIRECTORY);
    if (fd >= 0) {
        (void)fsync(fd);
        close(fd);
    }
}

}  // namespace
}  // namespace nsa

using namespace nsa;

extern "C" {

uint64_t ns_checkpoint_layout(void) { return layout_hash(9); }

uint64_t ns_checkpoint_layout_ex(int history) {
    return history == 9 || history == 17 ? layout_hash(uint64_t(history)) : 0;
}

void *ns_checkpoint_open(const char *path) {
    try {
        return read_checkpoint(path);
    } catch (const std::exception &e) {
        set_error(e.what());
        return nullptr;
    }
}

int ns_checkpoint_history(void *ckpt) {
    return ckpt ? int(static_cast<Checkpoint *>(ckpt)->history) : -1;
}

}  // extern "C"

template <class R>
static int get_checkpoint(void *ckpt, ns_checkpoint_header *header, const void **world, size_t *world_bytes,
                          const R **neurons, const ns_checkpoint_synapse **synapses) {
    try {
        if (!ckpt) throw std::runtime_error("Invalid checkpoint handle");
        const auto &c = *static_cast<Checkpoint *>(ckpt);
        if (c.history != history_of<R>())
            throw std::runtime_error(c.history == 17 ? "Checkpoint has 17-slot neuron records"
                                                     : "Checkpoint has 9-slot neuron records");
        if (header) *header = c.header;
        if (world) *world = c.world.data();
        if (world_bytes) *world_bytes = c.world.size();
        if (neurons) {
            if constexpr (history_of<R>() == 17) *neurons = c.neurons17.data();
            else *neurons = c.neurons.data();
        }
        if (synapses) *synapses = c.synapses.data();
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

extern "C" {

int ns_checkpoint_get(void *ckpt, ns_checkpoint_header *header, const void **world, size_t *world_bytes,
                      const ns_checkpoint_neuron **neurons, const ns_checkpoint_synapse **synapses) {
    return get_checkpoint(ckpt, header, world, world_bytes, neurons, synapses);
}

int ns_checkpoint_get17(void *ckpt, ns_checkpoint_header *header, const void **world, size_t *world_bytes,
                        const ns_checkpoint_neuron17 **neurons, const ns_checkpoint_synapse **synapses) {
    return get_checkpoint(ckpt, header, world, world_bytes, neurons, synapses);
}

void ns_checkpoint_close(void *ckpt) { delete static_cast<Checkpoint *>(ckpt); }

}  // extern "C"

template <class R>
static int save_checkpoint(const char *path, const ns_checkpoint_header *header, const void *world,
                           size_t world_bytes, const R *neurons, int neuron_count,
                           const ns_checkpoint_synapse *synapses, int synapse_count) {
    try {
        if (!header || neuron_count < 0 || synapse_count < 0)
            throw std::runtime_error("Invalid checkpoint arguments");
        write_checkpoint(path, *header, static_cast<const unsigned char *>(world), world_bytes, neurons,
                         uint64_t(neuron_count), synapses, uint64_t(synapse_count));
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

extern "C" {

int ns_checkpoint_save(const char *path, const ns_checkpoint_header *header, const void *world,
                       size_t world_bytes, const ns_checkpoint_neuron *neurons, int neuron_count,
                       const ns_checkpoint_synapse *synapses, int synapse_count) {
    return save_checkpoint(path, header, world, world_bytes, neurons, neuron_count, synapses, synapse_count);
}

int ns_checkpoint_save17(const char *path, const ns_checkpoint_header *header, const void *world,
                         size_t world_bytes, const ns_checkpoint_neuron17 *neurons, int neuron_count,
                         const ns_checkpoint_synapse *synapses, int synapse_count) {
    return save_checkpoint(path, header, world, world_bytes, neurons, neuron_count, synapses, synapse_count);
}

int ns_pool_save_checkpoint(void *pool, const char *path, const ns_checkpoint_header *header,
                            const void *world, size_t world_bytes) {
    try {
        if (!pool || !header) throw std::runtime_error("Invalid checkpoint arguments");
        ns_pool_stats st{};
        if (ns_pool_info(pool, &st)) return -1;
        std::vector<ns_checkpoint_synapse> s(size_t(std::max<int64_t>(st.synapse_records, 1)));
        int nn = 0, ns = 0;
        const size_t ncap = size_t(std::max<int64_t>(st.neuron_records, 1));
        if (ns_pool_history(pool) == 17) {
            std::vector<ns_checkpoint_neuron17> n(ncap);
            if (ns_pool_export_records17(pool, n.data(), int(n.size()), s.data(), int(s.size()), &nn, &ns))
                return -1;
            write_checkpoint(path, *header, static_cast<const unsigned char *>(world), world_bytes, n.data(),
                             uint64_t(nn), s.data(), uint64_t(ns));
        } else {
            std::vector<ns_checkpoint_neuron> n(ncap);
            if (ns_pool_export_records(pool, n.data(), int(n.size()), s.data(), int(s.size()), &nn, &ns))
                return -1;
            write_checkpoint(path, *header, static_cast<const unsigned char *>(world), world_bytes, n.data(),
                             uint64_t(nn), s.data(), uint64_t(ns));
        }
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

int ns_pool_load_checkpoint(void *pool, const char *path, ns_checkpoint_header *header, void *world,
                            size_t world_capacity, size_t *world_bytes) {
    try {
        if (!pool) throw std::runtime_error("Invalid checkpoint arguments");
        const int history = ns_pool_history(pool);
        if (history < 0) return -1;
        std::unique_ptr<Checkpoint> c(read_checkpoint(path));
        if (world && world_capacity < c->world.size()) throw std::runtime_error("World buffer too small");
        if (int(c->history) != history) throw std::runtime_error("Checkpoint ring length differs from the pool's");
        if (history == 17 ? ns_pool_import_records17(pool, c->neurons17.data(), int(c->neurons17.size()),
                                                     c->synapses.data(), int(c->synapses.size()))
                          : ns_pool_import_records(pool, c->neurons.data(), int(c->neurons.size()),
                                                   c->synapses.data(), int(c->synapses.size())))
            return -1;
        if (header) *header = c->header;
        if (world) std::memcpy(world, c->world.data(), c->world.size());
        if (world_bytes) *world_bytes = c->world.size();
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

}  // extern "C"
