// Launch interface of the tick kernels (host side of ns_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ns_device.h"

namespace nsa {

// Per-pack limits shared by the packer and the resident kernel's LDS arrays.
constexpr int kPackBrains = 128;   // brains per pack
constexpr int kPackMotors = 512;   // motor slots per pack
constexpr uint16_t kNoMotor = 0xffff;   // motor slot without a neuron (world ports)

// A pack: whole brains [b0, b0+nb) whose neurons [n0, n0+nn) and synapses
// [s0, s0+ns) are contiguous in the pools; nm = nb * outputs motor slots.
struct Pack {
    int32_t n0, nn, s0, ns, b0, nb, nm, pad;
};

// Element ranges for the plain per-tick kernels.
struct PlainRange {
    int64_t n_lo, n_hi, s_lo, s_hi;
    int brains;            // brain table size (need-mask decides who runs)
    int slots;             // motor slots covered (brains * outputs)
    const int32_t *slow;   // positions of type-2/3 neurons
    int slow_count;
};

using ResidentLauncher = hipError_t (*)(const Pool &, const Pack *, int, const int32_t *, int,
                                        const uint16_t *, int, const float *, float, float *, int,
                                        hipStream_t);
// (LDS tiles rewrite the slot order of the packs they run, see k_resident_lds)
using LdsLauncher = hipError_t (*)(const Pool &, const Pack *, int, const int32_t *, int,
                                   uint16_t *, int, const float *, float, float *, int, int,
                                   hipStream_t);
// A compiled tile.  Register-slot tiles (synapses_per_thread > 0) keep
// synapse state in registers; LDS tiles keep it in LDS with room for
// lds_plastic / lds_static synapses per pack (with / without plasticity
// traces).
struct TileShape {
    int threads, neurons_per_thread, synapses_per_thread, min_waves;
    int lds_plastic, lds_static;
    ResidentLauncher launch;
    LdsLauncher launch_plastic, launch_static;
};
extern const TileShape kTiles[];
extern const int kNumTiles;

void launch_reset(const Pool &p, int64_t nn, int64_t ns, int nb, const uint32_t *bseed,
                  const uint32_t *norig, const uint32_t *sorig, hipStream_t st);
// Reset of the slots list[0..count) with seeds list[count..2 count) (device
// memory), as launch_reset resets a run-flagged brain.
void launch_reset_slots(const Pool &p, const int32_t *list, int count, const uint32_t *norig,
                        const uint32_t *sorig, hipStream_t st);
// One staged copy of a slot placement: `bytes` bytes from offset `src` of the
// staged buffer to device address `dst`.
struct ScatterEntry {
    uint64_t dst, src, bytes;
};
void launch_scatter(const ScatterEntry *entries, int n, const unsigned char *base, hipStream_t st);
// Stable class order of the neuron slots of packs[list[0..count)] into perm
// (pack-relative indices at perm[pack.n0 ...]), as the pool build computes it.
void launch_pack_order(const uint32_t *nflags, const uint8_t *modbits, const Pack *packs,
                       const int32_t *list, int count,
                       uint16_t *perm, hipStream_t st);
void launch_fill(float *x, int64_t n, float v, hipStream_t st);
void launch_clear_step(float *out, int64_t nout, int32_t *status, int64_t nstatus, hipStream_t st);
void launch_fill_i32(int32_t *x, int64_t n, int32_t v, hipStream_t st);
void launch_add_counts(int32_t *total, const float *counts, int64_t n, hipStream_t st);
void launch_per_tick(const Pool &p, const PlainRange &r, int ticks, uint8_t need,
                     const float *input, float gain, float *out, int homeostasis,
                     hipStream_t st);
void launch_math_digest(int fn, uint64_t lo, uint64_t hi, int64_t chunk,
                        unsigned long long *digests, hipStream_t st);
void launch_world_homeostasis(const Pool &p, int64_t nn, const int32_t *age, hipStream_t st);

}  // namespace nsa
