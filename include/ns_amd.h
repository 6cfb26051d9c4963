/*
 * ns_amd.h — C ABI of the MI355X per-tick brain update (libns_amd.so).
 *
 * Plain C: pointers, sizes and POD records only.  Every entry point returns
 * 0 / non-null on success and -1 / NULL on failure, with the reason kept in a
 * thread-local string read by ns_error().  No exception crosses this ABI.
 *
 * Two surfaces replace the reference's two entry surfaces for the same
 * arithmetic (reference paths are relative to the neural-selection repo):
 *
 *   1. Batched episodes — drop-in for gym_adapter/native.cpp (genome handles)
 *      and gym_adapter/gpu.cu (ns_gpu_*), bound from Python by ctypes in
 *      gym_adapter/gpu.py:35-38 and gym_adapter/adapter.py:31-39.
 *   2. World pool — the neural part of src/main.cu's step loop
 *      (main.cu:2634-2768) over checkpointable neuron/synapse records
 *      (main.cu:164-177), replacing the kernels synapse_delivery,
 *      neuron_step, synapse_plasticity(_delivery), homeostasis_step.
 *
 * A third surface, serving, has no reference counterpart: it is the
 * continuous-batching episode loop that gym_adapter/batch_evaluate.py
 * approximates with lock-step masks.
 */
#ifndef NS_AMD_H
#define NS_AMD_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Record layouts.  Byte-identical to the reference's structs so that state */
/* moves between the reference and this library without conversion.        */
/* ------------------------------------------------------------------------ */

/* src/physiology.cuh:29-45 NeuronGene (124 B). */
typedef struct ns_neuron_gene {
    float capacitance_pf, leak_ns, rest_mv, threshold_mv, slope_mv;
    float adapt_ns, adapt_tau_ms, spike_adapt_pa, reset_mv;
    float afterdepol_spike_pa, afterdepol_tau_ms;
    uint8_t type, active, modulator, inhibitory_transmitter;
    float lp_scale[4];            /* LPConductance: calcium, calcium_k, h, leak */
    float prinz_g[8];             /* PrinzConductance */
    float prinz_ambient_leak_ns;
    int8_t sensory_channel;
    uint8_t pad0_[3];
    float receptor_baseline_hz, receptor_gain_hz;
    float receptor_event_ns, receptor_decay_ms;
    float target_affinity;
} ns_neuron_gene;

/* src/physiology.cuh:50-58 NeuronState (120 B).  `cell` is the union of
 * LPCell x[16] (lp_cell.cuh:16), PrinzCell {v, ca, m[7], h[4]}
 * (prinz_cell.cuh:18) and LFSCell {voltage, threshold, activated[3], open[3]}
 * (physiology.cuh:46-49). */
typedef struct ns_neuron_state {
    float voltage_mv, adaptation_pa, afterdepol_pa, inh_ns, recent_depol_mv;
    float log_ampa_scale;
    float dendrite_exc_ns[4];
    float cell[16];
    int32_t total_spikes;
    uint8_t spike, pad0_[3];
    float receptor_ns;
    uint32_t receptor_rng;
} ns_neuron_state;

/* src/physiology.cuh:59-65 SynapseGene (40 B). pre/post are brain-local
 * neuron indices on the batch surface and pool slots on the world surface. */
typedef struct ns_synapse_gene {
    float initial_weight_ns, utilization, recovery_ms, facilitation_ms;
    int32_t pre, post;
    uint8_t delay, active, plasticity, branch, release_sites, mod_receptors;
    uint8_t pad0_[2];
    float slow_recovery_ms, slow_depletion;
} ns_synapse_gene;

/* src/physiology.cuh:66-76 SynapseState (80 B). */
typedef struct ns_synapse_state {
    float weight_ns, running_utilization, pre_trace, post_trace;
    float eligible_ltp, eligible_ltd, da_age_ms;
    float spine_volume, structural_age_ms;
    float pre_age_ms, post_age_ms;
    uint32_t pad0_;
    int64_t last_arrival_tick;
    uint64_t ready_mask;
    uint32_t rng_state;
    uint8_t arrived, pad1_[3];
    float slow_resource;
    uint32_t pad2_;
} ns_synapse_state;

/* src/main.cu:164-171 CheckpointNeuron (344 B), HISTORY = 9. */
typedef struct ns_checkpoint_neuron {
    int32_t slot, owner, next, port;
    ns_neuron_gene gene;
    ns_neuron_state state;
    float pending_exc[4], pending_inh;
    uint64_t mod_event_tick[6];
    uint8_t spike_history[9];
    uint8_t pad0_[7];
} ns_checkpoint_neuron;

/* CheckpointNeuron of a world build with delays of 1-16 ms: main.cu:46
 * HISTORY = D + 1 = 17 ring slots (352 B).  The same fields at the same
 * offsets up to the ring.  A layout version of its own: HISTORY and the
 * record size enter checkpoint_layout (checkpoint_impl.cuh:21-32), so its
 * files carry ns_checkpoint_layout_ex(17). */
typedef struct ns_checkpoint_neuron17 {
    int32_t slot, owner, next, port;
    ns_neuron_gene gene;
    ns_neuron_state state;
    float pending_exc[4], pending_inh;
    uint64_t mod_event_tick[6];
    uint8_t spike_history[17];
    uint8_t pad0_[7];
} ns_checkpoint_neuron17;

/* src/main.cu:173-177 CheckpointSynapse (136 B). */
typedef struct ns_checkpoint_synapse {
    int32_t slot, owner, next;
    ns_synapse_gene gene;
    uint32_t pad0_;
    ns_synapse_state state;
} ns_checkpoint_synapse;

/* Algorithmic bytes per unit of work of the per-tick update, the roofline's
 * numerator (structure-of-arrays, minimal, per tick, against ~8 TB/s):
 *   neuron-update (AdEx types 0/1/4): 82 B of state read and written, 4 B
 *     log AMPA scale, 48 B of genes, 40 B of pending inputs read and cleared,
 *     4 B sense level;
 *   synapse-update, plasticity on: 20 B of genes, 40 B of plasticity state
 *     read and written, about 6 B of post-neuron gather; plasticity off: 12 B;
 *   event (an arrival): 58 B of short-term state read and written, 20 B of
 *     short-term genes. */
#define NS_BYTES_PER_NEURON_UPDATE 178
#define NS_BYTES_PER_SYNAPSE_UPDATE_PLASTIC 66
#define NS_BYTES_PER_SYNAPSE_UPDATE_STATIC 12
#define NS_BYTES_PER_EVENT 78

/* ------------------------------------------------------------------------ */
/* Errors                                                                    */
/* ------------------------------------------------------------------------ */

/* gym_adapter/native.cpp:45 — message of the last failure on this thread. */
const char *ns_error(void);

/* ------------------------------------------------------------------------ */
/* 1. Batched-episode surface                                               */
/* ------------------------------------------------------------------------ */

/* gym_adapter/native.cpp:46-82 ns_create.  Host-side genome handle; no
 * device work.  Neuron genes are count x 10 floats {profile, sensory
 * channel, active, inhibitory, modulator, capacitance multiplier, leak
 * multiplier, threshold mV, receptor gain Hz, receptor event nS}; synapse
 * genes are edges x 11 floats {pre, post, weight nS, active, delay 1..8,
 * branch, U, recovery ms, facilitation ms, plasticity 0..5, modulator
 * receptor mask}.  Same validation and messages as the reference. */
void *ns_create(int count, int edges, int channels, int outputs,
                const float *neuron_genes, const float *synapse_genes,
                const int *motors);

/* gym_adapter/native.cpp:83-87 — original indices used to seed reset RNGs
 * after subgraph compilation (seed + original*7919 / + original*104729). */
void ns_original_indices(void *brain, const uint32_t *neurons,
                         const uint32_t *synapses);

/* gym_adapter/native.cpp:88 */
void ns_destroy(void *brain);

/* gym_adapter/native.cpp:89-94 ns_reset: tick 0 and the initial state of
 * every neuron and synapse (physiology.cuh:179-202, 393-404) seeded
 * seed + original * 7919 / + original * 104729, with the original indices of
 * ns_original_indices when set, else the genome order.  Host only. */
void ns_reset(void *brain, uint32_t seed);

/* gym_adapter/native.cpp:95-109 ns_step: `ticks` (1..10000) one-ms ticks of
 * the handle's brain with constant encoded input (channels floats, finite, in
 * [0,1]) and input gain (>= 0); output = per-motor spike counts; the same
 * checks and messages.  The handle's state moves through a one-brain device
 * pool (made on the first call, held until ns_destroy) and back,
 * synchronously, on the same kernels as ns_gpu_step.  A nonfinite voltage
 * returns -1 ("Nonfinite neural state") with every tick run and that state
 * kept, as the reference's device batch does (gym_adapter/gpu.cu:48-64); the
 * reference's CPU path stops inside the tick instead. */
int ns_step(void *brain, const float *input, int ticks, float gain, float *output);

/* ns_step without allocating: no counterpart in the reference; for callers
 * that cannot allocate device memory (code under graph capture, tight
 * serving loops, harnesses that forbid it).  The same checks, kernels and
 * results as ns_step, with the one-brain pool laid out afresh on each call in
 * a static device array of NS_STEP_STATIC_BYTES per GPU (held until the
 * process ends) instead of memory of the handle's own.  Brains whose pool
 * does not fit (about 300 B per neuron and 120 B per synapse: 4,096 neurons
 * with 16,384 synapses fit) are refused with an error and the handle
 * unchanged.  Calls on one GPU share the array and the last call's contents
 * stay in it; this process's calls are serialised per GPU, and overlapping
 * calls from outside that lock (another copy of the library in the process)
 * would race. */
#define NS_STEP_STATIC_BYTES (4u << 20)
int ns_step_static(void *brain, const float *input, int ticks, float gain, float *output);

/* A handle's state (no reference counterpart): neuron and synapse records
 * in genome order, the spike shift registers (bit k = spike k ticks before
 * the last completed tick) and the tick.  Any pointer may be NULL. */
int ns_export(void *brain, ns_neuron_state *neurons, ns_synapse_state *synapses,
              uint32_t *spike_bits, int64_t *tick);

/* gym_adapter/gpu.cu:67-109 ns_gpu_create.  Borrows `count` handles made by
 * ns_create (or ns_create_ex); the caller may destroy them afterwards.  All
 * brains must share channels and outputs.  Uploads genes into the
 * structure-of-arrays device pool and builds the incoming CSR in ascending
 * edge order. */
void *ns_gpu_create(void **brains, int count);

/* gym_adapter/gpu.cu:110 */
void ns_gpu_destroy(void *batch);

/* gym_adapter/gpu.cu:111-113 — reset masked brains: tick 0, fresh state
 * seeded seeds[i] + original_neuron*7919 / + original_synapse*104729. */
int ns_gpu_reset(void *batch, const uint8_t *mask, const uint32_t *seeds);

/* gym_adapter/gpu.cu:114-123 — advance masked brains `ticks` (1..10000)
 * one-ms ticks with constant encoded input (count x channels, finite, in
 * [0,1]) and input gain (>= 0); output = per-motor spike counts
 * (count x outputs), status[i] = -1 if brain i reached a nonfinite voltage.
 * Synchronous, like the reference. */
int ns_gpu_step(void *batch, const float *input, const uint8_t *mask,
                int ticks, float gain, float *output, int *status);

/* ---- extensions of the batch surface (no reference counterpart) -------- */

/* ns_create with the delay bound raised from 8 up to NS_MAX_DELAY ticks
 * (a delay ring longer than the reference's 9 slots). */
#define NS_MAX_DELAY 32
void *ns_create_ex(int count, int edges, int channels, int outputs,
                   const float *neuron_genes, const float *synapse_genes,
                   const int *motors, int max_delay);

/* Build a batch from many genomes in one call: brain i has neuron_counts[i]
 * neurons and edge_counts[i] synapses; genes and motors are concatenated in
 * brain order with the ns_create encodings (synapse pre/post brain-local).
 * world_rules = 1 selects src/main.cu's neuron_step rules (230 pA per
 * sensory level, receptor events only for sensory channels < 6, inactive
 * neurons keep their history) instead of core.cuh's. */
void *ns_gpu_create_bulk(int count, const int *neuron_counts,
                         const int *edge_counts, int channels, int outputs,
                         const float *neuron_genes, const float *synapse_genes,
                         const int *motors, int max_delay, int world_rules);

/* Execution form of ns_gpu_step:
 *   NS_FORM_RESIDENT — one launch per call; each workgroup keeps a pack of
 *                      brains in registers/LDS for all ticks (default);
 *   NS_FORM_PER_TICK — plain form: delivery, neuron and plasticity kernels
 *                      launched once per tick over the global pools.
 * Both give bit-identical results. */
#define NS_FORM_RESIDENT 0
#define NS_FORM_PER_TICK 1
int ns_gpu_set_form(void *batch, int form);

/* Pack tuning: workgroup size, neurons per thread, synapses per thread
 * (register-slot kernel) or 0 with `lds_synapses` = the LDS kernel's
 * per-pack synapse capacity (plastic pools), and the launch bound's minimum
 * waves per SIMD (caps registers; 0 = any compiled variant).  Only compiled
 * shapes are accepted.  ns_gpu_autotune times every compiled shape on the
 * current pools (state is saved and restored around the trials) and keeps
 * the fastest. */
int ns_gpu_set_tiles(void *batch, int threads, int neurons_per_thread,
                     int synapses_per_thread, int min_waves, int lds_synapses);
int ns_gpu_autotune(void *batch, int ticks, int *threads,
                    int *neurons_per_thread, int *synapses_per_thread,
                    int *min_waves, int *lds_synapses);

/* Device-resident step: input/mask/output/status are device pointers and
 * `stream` is a hipStream_t (NULL = default stream).  Asynchronous.  No
 * input validation (the caller owns the device buffers). */
int ns_gpu_step_device(void *batch, const float *input, const uint8_t *mask,
                       int ticks, float gain, float *output, int *status,
                       void *stream);

/* Copy state out / in, AoS in genome order.  history9 is count_neurons x 9
 * bytes laid out like gym_adapter/core.cuh's history[i][9] ring
 * (slot = tick % 9).  ticks: one int64 per brain. Any pointer may be NULL.
 * Nine slots hold the whole delay history only for delays up to 8 ticks:
 * with a non-NULL history9, a batch made with a delay bound above 8
 * (ns_create_ex / ns_gpu_create_bulk max_delay > 8) is refused; its rings
 * travel as spike bit registers (ns_gpu_export/import_spike_bits). */
int ns_gpu_export(void *batch, ns_neuron_state *neurons,
                  ns_synapse_state *synapses, uint8_t *history9,
                  int64_t *ticks);
int ns_gpu_import(void *batch, const ns_neuron_state *neurons,
                  const ns_synapse_state *synapses, const uint8_t *history9,
                  const int64_t *ticks);

/* Raw 32-bit spike shift registers (bit k = spike k ticks before the last
 * completed tick), one per neuron in genome order.  Import writes them back,
 * so the whole ring of a batch with delays up to NS_MAX_DELAY round-trips.
 * No counterpart in the reference (its rings have 9 slots). */
int ns_gpu_export_spike_bits(void *batch, uint32_t *bits);
int ns_gpu_import_spike_bits(void *batch, const uint32_t *bits);

typedef struct ns_batch_info {
    int64_t brains, neurons, synapses;
    int64_t resident_packs, fallback_brains;
    int64_t events;             /* delivered arrivals since create */
    int threads, neurons_per_thread, synapses_per_thread, form;
    int64_t device_bytes;
    /* with timing on: HIP-event time of the resident tick kernel launches
     * and of the per-tick kernel sequences, on the stream they ran on */
    double resident_ms, per_tick_ms;
    int64_t resident_launches, per_tick_calls;
    int64_t resident_events;    /* arrivals delivered inside resident launches */
    int64_t fallback_neurons, fallback_synapses;   /* on the per-tick path */
    int64_t big_brains, big_neurons;   /* one workgroup each in the resident launch */
    int min_waves;
    int lds_synapses;           /* LDS synapse capacity per pack, 0 = register slots */
    /* slots (ns_gpu_reserve / ns_gpu_replace) */
    int64_t slots, free_slots;  /* pool brain slots; slots without a brain */
    int64_t slot_neurons, slot_synapses;   /* total slot capacity */
    int64_t placements, relayouts;         /* replacements placed in free slots / by rebuild */
    int64_t graph_captures, graph_launches;  /* ns_gpu_set_graph */
    int64_t staging_grows;      /* replacement staging buffers enlarged (host and device) */
    int64_t appended_slots;     /* slots added past the last one for newborns no free slot fit */
    /* neuron ticks the resident launch skipped: an AdEx cell without sensory
     * input whose previous tick changed no state, with no arrival (the tick
     * is then the identity) */
    int64_t skipped_neuron_ticks;
} ns_batch_info;
int ns_gpu_info(void *batch, ns_batch_info *info);

/* Brain turnover (births and deaths) without rebuilding the pools.  These
 * have no counterpart in the reference's batch ABI; they serve the world's
 * births/deaths (src/main.cu:364-454, 1302-1338: slot claims and genome
 * lists) and continuous batching.
 *
 * ns_gpu_reserve: rebuild the pools (state kept) with slot capacities rounded
 * up a geometric ladder of ratio 1 + slack and spare empty slots for a
 * fraction `spare` of the batch, spread over the brain sizes.
 *
 * ns_gpu_replace: brains indices[0..count) take the genomes `handles` (made by
 * ns_create; copied, the caller keeps them) with fresh state seeded like
 * ns_gpu_reset with seeds[j].  Each newborn goes to the best-fitting free slot
 * (spares, vacancies, or the slot it replaces); one that fits none gets a new
 * slot past the last one, in room the pool keeps for that (about 1/16 of its
 * neurons and synapses); the pools are rebuilt (state kept) only when that
 * room is short, or for LP/Prinz cells.  Enqueued on `stream`
 * (a hipStream_t, NULL = the batch's own stream) after earlier work there;
 * later ns_gpu_step_device calls on that stream see the new brains.  Results
 * are those of a batch created with the new genomes and reset with seeds. */
int ns_gpu_reserve(void *batch, float slack, float spare);
int ns_gpu_replace(void *batch, int count, const int *indices, void **handles,
                   const uint32_t *seeds, void *stream);

/* Record HIP events around every tick-kernel launch of this batch (adds one
 * event pair per launch; totals read back by ns_gpu_info). */
int ns_gpu_set_timing(void *batch, int on);

/* Graph mode (on = 1): each step (ns_gpu_step, ns_gpu_step_device) runs as
 * one hipGraph launch on its stream: the output clears and the whole tick
 * loop (the resident launch, or every per-tick kernel of the plain form,
 * plus the LP/Prinz slots' per-tick kernels) are captured once and replayed
 * while the step's arguments (ticks, input and output buffers, gain) and the
 * pool layout stay the same; a rebuild, a new tile or form, or other
 * arguments capture again.  Replacements and run masks stay eager work
 * between replays.  Results are those of eager steps.  No counterpart in the
 * reference; the per-tick form issues dozens of launches per step, which a
 * graph replays at once. */
int ns_gpu_set_graph(void *batch, int on);

/* Diagnostic: evaluate the device expf (fn = 0) or logf (fn = 1) on every
 * float bit pattern u in [lo, hi) and write one digest per `chunk` inputs:
 * sum of bits(f(u)) * (u | 1) mod 2^64, NaN results as 0x7fc00000.  Lets a
 * host check the device math against its libm without moving 2^32 floats. */
int ns_device_math_digest(int fn, uint64_t lo, uint64_t hi, int64_t chunk,
                          uint64_t *digests);

/* ------------------------------------------------------------------------ */
/* 2. World-pool surface (src/main.cu)                                       */
/* ------------------------------------------------------------------------ */

/* Pool over checkpoint records.  Capacities bound the record slots
 * (main.cu:33-35: 32768 animals, 4194304 neurons, 8388608 synapses).  Its
 * records have the reference's 9-slot rings (delays 1-8). */
void *ns_pool_create(int animal_capacity, int neuron_capacity,
                     int synapse_capacity);
/* The same with a ring of `history` slots: 9 (as ns_pool_create) or 17, the
 * layout of a build with delays of 1-16 ms (main.cu:46 HISTORY = D + 1),
 * whose records are ns_checkpoint_neuron17 and go through the *_records17
 * calls below; the 9-slot calls refuse such a pool and the 17-slot calls a
 * 9-slot pool.  An inactive neuron keeps its ring slot for `history` ticks
 * (main.cu:718). */
void *ns_pool_create_ex(int animal_capacity, int neuron_capacity,
                        int synapse_capacity, int history);
int ns_pool_history(void *pool);       /* 9 or 17; -1 on a bad handle */
void ns_pool_destroy(void *pool);

/* Load occupied neuron/synapse records (ascending slot order, as written by
 * checkpoint_impl.cuh:135-220) into per-animal device segments.  Validates
 * like checkpoint_impl.cuh:272-311.  Replaces all previous contents. */
int ns_pool_import_records(void *pool, const ns_checkpoint_neuron *neurons,
                           int neuron_count,
                           const ns_checkpoint_synapse *synapses,
                           int synapse_count);

/* Write the records back in ascending slot order (main.cu:197-264). */
int ns_pool_export_records(void *pool, ns_checkpoint_neuron *neurons,
                           int neuron_capacity,
                           ns_checkpoint_synapse *synapses,
                           int synapse_capacity, int *neuron_count,
                           int *synapse_count);

/* One world step of neural ticks (main.cu:2702-2728): ticks first_tick ..
 * first_tick+15 for every imported animal.  senses: animal_capacity x 19
 * levels (main.cu:729-737); motors: animal_capacity x 5 spike counts,
 * incremented (main.cu:743); frozen: animal_capacity flags or NULL
 * (main.cu:142-146).  Host pointers, synchronous. */
int ns_world_neural_step(void *pool, int64_t first_tick, const float *senses,
                         int *motors, const uint8_t *frozen);

/* ns_world_neural_step over device arrays, as main.cu keeps them (d.senses
 * written by update_senses, d.motors read by body_step): senses and motors are
 * device pointers (animal_capacity x 19 floats, animal_capacity x 5 ints,
 * incremented on the device), frozen a host array or NULL.  Enqueued on
 * `stream` (a hipStream_t, NULL = the pool's own) after earlier work there,
 * with no host synchronisation except on the first step after an import or
 * after skipped ticks (ring alignment) and when the run or freeze flags
 * change.  Every other ns_pool_ and ns_world_ call waits for it. */
int ns_world_neural_step_device(void *pool, int64_t first_tick,
                                const float *senses, int *motors,
                                const uint8_t *frozen, void *stream);

/* homeostasis_step (main.cu:1287-1300): age[a] in world steps. */
int ns_world_homeostasis(void *pool, const int *age, const uint8_t *frozen);

/* Births and deaths between world steps (main.cu:1302-1338 reclaim_dead_
 * genomes releases a dead animal's neuron and synapse slots; main.cu:364-454
 * claim_slot and the genome lists give a newborn its slots).  Animals
 * animals[0..animal_count) lose all their records; the given records (ascending
 * slot order per array, each owned by a listed animal, checked like
 * ns_pool_import_records, slots free once the listed animals' are released)
 * become their new brains, state and rings as the records hold them.  A
 * listed animal with no records just dies.  Each newborn goes to the best-
 * fitting free pool slot; the pool is rebuilt (state kept) only when none
 * fits or for LP/Prinz cells.  The result equals exporting the records,
 * applying the same change and importing them.  Synchronous. */
int ns_pool_replace_records(void *pool, int animal_count, const int *animals,
                            const ns_checkpoint_neuron *neurons,
                            int neuron_count,
                            const ns_checkpoint_synapse *synapses,
                            int synapse_count);

/* The record calls of a 17-slot pool (ns_pool_create_ex(..., 17)): the same
 * checks and results as ns_pool_import_records, ns_pool_export_records and
 * ns_pool_replace_records, with ns_checkpoint_neuron17 records and synapse
 * delays of 1-16. */
int ns_pool_import_records17(void *pool, const ns_checkpoint_neuron17 *neurons,
                             int neuron_count,
                             const ns_checkpoint_synapse *synapses,
                             int synapse_count);
int ns_pool_export_records17(void *pool, ns_checkpoint_neuron17 *neurons,
                             int neuron_capacity,
                             ns_checkpoint_synapse *synapses,
                             int synapse_capacity, int *neuron_count,
                             int *synapse_count);
int ns_pool_replace_records17(void *pool, int animal_count, const int *animals,
                              const ns_checkpoint_neuron17 *neurons,
                              int neuron_count,
                              const ns_checkpoint_synapse *synapses,
                              int synapse_count);

/* Slot reserve for ns_pool_replace_records, as ns_gpu_reserve: rebuilds the
 * pool (state kept) with capacities rounded up a ladder of ratio 1 + slack and
 * spare slots for a fraction `spare` of the animals.  It stays in force for
 * later imports.  No counterpart in the reference (its slots are single
 * neurons and synapses in one bitmap). */
int ns_pool_reserve(void *pool, float slack, float spare);

typedef struct ns_pool_stats {
    int64_t animals, neuron_records, synapse_records;
    int64_t slots, free_slots;          /* pool brain slots; without a brain */
    int64_t placements, relayouts;      /* newborns placed in free slots; rebuilds */
} ns_pool_stats;
int ns_pool_info(void *pool, ns_pool_stats *info);

/* ------------------------------------------------------------------------ */
/* 3. Serving: continuous batching of genome-evaluation episodes            */
/* ------------------------------------------------------------------------ */

/* Episode tasks (config.task):
 *  0  a seeded 4-d linear system x' = A x + B u + N(0, sigma^2) observed
 *     through an 8-sensor opponent rate code; two motor neurons drive u;
 *     steps_per_episode steps.  Result: return = -sum |x|^2, final_state = x,
 *     motor_spikes = the two motors' totals.
 *  1  maze_episodes (0: 10) seeded 31 x 31 depth-first mazes in a row with
 *     the brain's state kept throughout; 9 channels (walls N/E/S/W, goal
 *     E/W/S/N of the agent, reward) and 4 motors (move N/E/S/W: most spikes,
 *     ties to the lower motor, none = stay).  A maze ends at the goal (reward
 *     1 - 0.9 steps / horizon, gym_adapter/maze_gpu.cu:46, and the reward
 *     channel on for the next step) or after maze_horizon (0: 100) steps.
 *     Result: return = sum of rewards, final_state = {mazes solved, steps,
 *     mazes played, 0}, motor_spikes = totals of motors 0 and 1.
 * One device pool of slots holds every episode: size classes
 * ceil(16 * sqrt2^k) up to max_episode_neurons, the same number of slots per
 * class (at least one) within neuron_capacity and episode_capacity, synapse
 * capacity per slot in the ratio synapse_capacity / neuron_capacity (clamped
 * to [1, 32]).  A
 * request is admitted at a launch boundary into the best-fitting free slot;
 * brain and task state stay resident there until the episode ends. */
typedef struct ns_serve_config {
    int neuron_capacity, synapse_capacity, episode_capacity;
    int steps_per_episode;      /* env steps, 16 ticks each */
    int ticks_per_step;
    int steps_per_launch;       /* env steps per device launch */
    float noise_sigma;
    float input_gain;
    int max_episode_neurons;    /* largest slot class (0: 4096) */
    int task;                   /* 0 linear system, 1 mazes */
    int maze_episodes;          /* mazes per request (0: 10) */
    int maze_horizon;           /* steps per maze (0: 100) */
} ns_serve_config;

void *ns_serve_create(const ns_serve_config *config);
void ns_serve_destroy(void *server);

/* Queue one episode (genome in ns_create encoding with the task's channels
 * and outputs: 8 and 2 for the linear system, 9 and 4 for mazes).  Returns a
 * request id >= 0.  Thread-safe with respect to the serving thread. */
int64_t ns_serve_submit(void *server, int count, int edges,
                        const float *neuron_genes, const float *synapse_genes,
                        const int *motors, uint32_t seed, uint32_t task_seed);

typedef struct ns_serve_result {
    int64_t id;
    double submit_s, admit_s, finish_s;  /* seconds since server creation */
    float episode_return;                /* sum of -|x|^2 over steps */
    float final_state[4];
    int32_t motor_spikes[2];
    int32_t status;
} ns_serve_result;

/* Move up to `capacity` finished results into `out`; returns how many. */
int ns_serve_poll(void *server, ns_serve_result *out, int capacity);

/* Block until every submitted episode has finished. */
int ns_serve_drain(void *server);

typedef struct ns_serve_stats {
    int64_t submitted, admitted, finished, launches;   /* launches: launch rounds */
    int64_t resident_peak;
    double busy_s;              /* device time of the launch rounds (HIP events) */
} ns_serve_stats;
int ns_serve_stats_get(void *server, ns_serve_stats *stats);

/* Maze `index` of a maze request seeded `seed`, as the serving loop carves
 * it: rows[31] (bit x of rows[y] = cell (x, y) open) and goal[2] = (x, y).
 * Host only; for callers that replay or display episodes. */
void ns_maze_layout(uint32_t seed, uint32_t index, uint32_t *rows, uint32_t *goal);

/* ------------------------------------------------------------------------ */
/* 4. NSCHKP01 checkpoint files (src/checkpoint_impl.cuh)                    */
/* ------------------------------------------------------------------------ */

/* checkpoint_impl.cuh:6-19 CheckpointHeader (608 B). */
typedef struct ns_checkpoint_header {
    char magic[8];                     /* "NSCHKP01" */
    uint32_t version, endian;          /* 1, 0x01020304 */
    uint64_t layout;                   /* ns_checkpoint_layout() */
    int64_t step, checkpoint_every;
    unsigned long long seed;
    int32_t initial, trace_animal;
    float food_renewal;
    uint32_t flags, neuron_count, synapse_count;
    uint64_t csv_size, behavior_size, lives_size;
    char output[512];
} ns_checkpoint_header;

/* checkpoint_impl.cuh:21-32 checkpoint_layout: FNV-1a over the world build's
 * capacities and struct sizes (32768 animals, 4194304 neurons, 8388608
 * synapses, 19 senses, 5 motors, a 128 x 128 world, ...). */
uint64_t ns_checkpoint_layout(void);
/* The layout hash of a build with `history` ring slots (9: the value of
 * ns_checkpoint_layout; 17: HISTORY 17 and the 352-byte neuron record);
 * 0 for any other history. */
uint64_t ns_checkpoint_layout_ex(int history);

/* Read a whole checkpoint file with the reference's checks, in its order
 * (main.cu:1986-2000 header, checkpoint_impl.cuh:201-366 state, 368-383
 * header checks): incomplete or corrupt data, trailing data, an
 * incompatible header or layout, history counts, animal counts against the
 * record counts, and every record (ascending slots, owner alive, bounds)
 * are rejected with the reference's messages.  The handle holds the header,
 * the records, and the world sections as one block in file order: observer
 * totals and histories, animals, food, kind, senses, motors, counters,
 * traces, then the plasticity-cohort extension when flags & 64.  NULL on
 * failure.  Both layout versions are read (ns_checkpoint_layout_ex(9) and
 * (17)); ns_checkpoint_history tells which, and the records come out of
 * ns_checkpoint_get (9 slots) or ns_checkpoint_get17 (17 slots), the other
 * call failing.  The whole file is held in host memory: at the world's
 * capacities up to 1.44 GB of neuron records (1.48 GB with 17 slots) and
 * 1.14 GB of synapse records; ns_pool_load_checkpoint holds them a second
 * time while it groups them per animal. */
void *ns_checkpoint_open(const char *path);
int ns_checkpoint_history(void *checkpoint);   /* 9 or 17; -1 on a bad handle */
int ns_checkpoint_get(void *checkpoint, ns_checkpoint_header *header,
                      const void **world, size_t *world_bytes,
                      const ns_checkpoint_neuron **neurons,
                      const ns_checkpoint_synapse **synapses);
int ns_checkpoint_get17(void *checkpoint, ns_checkpoint_header *header,
                        const void **world, size_t *world_bytes,
                        const ns_checkpoint_neuron17 **neurons,
                        const ns_checkpoint_synapse **synapses);
void ns_checkpoint_close(void *checkpoint);

/* checkpoint_impl.cuh:135-199, 385-413 checkpoint_atomic_save: gzip level 1
 * into path.tmp, fsync, rename onto path, fsync of the directory.  The
 * header's record counts are set from the arguments; the world block is laid
 * out as ns_checkpoint_open returns it; records must be in ascending slot
 * order and owned by living animals whose counts match.  Nothing is written
 * when a check fails, and a write that fails removes path.tmp.  The header's
 * layout must be ns_checkpoint_layout() (ns_checkpoint_layout_ex(17) for
 * ns_checkpoint_save17). */
int ns_checkpoint_save(const char *path, const ns_checkpoint_header *header,
                       const void *world, size_t world_bytes,
                       const ns_checkpoint_neuron *neurons, int neuron_count,
                       const ns_checkpoint_synapse *synapses, int synapse_count);
int ns_checkpoint_save17(const char *path, const ns_checkpoint_header *header,
                         const void *world, size_t world_bytes,
                         const ns_checkpoint_neuron17 *neurons, int neuron_count,
                         const ns_checkpoint_synapse *synapses, int synapse_count);

/* The same with a world pool's records (ns_pool_export_records /
 * ns_pool_import_records, or their 17-slot forms: the file's layout version
 * must be the pool's).  Loading checks the whole file before the pool
 * changes, then loads the pool; world may be NULL (the world block is not
 * copied, *world_bytes still gives its size). */
int ns_pool_save_checkpoint(void *pool, const char *path,
                            const ns_checkpoint_header *header,
                            const void *world, size_t world_bytes);
int ns_pool_load_checkpoint(void *pool, const char *path,
                            ns_checkpoint_header *header, void *world,
                            size_t world_capacity, size_t *world_bytes);

/* ------------------------------------------------------------------------ */
/* 5. The world step around the neural ticks (src/main.cu)                   */
/* ------------------------------------------------------------------------ */

/* A world of food and animals whose brains are a world-rules batch
 * (ns_gpu_create_bulk with world_rules = 1, 19 sensory channels, 5 motors):
 * update_senses (main.cu:619-657) feeds the batch's input, its motor spike
 * counts drive body_step (main.cu:1146-1249), then bites (main.cu:1251-1285),
 * world homeostasis (main.cu:1287-1300) and regrowth (main.cu:1138-1144).
 * torus = 1 wraps the width x height grid (the C2 world); 0 clamps as the
 * reference's does.  Each food kind A/B/C covers `food_density` of the cells
 * at the start (main.cu:458-462 with 0.28 / 3 there); an empty cell regrows
 * with probability `regrowth` per step.  Where the reference resolves
 * contention by atomics (one food item, the prey bitten first, the bite that
 * kills) this step takes the animals in ascending index, so it is
 * deterministic; its sines and cosines are computed in double and rounded
 * once, not the reference's fast-math intrinsics.  Reproduction and the
 * behavior observers are not part of it. */
typedef struct ns_ecology_config {
    int width, height, torus;
    float food_density, regrowth;
    uint64_t seed;
    int instant_motors;                /* main.cu instant_motors: force = spikes / 16 */
} ns_ecology_config;

/* One animal's body (main.cu:75-84 Animal, the fields this step uses). */
typedef struct ns_world_animal {
    float x, y, heading, energy, health, fullness, pain;
    float taste[4], muscle[3];
    int32_t age, alive;
} ns_world_animal;

void *ns_ecology_create(void *batch, const ns_ecology_config *config);
void ns_ecology_destroy(void *ecology);
/* One world step: the 16 neural ticks of every living animal on its senses,
 * then body, bites, homeostasis, regrowth and the next senses.  `step`
 * numbers the regrowth draws (main.cu:1142).  Synchronous. */
int ns_ecology_step(void *ecology, int64_t step);
/* Animals (batch order), food stock and kind per cell (row-major), senses
 * (animals x 19) and the last motor spike counts (animals x 5); any may be
 * NULL. */
int ns_ecology_export(void *ecology, ns_world_animal *animals, int32_t *food, uint8_t *kind,
                      float *senses, float *motors);

#ifdef __cplusplus
}
#endif

#endif /* NS_AMD_H */
