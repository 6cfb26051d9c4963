"""ctypes binding of libns_amd.so (include/ns_amd.h).

The batch class mirrors gym_adapter/gpu.py's GpuPolicyBatch use of the
ns_gpu_* ABI: create from genomes, reset with per-brain seeds and a mask,
step with encoded inputs.  The library is required: importing this module
raises if lib/libns_amd.so has not been built, and there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as ct
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get('NS_AMD_LIB', HERE / 'lib' / 'libns_amd.so'))

FORM_RESIDENT = 0
FORM_PER_TICK = 1

_f = ct.POINTER(ct.c_float)
_u8 = ct.POINTER(ct.c_uint8)
_u32 = ct.POINTER(ct.c_uint32)
_i32 = ct.POINTER(ct.c_int)
_i64 = ct.POINTER(ct.c_int64)
_vp = ct.c_void_p


class BatchInfo(ct.Structure):
    _fields_ = [('brains', ct.c_int64), ('neurons', ct.c_int64), ('synapses', ct.c_int64),
                ('resident_packs', ct.c_int64), ('fallback_brains', ct.c_int64),
                ('events', ct.c_int64), ('threads', ct.c_int), ('neurons_per_thread', ct.c_int),
                ('synapses_per_thread', ct.c_int), ('form', ct.c_int),
                ('device_bytes', ct.c_int64), ('resident_ms', ct.c_double),
                ('per_tick_ms', ct.c_double), ('resident_launches', ct.c_int64),
                ('per_tick_calls', ct.c_int64), ('resident_events', ct.c_int64),
                ('fallback_neurons', ct.c_int64), ('fallback_synapses', ct.c_int64),
                ('big_brains', ct.c_int64), ('big_neurons', ct.c_int64),
                ('min_waves', ct.c_int), ('lds_synapses', ct.c_int),
                ('slots', ct.c_int64), ('free_slots', ct.c_int64), ('slot_neurons', ct.c_int64),
                ('slot_synapses', ct.c_int64), ('placements', ct.c_int64),
                ('relayouts', ct.c_int64), ('graph_captures', ct.c_int64),
                ('graph_launches', ct.c_int64), ('staging_grows', ct.c_int64),
                ('appended_slots', ct.c_int64), ('skipped_neuron_ticks', ct.c_int64)]


class ServeConfig(ct.Structure):
    _fields_ = [('neuron_capacity', ct.c_int), ('synapse_capacity', ct.c_int),
                ('episode_capacity', ct.c_int), ('steps_per_episode', ct.c_int),
                ('ticks_per_step', ct.c_int), ('steps_per_launch', ct.c_int),
                ('noise_sigma', ct.c_float), ('input_gain', ct.c_float),
                ('max_episode_neurons', ct.c_int), ('task', ct.c_int),
                ('maze_episodes', ct.c_int), ('maze_horizon', ct.c_int)]


TASK_LINEAR, TASK_MAZE = 0, 1


class ServeResult(ct.Structure):
    _fields_ = [('id', ct.c_int64), ('submit_s', ct.c_double), ('admit_s', ct.c_double),
                ('finish_s', ct.c_double), ('episode_return', ct.c_float),
                ('final_state', ct.c_float * 4), ('motor_spikes', ct.c_int32 * 2),
                ('status', ct.c_int32)]


class ServeStats(ct.Structure):
    _fields_ = [('submitted', ct.c_int64), ('admitted', ct.c_int64), ('finished', ct.c_int64),
                ('launches', ct.c_int64), ('resident_peak', ct.c_int64), ('busy_s', ct.c_double)]


class PoolStats(ct.Structure):
    _fields_ = [('animals', ct.c_int64), ('neuron_records', ct.c_int64), ('synapse_records', ct.c_int64),
                ('slots', ct.c_int64), ('free_slots', ct.c_int64), ('placements', ct.c_int64),
                ('relayouts', ct.c_int64)]


class CheckpointHeader(ct.Structure):
    _fields_ = [('magic', ct.c_char * 8), ('version', ct.c_uint32), ('endian', ct.c_uint32),
                ('layout', ct.c_uint64), ('step', ct.c_int64), ('checkpoint_every', ct.c_int64),
                ('seed', ct.c_ulonglong), ('initial', ct.c_int32), ('trace_animal', ct.c_int32),
                ('food_renewal', ct.c_float), ('flags', ct.c_uint32), ('neuron_count', ct.c_uint32),
                ('synapse_count', ct.c_uint32), ('csv_size', ct.c_uint64), ('behavior_size', ct.c_uint64),
                ('lives_size', ct.c_uint64), ('output', ct.c_char * 512)]


class EcologyConfig(ct.Structure):
    _fields_ = [('width', ct.c_int), ('height', ct.c_int), ('torus', ct.c_int),
                ('food_density', ct.c_float), ('regrowth', ct.c_float), ('seed', ct.c_uint64),
                ('instant_motors', ct.c_int)]


WORLD_ANIMAL = np.dtype([('x', '<f4'), ('y', '<f4'), ('heading', '<f4'), ('energy', '<f4'),
                         ('health', '<f4'), ('fullness', '<f4'), ('pain', '<f4'),
                         ('taste', '<f4', 4), ('muscle', '<f4', 3), ('age', '<i4'),
                         ('alive', '<i4')])


# Exported symbols of include/ns_amd.h with their ctypes signatures.
SIGNATURES = {
    'ns_error': ([], ct.c_char_p),
    'ns_create': ([ct.c_int] * 4 + [_f, _f, _i32], _vp),
    'ns_create_ex': ([ct.c_int] * 4 + [_f, _f, _i32, ct.c_int], _vp),
    'ns_original_indices': ([_vp, _u32, _u32], None),
    'ns_destroy': ([_vp], None),
    'ns_reset': ([_vp, ct.c_uint32], None),
    'ns_step': ([_vp, _f, ct.c_int, ct.c_float, _f], ct.c_int),
    'ns_export': ([_vp, _vp, _vp, _u32, _i64], ct.c_int),
    'ns_step_static': ([_vp, _f, ct.c_int, ct.c_float, _f], ct.c_int),
    'ns_gpu_create': ([ct.POINTER(_vp), ct.c_int], _vp),
    'ns_gpu_create_bulk': ([ct.c_int, _i32, _i32, ct.c_int, ct.c_int, _f, _f, _i32, ct.c_int,
                            ct.c_int], _vp),
    'ns_gpu_destroy': ([_vp], None),
    'ns_gpu_reset': ([_vp, _u8, _u32], ct.c_int),
    'ns_gpu_step': ([_vp, _f, _u8, ct.c_int, ct.c_float, _f, _i32], ct.c_int),
    'ns_gpu_step_device': ([_vp, _vp, _vp, ct.c_int, ct.c_float, _vp, _vp, _vp], ct.c_int),
    'ns_gpu_set_form': ([_vp, ct.c_int], ct.c_int),
    'ns_gpu_set_tiles': ([_vp, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_int], ct.c_int),
    'ns_gpu_autotune': ([_vp, ct.c_int, _i32, _i32, _i32, _i32, _i32], ct.c_int),
    'ns_gpu_export': ([_vp, _vp, _vp, _u8, _i64], ct.c_int),
    'ns_gpu_import': ([_vp, _vp, _vp, _u8, _i64], ct.c_int),
    'ns_gpu_export_spike_bits': ([_vp, _u32], ct.c_int),
    'ns_gpu_import_spike_bits': ([_vp, _u32], ct.c_int),
    'ns_gpu_info': ([_vp, ct.POINTER(BatchInfo)], ct.c_int),
    'ns_gpu_set_timing': ([_vp, ct.c_int], ct.c_int),
    'ns_gpu_reserve': ([_vp, ct.c_float, ct.c_float], ct.c_int),
    'ns_gpu_set_graph': ([_vp, ct.c_int], ct.c_int),
    'ns_gpu_replace': ([_vp, ct.c_int, _i32, ct.POINTER(_vp), _u32, _vp], ct.c_int),
    'ns_device_math_digest': ([ct.c_int, ct.c_uint64, ct.c_uint64, ct.c_int64,
                               ct.POINTER(ct.c_uint64)], ct.c_int),
    'ns_pool_create': ([ct.c_int, ct.c_int, ct.c_int], _vp),
    'ns_pool_create_ex': ([ct.c_int, ct.c_int, ct.c_int, ct.c_int], _vp),
    'ns_pool_history': ([_vp], ct.c_int),
    'ns_pool_import_records17': ([_vp, _vp, ct.c_int, _vp, ct.c_int], ct.c_int),
    'ns_pool_export_records17': ([_vp, _vp, ct.c_int, _vp, ct.c_int, _i32, _i32], ct.c_int),
    'ns_pool_replace_records17': ([_vp, ct.c_int, _i32, _vp, ct.c_int, _vp, ct.c_int], ct.c_int),
    'ns_pool_destroy': ([_vp], None),
    'ns_pool_import_records': ([_vp, _vp, ct.c_int, _vp, ct.c_int], ct.c_int),
    'ns_pool_export_records': ([_vp, _vp, ct.c_int, _vp, ct.c_int, _i32, _i32], ct.c_int),
    'ns_world_neural_step': ([_vp, ct.c_int64, _f, _i32, _u8], ct.c_int),
    'ns_world_homeostasis': ([_vp, _i32, _u8], ct.c_int),
    'ns_world_neural_step_device': ([_vp, ct.c_int64, _vp, _vp, _u8, _vp], ct.c_int),
    'ns_pool_replace_records': ([_vp, ct.c_int, _i32, _vp, ct.c_int, _vp, ct.c_int], ct.c_int),
    'ns_pool_reserve': ([_vp, ct.c_float, ct.c_float], ct.c_int),
    'ns_pool_info': ([_vp, ct.POINTER(PoolStats)], ct.c_int),
    'ns_serve_create': ([ct.POINTER(ServeConfig)], _vp),
    'ns_serve_destroy': ([_vp], None),
    'ns_serve_submit': ([_vp, ct.c_int, ct.c_int, _f, _f, _i32, ct.c_uint32, ct.c_uint32],
                        ct.c_int64),
    'ns_serve_poll': ([_vp, ct.POINTER(ServeResult), ct.c_int], ct.c_int),
    'ns_serve_drain': ([_vp], ct.c_int),
    'ns_serve_stats_get': ([_vp, ct.POINTER(ServeStats)], ct.c_int),
    'ns_checkpoint_layout': ([], ct.c_uint64),
    'ns_checkpoint_layout_ex': ([ct.c_int], ct.c_uint64),
    'ns_checkpoint_open': ([ct.c_char_p], _vp),
    'ns_checkpoint_history': ([_vp], ct.c_int),
    'ns_checkpoint_get': ([_vp, ct.POINTER(CheckpointHeader), ct.POINTER(_vp), ct.POINTER(ct.c_size_t),
                           ct.POINTER(_vp), ct.POINTER(_vp)], ct.c_int),
    'ns_checkpoint_get17': ([_vp, ct.POINTER(CheckpointHeader), ct.POINTER(_vp), ct.POINTER(ct.c_size_t),
                             ct.POINTER(_vp), ct.POINTER(_vp)], ct.c_int),
    'ns_checkpoint_close': ([_vp], None),
    'ns_checkpoint_save': ([ct.c_char_p, ct.POINTER(CheckpointHeader), _vp, ct.c_size_t, _vp, ct.c_int, _vp,
                            ct.c_int], ct.c_int),
    'ns_checkpoint_save17': ([ct.c_char_p, ct.POINTER(CheckpointHeader), _vp, ct.c_size_t, _vp, ct.c_int, _vp,
                              ct.c_int], ct.c_int),
    'ns_pool_save_checkpoint': ([_vp, ct.c_char_p, ct.POINTER(CheckpointHeader), _vp, ct.c_size_t], ct.c_int),
    'ns_pool_load_checkpoint': ([_vp, ct.c_char_p, ct.POINTER(CheckpointHeader), _vp, ct.c_size_t,
                                 ct.POINTER(ct.c_size_t)], ct.c_int),
    'ns_ecology_create': ([_vp, ct.POINTER(EcologyConfig)], _vp),
    'ns_ecology_destroy': ([_vp], None),
    'ns_ecology_step': ([_vp, ct.c_int64], ct.c_int),
    'ns_ecology_export': ([_vp, _vp, _i32, _u8, _f, _f], ct.c_int),
    'ns_maze_layout': ([ct.c_uint32, ct.c_uint32, ct.POINTER(ct.c_uint32), ct.POINTER(ct.c_uint32)],
                       None),
}

_lib = None


def lib():
    """Load libns_amd.so; raises if it is missing (no CPU fallback exists)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f'{LIB_PATH} is not built: run make -C neural-selection_amd '
                               '(or __graft_entry__.build())')
        handle = ct.CDLL(str(LIB_PATH))
        for name, (args, res) in SIGNATURES.items():
            if not hasattr(handle, name):
                continue
            fn = getattr(handle, name)
            fn.argtypes = args
            fn.restype = res
        _lib = handle
    return _lib


def _ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)


def error():
    return lib().ns_error().decode()


def device_math_digest(fn, lo, hi, chunk):
    n = (hi - lo + chunk - 1) // chunk
    out = np.zeros(n, np.uint64)
    rc = lib().ns_device_math_digest(fn, lo, hi, chunk, out.ctypes.data_as(ct.POINTER(ct.c_uint64)))
    if rc:
        raise RuntimeError(error())
    return out


class GenomeHandles:
    """ns_create handles for a list of gym-encoded genomes (neuron rows,
    synapse rows, motors), kept for ns_gpu_replace."""

    def __init__(self, genomes, channels, max_delay=8):
        L = lib()
        self.lib = L
        self.ptrs = []
        self.counts, self.edges = [], []
        self.max_delay = max_delay
        try:
            for nr, sr, mo in genomes:
                nr = np.ascontiguousarray(nr, np.float32)
                sr = np.ascontiguousarray(np.asarray(sr, np.float32).reshape(-1, 11))
                mo = np.ascontiguousarray(mo, np.int32)
                h = L.ns_create_ex(len(nr), len(sr), channels, len(mo), _ptr(nr, _f), _ptr(sr, _f),
                                   _ptr(mo, _i32), max_delay)
                if not h:
                    raise RuntimeError(error())
                self.ptrs.append(h)
                self.counts.append(len(nr))
                self.edges.append(len(sr))
        except Exception:
            self.close()
            raise

    def close(self):
        for h in self.ptrs:
            self.lib.ns_destroy(h)
        self.ptrs = []


class GpuBatch:
    """Device batch of brains (ns_gpu_*), gym-encoded genomes.

    genomes: list of (neuron rows [n,10], synapse rows [m,11], motors [outputs]).
    """

    def __init__(self, genomes=None, channels=None, max_delay=8, world_rules=False,
                 workload=None, handles=False):
        self.ptr = None
        L = lib()
        self.lib = L
        if workload is not None:
            counts = workload.neuron_counts
            edges = workload.edge_counts
            neurons, synapses, motors = workload.neurons, workload.synapses, workload.motors
            channels, outputs, max_delay = workload.channels, workload.outputs, workload.max_delay
        else:
            counts = np.array([len(g[0]) for g in genomes], np.int32)
            edges = np.array([len(g[1]) for g in genomes], np.int32)
            outputs = len(genomes[0][2])
            neurons = np.concatenate([np.asarray(g[0], np.float32) for g in genomes])
            synapses = (np.concatenate([np.asarray(g[1], np.float32).reshape(-1, 11)
                                        for g in genomes]) if edges.sum() else
                        np.zeros((0, 11), np.float32))
            motors = np.concatenate([np.asarray(g[2], np.int32) for g in genomes])
        self.count = len(counts)
        self.channels = channels
        self.outputs = outputs
        self.max_delay = max_delay
        self.neuron_counts = np.ascontiguousarray(counts, np.int32)
        self.edge_counts = np.ascontiguousarray(edges, np.int32)
        if handles:
            # The reference's two-step path: ns_create per genome, then ns_gpu_create.
            hs = []
            n0 = s0 = 0
            try:
                for i in range(self.count):
                    n, s = int(counts[i]), int(edges[i])
                    ng = np.ascontiguousarray(neurons[n0:n0 + n], np.float32)
                    sg = np.ascontiguousarray(synapses[s0:s0 + s], np.float32)
                    mo = np.ascontiguousarray(motors[i * outputs:(i + 1) * outputs], np.int32)
                    h = L.ns_create_ex(n, s, channels, outputs, _ptr(ng, _f), _ptr(sg, _f),
                                       _ptr(mo, _i32), max_delay)
                    if not h:
                        raise RuntimeError(error())
                    hs.append(h)
                    n0 += n
                    s0 += s
                arr = (_vp * self.count)(*hs)
                self.ptr = L.ns_gpu_create(arr, self.count)
            finally:
                for h in hs:
                    L.ns_destroy(h)
        else:
            neurons = np.ascontiguousarray(neurons, np.float32)
            synapses = np.ascontiguousarray(synapses, np.float32)
            motors = np.ascontiguousarray(motors, np.int32)
            self.ptr = L.ns_gpu_create_bulk(self.count, _ptr(self.neuron_counts, _i32),
                                            _ptr(self.edge_counts, _i32), channels, outputs,
                                            _ptr(neurons, _f), _ptr(synapses, _f),
                                            _ptr(motors, _i32), max_delay, int(world_rules))
        if not self.ptr:
            raise RuntimeError(error())
        self.output = np.zeros((self.count, outputs), np.float32)
        self.status = np.zeros(self.count, np.int32)

    def _check(self, rc):
        if rc:
            raise RuntimeError(error())

    def reset(self, seeds, mask=None):
        seeds = np.ascontiguousarray(seeds, np.uint32)
        mask = (np.ones(self.count, np.uint8) if mask is None
                else np.ascontiguousarray(mask, np.uint8))
        self._check(self.lib.ns_gpu_reset(self.ptr, _ptr(mask, _u8), _ptr(seeds, _u32)))

    def step(self, inputs, ticks=16, gain=230.0, mask=None):
        inputs = np.ascontiguousarray(inputs, np.float32)
        mask = (np.ones(self.count, np.uint8) if mask is None
                else np.ascontiguousarray(mask, np.uint8))
        self._check(self.lib.ns_gpu_step(self.ptr, _ptr(inputs, _f), _ptr(mask, _u8), ticks,
                                         gain, _ptr(self.output, _f), _ptr(self.status, _i32)))
        return self.output.copy(), self.status.copy()

    def step_device(self, input_ptr, ticks, gain, output_ptr, stream=None, mask_ptr=None):
        self._check(self.lib.ns_gpu_step_device(self.ptr, input_ptr, mask_ptr, ticks, gain,
                                                output_ptr, None, stream))

    def set_form(self, form):
        self._check(self.lib.ns_gpu_set_form(self.ptr, form))

    def set_tiles(self, threads, kn, ks, min_waves=0, lds_synapses=0):
        self._check(self.lib.ns_gpu_set_tiles(self.ptr, threads, kn, ks, min_waves, lds_synapses))

    def autotune(self, ticks=16):
        t, kn, ks, mw, ld = ct.c_int(), ct.c_int(), ct.c_int(), ct.c_int(), ct.c_int()
        self._check(self.lib.ns_gpu_autotune(self.ptr, ticks, ct.byref(t), ct.byref(kn),
                                             ct.byref(ks), ct.byref(mw), ct.byref(ld)))
        return t.value, kn.value, ks.value, mw.value, ld.value

    def set_timing(self, on=True):
        self._check(self.lib.ns_gpu_set_timing(self.ptr, int(on)))

    def set_graph(self, on=True):
        """Run each step as one replayed hipGraph launch (ns_gpu_set_graph)."""
        self._check(self.lib.ns_gpu_set_graph(self.ptr, int(on)))

    def reserve(self, slack=0.19, spare=0.03):
        """Slot reserve for replacements (ns_gpu_reserve); state is kept."""
        self._check(self.lib.ns_gpu_reserve(self.ptr, slack, spare))

    def replace(self, indices, handles, which, seeds, stream=None):
        """Brains `indices` take genomes handles.ptrs[which[j]] with fresh state
        seeded by seeds[j] (ns_gpu_replace), stream-ordered on `stream`."""
        idx = np.ascontiguousarray(indices, np.int32)
        which = list(which)
        arr = (_vp * len(idx))(*[handles.ptrs[w] for w in which])
        sd = np.ascontiguousarray(seeds, np.uint32)
        self._check(self.lib.ns_gpu_replace(self.ptr, len(idx), _ptr(idx, _i32), arr, _ptr(sd, _u32),
                                            stream))
        for i, w in zip(idx, which):
            self.neuron_counts[i] = handles.counts[w]
            self.edge_counts[i] = handles.edges[w]
        if len(idx):
            self.max_delay = max(self.max_delay, handles.max_delay)

    def info(self):
        i = BatchInfo()
        self._check(self.lib.ns_gpu_info(self.ptr, ct.byref(i)))
        return {k: getattr(i, k) for k, _ in BatchInfo._fields_}

    def export(self):
        """(neuron states, synapse states, history9, ticks) in genome order;
        history9 is None when the delay bound exceeds 8 (nine slots would not
        hold the ring: read spike_bits() instead)."""
        from_dtypes = _state_dtypes()
        ns = np.zeros(int(self.neuron_counts.sum()), from_dtypes[0])
        ss = np.zeros(int(self.edge_counts.sum()), from_dtypes[1])
        h9 = (np.zeros((int(self.neuron_counts.sum()), 9), np.uint8) if self.max_delay <= 8
              else None)
        ticks = np.zeros(self.count, np.int64)
        self._check(self.lib.ns_gpu_export(self.ptr, ns.ctypes.data, ss.ctypes.data,
                                           _ptr(h9, _u8), _ptr(ticks, _i64)))
        return ns, ss, h9, ticks

    def import_state(self, ns=None, ss=None, h9=None, ticks=None):
        self._check(self.lib.ns_gpu_import(
            self.ptr, None if ns is None else ns.ctypes.data,
            None if ss is None else ss.ctypes.data, _ptr(h9, _u8), _ptr(ticks, _i64)))

    def spike_bits(self):
        bits = np.zeros(int(self.neuron_counts.sum()), np.uint32)
        self._check(self.lib.ns_gpu_export_spike_bits(self.ptr, _ptr(bits, _u32)))
        return bits

    def import_spike_bits(self, bits):
        bits = np.ascontiguousarray(bits, np.uint32)
        self._check(self.lib.ns_gpu_import_spike_bits(self.ptr, _ptr(bits, _u32)))

    def close(self):
        if self.ptr:
            self.lib.ns_gpu_destroy(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        self.close()


class Ecology:
    """The world step around a world-rules batch (ns_ecology_*)."""

    def __init__(self, batch, width=256, height=256, torus=True, food_density=0.05,
                 regrowth=1e-3, seed=1, instant_motors=False):
        self.lib = lib()
        self.batch = batch
        self.count = batch.count
        self.cells = width * height
        cfg = EcologyConfig(width, height, int(torus), food_density, regrowth, seed,
                            int(instant_motors))
        self.ptr = self.lib.ns_ecology_create(batch.ptr, ct.byref(cfg))
        if not self.ptr:
            raise RuntimeError(error())

    def step(self, step):
        if self.lib.ns_ecology_step(self.ptr, step):
            raise RuntimeError(error())

    def export(self):
        an = np.zeros(self.count, WORLD_ANIMAL)
        food = np.zeros(self.cells, np.int32)
        kind = np.zeros(self.cells, np.uint8)
        senses = np.zeros((self.count, 19), np.float32)
        motors = np.zeros((self.count, 5), np.float32)
        if self.lib.ns_ecology_export(self.ptr, an.ctypes.data, _ptr(food, _i32), _ptr(kind, _u8),
                                      _ptr(senses, _f), _ptr(motors, _f)):
            raise RuntimeError(error())
        return an, food, kind, senses, motors

    def close(self):
        if self.ptr:
            self.lib.ns_ecology_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()


def _state_dtypes():
    neuron = np.dtype({
        'names': ['f', 'dend', 'cell', 'total_spikes', 'spike', 'receptor_ns', 'receptor_rng'],
        'formats': [('<f4', 6), ('<f4', 4), ('<f4', 16), '<i4', 'u1', '<f4', '<u4'],
        'offsets': [0, 24, 40, 104, 108, 112, 116], 'itemsize': 120})
    synapse = np.dtype({
        'names': ['f', 'last_arrival_tick', 'ready_mask', 'rng_state', 'arrived', 'slow_resource'],
        'formats': [('<f4', 11), '<i8', '<u8', '<u4', 'u1', '<f4'],
        'offsets': [0, 48, 56, 64, 68, 72], 'itemsize': 80})
    return neuron, synapse
