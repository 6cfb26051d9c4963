"""ctypes bindings of the CPU oracle (liboracle.so) and of the reference's own
CPU path host-compiled into oracle/_ref/libns_ref.so — TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as ct
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_LIB = HERE / 'liboracle.so'
REF_LIB = HERE / '_ref' / 'libns_ref.so'

# Record dtypes matching include/ns_amd.h (and the reference structs).
NEURON_GENE = np.dtype({
    'names': ['f', 'type', 'active', 'modulator', 'inhibitory', 'lp_scale', 'prinz_g',
              'prinz_ambient', 'sensory_channel', 'receptor', 'target_affinity'],
    'formats': [('<f4', 11), 'u1', 'u1', 'u1', 'u1', ('<f4', 4), ('<f4', 8), '<f4', 'i1',
                ('<f4', 4), '<f4'],
    'offsets': [0, 44, 45, 46, 47, 48, 64, 96, 100, 104, 120], 'itemsize': 124})
NEURON_STATE = np.dtype({
    'names': ['f', 'dend', 'cell', 'total_spikes', 'spike', 'receptor_ns', 'receptor_rng'],
    'formats': [('<f4', 6), ('<f4', 4), ('<f4', 16), '<i4', 'u1', '<f4', '<u4'],
    'offsets': [0, 24, 40, 104, 108, 112, 116], 'itemsize': 120})
SYNAPSE_GENE = np.dtype({
    'names': ['initial_weight_ns', 'utilization', 'recovery_ms', 'facilitation_ms', 'pre',
              'post', 'delay', 'active', 'plasticity', 'branch', 'release_sites',
              'mod_receptors', 'slow_recovery_ms', 'slow_depletion'],
    'formats': ['<f4', '<f4', '<f4', '<f4', '<i4', '<i4', 'u1', 'u1', 'u1', 'u1', 'u1', 'u1',
                '<f4', '<f4'],
    'offsets': [0, 4, 8, 12, 16, 20, 24, 25, 26, 27, 28, 29, 32, 36], 'itemsize': 40})
SYNAPSE_STATE = np.dtype({
    'names': ['f', 'last_arrival_tick', 'ready_mask', 'rng_state', 'arrived', 'slow_resource'],
    'formats': [('<f4', 11), '<i8', '<u8', '<u4', 'u1', '<f4'],
    'offsets': [0, 48, 56, 64, 68, 72], 'itemsize': 80})

# Byte masks of the padding inside each state record (excluded from compares).
def _pad_mask(dtype):
    m = np.zeros(dtype.itemsize, bool)
    for name in dtype.names:
        fmt, off = dtype.fields[name][:2]
        m[off:off + fmt.itemsize] = True
    return m


NEURON_GENE_BYTES = _pad_mask(NEURON_GENE)
NEURON_STATE_BYTES = _pad_mask(NEURON_STATE)
SYNAPSE_STATE_BYTES = _pad_mask(SYNAPSE_STATE)

_f = ct.POINTER(ct.c_float)
_u8 = ct.POINTER(ct.c_uint8)
_u32 = ct.POINTER(ct.c_uint32)
_i32 = ct.POINTER(ct.c_int)
_i64 = ct.POINTER(ct.c_int64)
_u64 = ct.POINTER(ct.c_uint64)
_vp = ct.c_void_p


def ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)


_oracle = None


def oracle_lib():
    global _oracle
    if _oracle is None:
        if not ORACLE_LIB.exists():
            raise RuntimeError(f'{ORACLE_LIB} is not built (make -C oracle)')
        lib = ct.CDLL(str(ORACLE_LIB))
        lib.ora_error.restype = ct.c_char_p
        lib.ora_create.argtypes = [ct.c_int] * 4 + [_f, _f, _i32, ct.c_int]
        lib.ora_create.restype = _vp
        lib.ora_create_records.argtypes = [ct.c_int, ct.c_int, ct.c_int, _vp, _vp, _i32, _u8,
                                           ct.c_int, _i32, ct.c_int, ct.c_int]
        lib.ora_create_records.restype = _vp
        lib.ora_create_world.argtypes = [ct.c_int, ct.c_int, _vp, _vp, _i32, ct.c_int]
        lib.ora_create_world.restype = _vp
        lib.ora_destroy.argtypes = [_vp]
        lib.ora_genes.argtypes = [_vp, _vp, _vp]
        lib.ora_ecology_create.argtypes = [ct.c_int] * 4 + [ct.c_float, ct.c_float, ct.c_uint64,
                                                             ct.c_int, _i32]
        lib.ora_ecology_create.restype = _vp
        lib.ora_ecology_destroy.argtypes = [_vp]
        lib.ora_ecology_step.argtypes = [_vp, ct.POINTER(_vp), ct.c_int64]
        lib.ora_ecology_export.argtypes = [_vp, _vp, _i32, _u8, _f, _f]
        for name in ('ora_world_sin', 'ora_world_cos'):
            getattr(lib, name).argtypes = [ct.c_float]
            getattr(lib, name).restype = ct.c_float
        lib.ora_original_indices.argtypes = [_vp, _u32, _u32]
        lib.ora_reset.argtypes = [_vp, ct.c_uint32]
        lib.ora_step.argtypes = [_vp, _f, ct.c_int, ct.c_float, _f, _u8]
        lib.ora_step.restype = ct.c_int
        lib.ora_set_tick.argtypes = [_vp, ct.c_int64]
        lib.ora_set_frozen.argtypes = [_vp, ct.c_int]
        lib.ora_world_homeostasis.argtypes = [_vp, ct.c_int]
        lib.ora_export.argtypes = [_vp, _vp, _vp, _u8, _u8, _i64, _u64]
        lib.ora_import.argtypes = [_vp, _vp, _vp, _u8, ct.c_int64]
        lib.ora_ring.argtypes = [_vp]
        lib.ora_ring.restype = ct.c_int
        lib.ora_batch_step.argtypes = [ct.POINTER(_vp), ct.c_int, _f, ct.c_int, _u8, ct.c_int,
                                       ct.c_float, _f, ct.c_int, _i32, ct.c_int]
        lib.ora_batch_step.restype = ct.c_int
        lib.ora_linear_episode.argtypes = [_vp, ct.c_uint32, ct.c_int, ct.c_float, ct.c_int, ct.c_float,
                                           _f, _f, _i32]
        lib.ora_linear_episode.restype = ct.c_int
        lib.ora_maze_request.argtypes = [_vp, ct.c_uint32, ct.c_int, ct.c_int, ct.c_int, ct.c_float,
                                         _f, _f, _i32]
        lib.ora_maze_request.restype = ct.c_int
        lib.ora_maze_layout.argtypes = [ct.c_uint32, ct.c_uint32, _u32, _u32]
        lib.ora_world_layout.argtypes = [_u64]
        lib.ora_checkpoint_neuron_layout.argtypes = [_u64]
        for name in ('ora_fn_expf', 'ora_fn_logf', 'ora_fn_branch_io'):
            getattr(lib, name).argtypes = [ct.c_float]
            getattr(lib, name).restype = ct.c_float
        lib.ora_branch_io_nonzero.argtypes = [ct.c_uint32, ct.c_uint32, ct.c_uint32]
        lib.ora_branch_io_nonzero.restype = ct.c_uint64
        _fn_signatures(lib, 'ora_fn_')
        _oracle = lib
    return _oracle


def _fn_signatures(lib, p):
    getattr(lib, p + 'profile').argtypes = [ct.c_int, _vp]
    getattr(lib, p + 'initial_neuron').argtypes = [_vp, ct.c_uint32, _vp]
    getattr(lib, p + 'initial_synapse').argtypes = [_vp, ct.c_uint32, _vp]
    getattr(lib, p + 'receptor_current').argtypes = [_vp, _vp, ct.c_float]
    getattr(lib, p + 'receptor_current').restype = ct.c_float
    getattr(lib, p + 'neuron_advance').argtypes = [_vp, _vp, ct.c_float, _f, ct.c_float]
    getattr(lib, p + 'synapse_release').argtypes = [_vp, _vp, ct.c_int64]
    getattr(lib, p + 'synapse_release').restype = ct.c_float
    getattr(lib, p + 'plasticity').argtypes = [_vp, _vp, ct.c_int, ct.c_int, ct.c_int,
                                               ct.c_float, ct.c_int]
    getattr(lib, p + 'homeostasis').argtypes = [_vp, _vp, ct.c_float, ct.c_float]


_ref = None


def ref_available():
    return REF_LIB.exists()


def ref_lib():
    """The reference's own CPU path (gym_adapter/native.cpp + physiology.cuh)."""
    global _ref
    if _ref is None:
        if not REF_LIB.exists():
            raise RuntimeError(f'{REF_LIB} is not built (make -C oracle ref)')
        lib = ct.CDLL(str(REF_LIB))
        lib.ns_error.restype = ct.c_char_p
        lib.ns_create.argtypes = [ct.c_int] * 4 + [_f, _f, _i32]
        lib.ns_create.restype = _vp
        lib.ns_destroy.argtypes = [_vp]
        lib.ns_original_indices.argtypes = [_vp, _u32, _u32]
        lib.ns_reset.argtypes = [_vp, ct.c_uint32]
        lib.ns_step.argtypes = [_vp, _f, ct.c_int, ct.c_float, _f]
        lib.ns_step.restype = ct.c_int
        lib.ref_brain_export.argtypes = [_vp, _vp, _vp, _u8, _i64]
        lib.ref_brain_genes.argtypes = [_vp, _vp, _vp]
        _fn_signatures(lib, 'ref_fn_')
        lib.ref_fn_plasticity_quiet.argtypes = [_vp, _vp, ct.c_int]
        lib.ref_fn_dendritic_transfer.argtypes = [ct.c_float]
        lib.ref_fn_dendritic_transfer.restype = ct.c_float
        _ref = lib
    return _ref


class OracleBrain:
    """One brain in the oracle (gym encoding), mirroring native.cpp's handle."""

    def __init__(self, neurons, synapses, motors, channels, max_delay=8):
        lib = oracle_lib()
        neurons = np.ascontiguousarray(neurons, np.float32)
        synapses = np.ascontiguousarray(synapses, np.float32)
        motors = np.ascontiguousarray(motors, np.int32)
        self.count, self.edges, self.outputs = len(neurons), len(synapses), len(motors)
        self.channels = channels
        self.ptr = lib.ora_create(self.count, self.edges, channels, self.outputs,
                                  ptr(neurons, _f), ptr(synapses, _f), ptr(motors, _i32), max_delay)
        if not self.ptr:
            raise RuntimeError(lib.ora_error().decode())
        self.lib = lib

    def reset(self, seed):
        self.lib.ora_reset(self.ptr, int(seed) & 0xffffffff)

    def step(self, inputs, ticks, gain, raster=False):
        out = np.zeros(self.outputs, np.float32)
        r = np.zeros((ticks, self.count), np.uint8) if raster else None
        rc = self.lib.ora_step(self.ptr, ptr(np.ascontiguousarray(inputs, np.float32), _f), ticks,
                               gain, ptr(out, _f), ptr(r, _u8))
        return (out, rc, r) if raster else (out, rc)

    def export(self):
        ns = np.zeros(self.count, NEURON_STATE)
        ss = np.zeros(self.edges, SYNAPSE_STATE)
        h9 = np.zeros((self.count, 9), np.uint8)
        tick = np.zeros(1, np.int64)
        self.lib.ora_export(self.ptr, ns.ctypes.data, ss.ctypes.data, ptr(h9, _u8), None,
                            ptr(tick, _i64), None)
        return ns, ss, h9, int(tick[0])

    def genes(self):
        ng = np.zeros(self.count, NEURON_GENE)
        sg = np.zeros(self.edges, SYNAPSE_GENE)
        self.lib.ora_genes(self.ptr, ng.ctypes.data, sg.ctypes.data)
        return ng, sg

    def close(self):
        if self.ptr:
            self.lib.ora_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()


class RefBrain:
    """One brain in the reference's own CPU path (oracle/_ref)."""

    def __init__(self, neurons, synapses, motors, channels):
        lib = ref_lib()
        neurons = np.ascontiguousarray(neurons, np.float32)
        synapses = np.ascontiguousarray(synapses, np.float32)
        motors = np.ascontiguousarray(motors, np.int32)
        self.count, self.edges, self.outputs = len(neurons), len(synapses), len(motors)
        self.ptr = lib.ns_create(self.count, self.edges, channels, self.outputs,
                                 ptr(neurons, _f), ptr(synapses, _f), ptr(motors, _i32))
        if not self.ptr:
            raise RuntimeError(lib.ns_error().decode())
        self.lib = lib

    def reset(self, seed):
        self.lib.ns_reset(self.ptr, int(seed) & 0xffffffff)

    def step(self, inputs, ticks, gain):
        out = np.zeros(self.outputs, np.float32)
        rc = self.lib.ns_step(self.ptr, ptr(np.ascontiguousarray(inputs, np.float32), _f), ticks,
                              gain, ptr(out, _f))
        return out, rc

    def export(self):
        ns = np.zeros(self.count, NEURON_STATE)
        ss = np.zeros(self.edges, SYNAPSE_STATE)
        h9 = np.zeros((self.count, 9), np.uint8)
        tick = np.zeros(1, np.int64)
        self.lib.ref_brain_export(self.ptr, ns.ctypes.data, ss.ctypes.data, ptr(h9, _u8),
                                  ptr(tick, _i64))
        return ns, ss, h9, int(tick[0])

    def close(self):
        if self.ptr:
            self.lib.ns_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()


WORLD_ANIMAL = np.dtype([('x', '<f4'), ('y', '<f4'), ('heading', '<f4'), ('energy', '<f4'),
                         ('health', '<f4'), ('fullness', '<f4'), ('pain', '<f4'),
                         ('taste', '<f4', 4), ('muscle', '<f4', 3), ('age', '<i4'),
                         ('alive', '<i4')])


def world_brain(neurons, synapses, motors, seed, max_delay=8, channels=19):
    """An oracle brain under the world rules (main.cu:710-746) from a gym
    genome: the gene keeps the row's sensory channel, motor j sits on port
    19 + j; reset with `seed` as the gym path seeds."""
    lib = oracle_lib()
    gym = OracleBrain(neurons, synapses, motors, channels, max_delay=max_delay)
    ng, sg = gym.genes()
    gym.close()
    ng['sensory_channel'] = np.asarray(neurons, np.float32)[:, 1].astype(np.int8)
    ports = np.full(len(ng), -1, np.int32)
    for j, m in enumerate(motors):
        ports[m] = 19 + j
    h = lib.ora_create_world(len(ng), len(sg), ng.ctypes.data, sg.ctypes.data, ptr(ports, _i32),
                             max_delay)
    lib.ora_reset(h, int(seed) & 0xffffffff)
    return h


class OracleEcology:
    """The oracle's restatement of the world step, with world-rule brains."""

    def __init__(self, w, width, height, torus, food_density, regrowth, seed, instant=False):
        lib = oracle_lib()
        self.lib = lib
        seeds = w.seeds()
        self.brains = [world_brain(*w.brain(i), seeds[i], max(8, w.max_delay)) for i in range(w.count)]
        active = np.ascontiguousarray([int((w.brain(i)[0][:, 2] > 0).sum()) for i in range(w.count)],
                                      np.int32)
        self.count, self.cells = w.count, width * height
        self.ptr = lib.ora_ecology_create(w.count, width, height, int(torus), food_density, regrowth,
                                          seed, int(instant), ptr(active, _i32))
        self.arr = (_vp * w.count)(*self.brains)

    def step(self, step):
        self.lib.ora_ecology_step(self.ptr, self.arr, step)

    def export(self):
        an = np.zeros(self.count, WORLD_ANIMAL)
        food = np.zeros(self.cells, np.int32)
        kind = np.zeros(self.cells, np.uint8)
        senses = np.zeros((self.count, 19), np.float32)
        motors = np.zeros((self.count, 5), np.float32)
        self.lib.ora_ecology_export(self.ptr, an.ctypes.data, ptr(food, _i32), ptr(kind, _u8),
                                    ptr(senses, _f), ptr(motors, _f))
        return an, food, kind, senses, motors

    def close(self):
        if self.ptr:
            self.lib.ora_ecology_destroy(self.ptr)
            for h in self.brains:
                self.lib.ora_destroy(h)
            self.ptr = None

    def __del__(self):
        self.close()


def cat_records(parts, dtype):
    """Concatenate record arrays keeping the padded C layout (numpy's own
    concatenate may drop the padding of structured dtypes)."""
    out = np.zeros(sum(len(p) for p in parts), dtype)
    raw = out.view(np.uint8).reshape(len(out), dtype.itemsize)
    i = 0
    for p in parts:
        raw[i:i + len(p)] = np.ascontiguousarray(p).view(np.uint8).reshape(len(p), dtype.itemsize)
        i += len(p)
    return out


def state_bytes(arr, mask):
    """Record bytes with padding bytes zeroed, for bit-exact comparison."""
    b = np.ascontiguousarray(arr).view(np.uint8).reshape(len(arr), len(mask)).copy()
    b[:, ~mask] = 0
    return b
