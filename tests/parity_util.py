"""Shared helpers for the parity tests: run a workload through the device
batch and through the CPU oracle on the same seeded inputs and compare."""
import ctypes as ct

import numpy as np

import ns_amd
import oracle as O


def oracle_brains(w):
    brains = []
    seeds = w.seeds()
    for i in range(w.count):
        n, s, m = w.brain(i)
        b = O.OracleBrain(n, s, m, w.channels, max_delay=max(8, w.max_delay))
        b.reset(seeds[i])
        brains.append(b)
    return brains


def oracle_step_all(brains, x, ticks, gain, mask=None):
    outs, rcs = [], []
    for i, b in enumerate(brains):
        if mask is not None and not mask[i]:
            outs.append(np.zeros(b.outputs, np.float32))
            rcs.append(0)
            continue
        o, rc = b.step(x[i], ticks, gain)
        outs.append(o)
        rcs.append(rc)
    return np.stack(outs), np.array(rcs)


def oracle_export_all(brains):
    parts = [b.export() for b in brains]
    ns = O.cat_records([p[0] for p in parts], O.NEURON_STATE)
    ss = O.cat_records([p[1] for p in parts], O.SYNAPSE_STATE) if sum(len(p[1]) for p in parts) else np.zeros(0, O.SYNAPSE_STATE)
    h9 = np.concatenate([p[2] for p in parts])
    ticks = np.array([p[3] for p in parts], np.int64)
    return ns, ss, h9, ticks


def assert_state_equal(dev, ora, where=''):
    dns, dss, dh9, dticks = dev
    ons, oss, oh9, oticks = ora
    assert np.array_equal(dticks, oticks), f'{where}: ticks differ'
    a = O.state_bytes(dns.view(O.NEURON_STATE), O.NEURON_STATE_BYTES)
    b = O.state_bytes(ons, O.NEURON_STATE_BYTES)
    if not np.array_equal(a, b):
        bad = np.flatnonzero((a != b).any(axis=1))
        raise AssertionError(f'{where}: {len(bad)} neuron states differ, first {bad[:8]}')
    if len(oss):
        a = O.state_bytes(dss.view(O.SYNAPSE_STATE), O.SYNAPSE_STATE_BYTES)
        b = O.state_bytes(oss, O.SYNAPSE_STATE_BYTES)
        if not np.array_equal(a, b):
            bad = np.flatnonzero((a != b).any(axis=1))
            raise AssertionError(f'{where}: {len(bad)} synapse states differ, first {bad[:8]}')
    if dh9 is not None:
        assert np.array_equal(dh9, oh9), f'{where}: history rings differ'


def oracle_spike_bits(brain):
    """The oracle brain's delay ring as the device's spike register: bit k =
    spike k ticks before the last completed tick, for k below the ring length
    (and 32).  Returns (bits, mask of meaningful bits)."""
    n, R = brain.count, brain.lib.ora_ring(brain.ptr)
    ring = np.zeros((n, R), np.uint8)
    tick = np.zeros(1, np.int64)
    brain.lib.ora_export(brain.ptr, None, None, None, O.ptr(ring, O._u8), O.ptr(tick, O._i64), None)
    T = int(tick[0])
    bits = np.zeros(n, np.uint32)
    for k in range(min(R, 32)):
        tau = T - 1 - k
        if tau >= 0:
            bits |= ring[:, tau % R].astype(np.uint32) << np.uint32(k)
    mask = np.uint32(0xffffffff) if R >= 32 else np.uint32((1 << R) - 1)
    return bits, mask


def assert_bits_equal(dev_bits, brains, where=''):
    """Device spike registers (genome order) against the oracle brains' rings,
    on every bit the oracle's ring holds (delays up to its bound)."""
    off = 0
    for i, b in enumerate(brains):
        ob, mask = oracle_spike_bits(b)
        d = dev_bits[off:off + b.count]
        if not np.array_equal(d & mask, ob & mask):
            bad = np.flatnonzero((d & mask) != (ob & mask))
            raise AssertionError(f'{where}: brain {i}: {len(bad)} spike registers differ, '
                                 f'first {bad[:8]}')
        off += b.count


def oracle_step_threaded(brains, x, ticks, gain, threads=16):
    """All brains stepped by the oracle's threaded batch (independent brains
    sharded over host threads)."""
    lib = O.oracle_lib()
    arr = (O._vp * len(brains))(*[b.ptr for b in brains])
    x = np.ascontiguousarray(x, np.float32)
    outs = np.zeros((len(brains), brains[0].outputs), np.float32)
    status = np.zeros(len(brains), np.int32)
    rc = lib.ora_batch_step(arr, len(brains), O.ptr(x, O._f), x.shape[1], None, ticks, gain,
                            O.ptr(outs, O._f), brains[0].outputs, O.ptr(status, O._i32), threads)
    assert rc == 0
    return outs, status


class Hip:
    """The HIP runtime libns_amd runs on (the same loaded library): device
    buffers and a stream for the device-pointer entry points."""

    def __init__(self):
        ns_amd.lib()
        self.h = ct.CDLL('libamdhip64.so.7')
        self.h.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
        self.h.hipFree.argtypes = [ct.c_void_p]
        self.h.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
        self.h.hipStreamCreate.argtypes = [ct.POINTER(ct.c_void_p)]
        self.h.hipStreamDestroy.argtypes = [ct.c_void_p]
        self.h.hipStreamSynchronize.argtypes = [ct.c_void_p]
        for f in ('hipMalloc', 'hipFree', 'hipMemcpy', 'hipStreamCreate', 'hipStreamDestroy',
                  'hipStreamSynchronize'):
            getattr(self.h, f).restype = ct.c_int

    def alloc(self, nbytes):
        p = ct.c_void_p()
        assert self.h.hipMalloc(ct.byref(p), nbytes) == 0
        return p.value

    def put(self, dst, arr):
        assert self.h.hipMemcpy(dst, arr.ctypes.data, arr.nbytes, 1) == 0      # host to device

    def free(self, p):
        if p:
            assert self.h.hipFree(p) == 0

    def stream(self):
        s = ct.c_void_p()
        assert self.h.hipStreamCreate(ct.byref(s)) == 0
        return s.value

    def sync(self, stream):
        assert self.h.hipStreamSynchronize(stream) == 0

    def get(self, src, arr):
        assert self.h.hipMemcpy(arr.ctypes.data, src, arr.nbytes, 2) == 0      # device to host
        return arr
