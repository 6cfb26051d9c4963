"""Genome handles with state (ns_reset / ns_export, and ns_step's checks) on
the CPU: the host reset equals the oracle's reset byte for byte, with and
without original indices (gym_adapter/native.cpp:89-94), and ns_step rejects
what the reference's ns_step rejects, with its messages, before any device
work (native.cpp:97-98)."""
import ctypes as ct

import numpy as np
import pytest

import ns_amd
import ns_workloads as W
import oracle as O

_f = ct.POINTER(ct.c_float)


def _handle(nr, sr, mo, channels, max_delay=8):
    L = ns_amd.lib()
    h = L.ns_create_ex(len(nr), len(sr), channels, len(mo), nr.ctypes.data_as(_f), sr.ctypes.data_as(_f),
                       mo.ctypes.data_as(ct.POINTER(ct.c_int)), max_delay)
    assert h, ns_amd.error()
    return h


def handle_state(h, n, m):
    ns = np.zeros(n, O.NEURON_STATE)
    ss = np.zeros(m, O.SYNAPSE_STATE)
    bits = np.zeros(n, np.uint32)
    tick = np.zeros(1, np.int64)
    assert ns_amd.lib().ns_export(h, ns.ctypes.data, ss.ctypes.data, bits.ctypes.data_as(ns_amd._u32),
                                  tick.ctypes.data_as(ns_amd._i64)) == 0
    return ns, ss, bits, int(tick[0])


@pytest.mark.parametrize('originals', [False, True])
def test_reset_equals_the_oracle_reset(originals):
    L = ns_amd.lib()
    w = W.small_mix(seed=91, brains=4, types=(0, 1, 2, 3, 4, 5, 6))
    rng = np.random.default_rng(5)
    for i in range(w.count):
        nr, sr, mo = w.brain(i)
        h = _handle(nr, sr, mo, w.channels)
        ob = O.OracleBrain(nr, sr, mo, w.channels)
        try:
            if originals:
                on = np.ascontiguousarray(rng.permutation(len(nr)) + 3, np.uint32)
                osyn = np.ascontiguousarray(rng.permutation(len(sr)) + 11, np.uint32)
                L.ns_original_indices(h, on.ctypes.data_as(ns_amd._u32), osyn.ctypes.data_as(ns_amd._u32))
                ob.lib.ora_original_indices(ob.ptr, on.ctypes.data_as(O._u32), osyn.ctypes.data_as(O._u32))
            seed = 1000 + 17 * i
            L.ns_reset(h, seed)
            ob.reset(seed)
            ns, ss, bits, tick = handle_state(h, len(nr), len(sr))
            ons, oss, oh9, otick = ob.export()
            assert tick == otick == 0 and not bits.any() and not oh9.any()
            assert np.array_equal(O.state_bytes(ns, O.NEURON_STATE_BYTES),
                                  O.state_bytes(ons, O.NEURON_STATE_BYTES)), i
            assert np.array_equal(O.state_bytes(ss, O.SYNAPSE_STATE_BYTES),
                                  O.state_bytes(oss, O.SYNAPSE_STATE_BYTES)), i
        finally:
            L.ns_destroy(h)
            ob.close()


def test_fresh_handle_state_is_zero():
    """native.cpp:46-82 value-initialises the state vectors of a new handle."""
    nr, sr, mo = W.small_mix(seed=92, brains=1).brain(0)
    h = _handle(nr, sr, mo, 4)
    try:
        ns, ss, bits, tick = handle_state(h, len(nr), len(sr))
        assert not ns.view(np.uint8).any() and not ss.view(np.uint8).any() and not bits.any()
        assert tick == 0
    finally:
        ns_amd.lib().ns_destroy(h)


@pytest.mark.parametrize('ticks,gain,level,message', [
    (0, 230.0, 0.5, 'Invalid neural integration interval'),
    (10001, 230.0, 0.5, 'Invalid neural integration interval'),
    (16, -1.0, 0.5, 'Invalid neural integration interval'),
    (16, float('nan'), 0.5, 'Invalid neural integration interval'),
    (16, 230.0, 1.5, 'Invalid encoded observation'),
    (16, 230.0, float('inf'), 'Invalid encoded observation'),
])
def test_step_checks_like_the_reference(ticks, gain, level, message):
    L = ns_amd.lib()
    nr, sr, mo = W.small_mix(seed=93, brains=1).brain(0)
    h = _handle(nr, sr, mo, 4)
    try:
        L.ns_reset(h, 3)
        x = np.full(4, level, np.float32)
        out = np.zeros(len(mo), np.float32)
        assert L.ns_step(h, x.ctypes.data_as(_f), ticks, gain, out.ctypes.data_as(_f)) == -1
        assert ns_amd.error() == message
        if O.ref_available():
            r = O.RefBrain(nr, sr, mo, 4)
            r.reset(3)
            rc = r.lib.ns_step(r.ptr, x.ctypes.data_as(_f), ticks, gain, out.ctypes.data_as(_f))
            assert rc == -1 and r.lib.ns_error().decode() == message
            r.close()
    finally:
        L.ns_destroy(h)
