"""ns_step_static (one-brain pool in a static device array, no allocation)
against ns_step (the handle's own device pool) and the oracle, through the C
ABI: two calls in a row, the second on a smaller brain with fewer ticks, then
the larger brain again; and a brain past the array's size is refused with the
handle unchanged."""
import ctypes as ct

import numpy as np
import pytest

import ns_amd
import ns_workloads as W
import oracle as O
from parity_util import assert_bits_equal
from test_handles import _handle, handle_state

pytestmark = pytest.mark.gpu
_f = ct.POINTER(ct.c_float)


def _genome(seed, n, fanout=4.0):
    rng = np.random.default_rng(seed)
    w = W._assemble('one', rng, [n], fanout, W._uniform_types([0, 1, 4, 5, 6]), 0.2, 6, 6, 2, 16,
                    -0.5, 0.6, W._stdp_mix, 0.08, seed, 0.5, 230.0, 16, 1)
    return w.brain(0)


def _same_state(ha, hb, n, m):
    a, b = handle_state(ha, n, m), handle_state(hb, n, m)
    assert np.array_equal(a[0].view(np.uint8), b[0].view(np.uint8))
    assert np.array_equal(a[1].view(np.uint8), b[1].view(np.uint8))
    assert np.array_equal(a[2], b[2]) and a[3] == b[3]


def test_static_step_equals_step_and_oracle_on_two_calls():
    L = ns_amd.lib()
    big, small = _genome(101, 3000), _genome(102, 40)
    print('[no-alloc] brains of 3000 and 40 neurons', flush=True)
    hs = {}
    try:
        for name, (nr, sr, mo) in (('big', big), ('small', small)):
            a = _handle(nr, sr, mo, 6, 16)
            b = _handle(nr, sr, mo, 6, 16)
            o = O.OracleBrain(nr, sr, mo, 6, max_delay=16)
            for h in (a, b):
                L.ns_reset(h, 77)
            o.reset(77)
            hs[name] = (a, b, o, len(nr), len(sr), len(mo))
        rng = np.random.default_rng(5)
        for name, ticks in (('big', 16), ('small', 5), ('big', 16)):
            a, b, o, n, m, outs = hs[name]
            x = np.ascontiguousarray((rng.random(6) < 0.5).astype(np.float32))
            oa, ob = np.zeros(outs, np.float32), np.zeros(outs, np.float32)
            assert L.ns_step(a, x.ctypes.data_as(_f), ticks, 230.0, oa.ctypes.data_as(_f)) == 0, ns_amd.error()
            assert L.ns_step_static(b, x.ctypes.data_as(_f), ticks, 230.0, ob.ctypes.data_as(_f)) == 0, \
                ns_amd.error()
            oo, rc = o.step(x, ticks, 230.0)
            assert np.array_equal(oa, ob) and np.array_equal(ob, oo), (name, oa, ob, oo)
            _same_state(a, b, n, m)
            ns, ss, bits, tick = handle_state(b, n, m)
            ons, oss, _, otick = o.export()
            assert tick == otick
            assert np.array_equal(O.state_bytes(ns, O.NEURON_STATE_BYTES),
                                  O.state_bytes(ons, O.NEURON_STATE_BYTES)), name
            assert np.array_equal(O.state_bytes(ss, O.SYNAPSE_STATE_BYTES),
                                  O.state_bytes(oss, O.SYNAPSE_STATE_BYTES)), name
            assert_bits_equal(bits, [o], name)
        a, b, o, n, m, outs = hs['big']
        assert handle_state(b, n, m)[0]['total_spikes'].sum() > 0
    finally:
        for a, b, o, *_ in hs.values():
            L.ns_destroy(a)
            L.ns_destroy(b)
            o.close()


def test_static_step_refuses_a_brain_past_the_array():
    L = ns_amd.lib()
    nr, sr, mo = _genome(103, 12000)
    h = _handle(nr, sr, mo, 6, 16)
    try:
        L.ns_reset(h, 3)
        before = handle_state(h, len(nr), len(sr))
        x = np.zeros(6, np.float32)
        out = np.zeros(len(mo), np.float32)
        assert L.ns_step_static(h, x.ctypes.data_as(_f), 16, 230.0, out.ctypes.data_as(_f)) == -1
        assert 'exceeds the static device pool of 4194304 bytes' in ns_amd.error()
        after = handle_state(h, len(nr), len(sr))
        assert np.array_equal(before[0].view(np.uint8), after[0].view(np.uint8)) and before[3] == after[3]
    finally:
        L.ns_destroy(h)
