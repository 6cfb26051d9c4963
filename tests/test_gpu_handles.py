"""ns_reset / ns_step (gym_adapter/native.cpp:89-109) on the device, through
the C ABI: every call's outputs and the handle's whole state (records, spike
registers, tick) equal the oracle brain's, for every cell type, plasticity,
modulators and a 16 ms delay bound; and the reference's two-step batch path
(ns_create + ns_reset per genome, then ns_gpu_create over the handles, as
gym_adapter/gpu.py:29-45 does) runs unchanged on the same handles."""
import ctypes as ct

import numpy as np
import pytest

import ns_amd
import ns_workloads as W
import oracle as O
from parity_util import assert_bits_equal, oracle_brains, oracle_step_all
from test_handles import _handle, handle_state

pytestmark = pytest.mark.gpu
_f = ct.POINTER(ct.c_float)


@pytest.mark.parametrize('types,max_delay', [((0, 1, 4, 5, 6), 8), ((0, 1, 2, 3, 4, 5), 8),
                                            ((0, 1, 4, 5, 6), 16)])
def test_step_equals_the_oracle(types, max_delay):
    L = ns_amd.lib()
    w = W.small_mix(seed=95, brains=5, calls=10, types=types, max_delay=max_delay, lo=10, hi=70)
    print(f'[handles] types={types} max_delay={max_delay} neurons={w.total_neurons}', flush=True)
    seeds = w.seeds()
    for i in range(w.count):
        nr, sr, mo = w.brain(i)
        h = _handle(nr, sr, mo, w.channels, max_delay)
        ob = O.OracleBrain(nr, sr, mo, w.channels, max_delay=max_delay)
        try:
            L.ns_reset(h, int(seeds[i]))
            ob.reset(seeds[i])
            for c in range(w.calls):
                x = np.ascontiguousarray(w.inputs(c)[i])
                ticks = 16 if c % 3 else 7
                out = np.zeros(len(mo), np.float32)
                assert L.ns_step(h, x.ctypes.data_as(_f), ticks, w.gain, out.ctypes.data_as(_f)) == 0, \
                    ns_amd.error()
                oo, rc = ob.step(x, ticks, w.gain)
                assert rc == 0 and np.array_equal(out, oo), (i, c)
            ns, ss, bits, tick = handle_state(h, len(nr), len(sr))
            ons, oss, _, otick = ob.export()
            assert tick == otick
            assert np.array_equal(O.state_bytes(ns, O.NEURON_STATE_BYTES),
                                  O.state_bytes(ons, O.NEURON_STATE_BYTES)), i
            assert np.array_equal(O.state_bytes(ss, O.SYNAPSE_STATE_BYTES),
                                  O.state_bytes(oss, O.SYNAPSE_STATE_BYTES)), i
            assert_bits_equal(bits, [ob], f'brain {i}')
        finally:
            L.ns_destroy(h)
            ob.close()


def test_reference_two_step_batch_path_on_stateful_handles():
    """gpu.py:29-45: handles made and reset one by one, a batch made over
    them (it copies the genomes), the handles freed, the batch stepped."""
    L = ns_amd.lib()
    w = W.small_mix(seed=96, brains=6, calls=6, types=(0, 1, 4, 5))
    hs = []
    for i in range(w.count):
        nr, sr, mo = w.brain(i)
        hs.append(_handle(nr, sr, mo, w.channels))
        L.ns_reset(hs[-1], 7)
    arr = (ct.c_void_p * len(hs))(*hs)
    batch = L.ns_gpu_create(arr, len(hs))
    for h in hs:
        L.ns_destroy(h)
    assert batch, ns_amd.error()
    try:
        seeds = np.ascontiguousarray(w.seeds(), np.uint32)
        mask = np.ones(w.count, np.uint8)
        assert L.ns_gpu_reset(batch, mask.ctypes.data_as(ns_amd._u8), seeds.ctypes.data_as(ns_amd._u32)) == 0
        brains = oracle_brains(w)
        out = np.zeros((w.count, w.outputs), np.float32)
        status = np.zeros(w.count, np.int32)
        for c in range(w.calls):
            x = np.ascontiguousarray(w.inputs(c))
            assert L.ns_gpu_step(batch, x.ctypes.data_as(_f), mask.ctypes.data_as(ns_amd._u8), 16, w.gain,
                                 out.ctypes.data_as(_f), status.ctypes.data_as(ns_amd._i32)) == 0
            oo, _ = oracle_step_all(brains, x, 16, w.gain)
            assert np.array_equal(out, oo), c
    finally:
        L.ns_gpu_destroy(batch)


def test_batch_keeps_the_genomes_it_was_made_from():
    """A batch shares its genomes with the handles it was made from; a handle
    that then changes its genome (its original indices, which seed the
    initial states) takes its own copy, so the batch still steps the genomes
    it was made from: outputs and states equal a batch made directly from
    the same genomes."""
    w = W.small_mix(seed=5, brains=6, calls=4, types=(0, 1, 4, 5), lo=20, hi=60)
    print(f'[handles] shared genomes: brains={w.count}', flush=True)
    hs = ns_amd.GenomeHandles([w.brain(i) for i in range(w.count)], w.channels, max_delay=w.max_delay)
    ref = ns_amd.GpuBatch(workload=w)
    b = ns_amd.GpuBatch(workload=w)
    try:
        L = b.lib
        L.ns_gpu_destroy(b.ptr)
        b.ptr = L.ns_gpu_create((ct.c_void_p * w.count)(*hs.ptrs), w.count)
        assert b.ptr, ns_amd.error()
        n0, s0 = int(w.neuron_counts[0]), int(w.edge_counts[0])
        perm_n = np.ascontiguousarray(np.arange(n0)[::-1] + 1000, np.uint32)
        perm_s = np.ascontiguousarray(np.arange(max(s0, 1))[::-1] + 5000, np.uint32)
        L.ns_original_indices(hs.ptrs[0], perm_n.ctypes.data_as(ns_amd._u32),
                              perm_s.ctypes.data_as(ns_amd._u32))
        for batch in (ref, b):
            batch.reset(w.seeds())
        for call in range(4):
            x = w.inputs(call)
            o_ref, st_ref = ref.step(x, w.ticks, w.gain)
            o_b, st_b = b.step(x, w.ticks, w.gain)
            assert np.array_equal(o_ref, o_b), f'call {call}: outputs differ'
        for a_ref, a_b in zip(ref.export(), b.export()):
            if a_ref is not None:
                assert a_ref.tobytes() == a_b.tobytes(), 'batch state differs'
    finally:
        ref.close()
        b.close()
        hs.close()
