"""Error paths of the extensions: invalid replacements and serving requests
return -1 (or NULL) with a message, change nothing, and leave the batch or
server usable."""
import ctypes as ct

import numpy as np
import pytest

import ns_amd
import ns_workloads as W
from parity_util import assert_state_equal, oracle_brains, oracle_export_all, oracle_step_all

pytestmark = pytest.mark.gpu


def test_invalid_replacements_change_nothing():
    w = W.small_mix(seed=51, brains=6, calls=4, types=(0, 1, 4, 5))
    print(f'[errors] replace on {w.count} brains', flush=True)
    batch = ns_amd.GpuBatch(workload=w)
    good = ns_amd.GenomeHandles([w.brain(0)], w.channels, max_delay=8)
    other = ns_amd.GenomeHandles([w.brain(1)], w.channels + 1, max_delay=8)
    try:
        batch.reserve(0.19, 0.25)
        batch.reset(w.seeds())
        brains = oracle_brains(w)
        cases = [
            (([2, 2], good, [0, 0], [1, 2]), 'Invalid brain index'),       # duplicate
            (([w.count], good, [0], [1]), 'Invalid brain index'),          # out of range
            (([-1], good, [0], [1]), 'Invalid brain index'),
            (([3], other, [0], [1]), 'Batch ports differ'),                # other channels
        ]
        for args, message in cases:
            with pytest.raises(RuntimeError, match=message):
                batch.replace(*args)
        L = batch.lib
        idx = np.array([1], np.int32)
        null = (ct.c_void_p * 1)(None)
        seeds = np.array([5], np.uint32)
        assert L.ns_gpu_replace(batch.ptr, 1, idx.ctypes.data_as(ct.POINTER(ct.c_int)), null,
                                seeds.ctypes.data_as(ct.POINTER(ct.c_uint32)), None) == -1
        assert 'handle' in ns_amd.error()
        assert L.ns_gpu_replace(batch.ptr, -1, None, None, None, None) == -1
        assert L.ns_gpu_set_graph(batch.ptr, 2) == -1
        assert 'graph mode' in ns_amd.error()
        # nothing changed: the batch still equals the untouched oracle brains
        assert_state_equal(batch.export(), oracle_export_all(brains), 'after failed calls')
        for call in range(w.calls):
            x = w.inputs(call)
            out, _ = batch.step(x, w.ticks, w.gain)
            oo, _ = oracle_step_all(brains, x, w.ticks, w.gain)
            assert np.array_equal(out, oo)
    finally:
        good.close()
        other.close()
        batch.close()


def _config(**kw):
    c = dict(neuron_capacity=20000, synapse_capacity=100000, episode_capacity=64, steps_per_episode=4,
             ticks_per_step=16, steps_per_launch=2, noise_sigma=0.0, input_gain=230.0,
             max_episode_neurons=256, task=0, maze_episodes=0, maze_horizon=0)
    c.update(kw)
    return ns_amd.ServeConfig(**c)


def test_serving_rejects_bad_configs_and_requests():
    L = ns_amd.lib()
    for bad in (dict(task=2), dict(maze_horizon=-1), dict(ticks_per_step=0), dict(noise_sigma=-1.0)):
        cfg = _config(**bad)
        assert not L.ns_serve_create(ct.byref(cfg))
        assert 'Invalid serving configuration' in ns_amd.error()
    cfg = _config(task=ns_amd.TASK_MAZE, maze_episodes=1, maze_horizon=4)
    srv = L.ns_serve_create(ct.byref(cfg))
    assert srv, ns_amd.error()
    f = ct.POINTER(ct.c_float)
    try:
        rng = np.random.default_rng(2)
        nr, sr, mo = W.maze_genome(rng, 40)
        nr = nr.copy()
        nr[3, 1] = 9                                  # sensor channel 9: a maze has channels 0..8
        assert L.ns_serve_submit(srv, len(nr), len(sr), nr.ctypes.data_as(f), sr.ctypes.data_as(f),
                                 mo.ctypes.data_as(ct.POINTER(ct.c_int)), 1, 2) == -1
        assert 'Invalid neuron gene' in ns_amd.error()
        mo = mo.copy()
        nr, sr, _ = W.maze_genome(rng, 40)
        mo[2] = len(nr)                               # motor outside the brain
        assert L.ns_serve_submit(srv, len(nr), len(sr), nr.ctypes.data_as(f), sr.ctypes.data_as(f),
                                 mo.ctypes.data_as(ct.POINTER(ct.c_int)), 1, 2) == -1
        assert 'Invalid motor index' in ns_amd.error()
        big = W.maze_genome(rng, 600)                 # larger than the largest slot (256 neurons)
        nr, sr, mo = big
        assert L.ns_serve_submit(srv, len(nr), len(sr), nr.ctypes.data_as(f), sr.ctypes.data_as(f),
                                 mo.ctypes.data_as(ct.POINTER(ct.c_int)), 1, 2) == -1
        assert 'largest slot' in ns_amd.error()
        ok = W.maze_genome(rng, 40)                   # the server still serves a good request
        nr, sr, mo = ok
        assert L.ns_serve_submit(srv, len(nr), len(sr), nr.ctypes.data_as(f), sr.ctypes.data_as(f),
                                 mo.ctypes.data_as(ct.POINTER(ct.c_int)), 1, 2) >= 0
        assert L.ns_serve_drain(srv) == 0, ns_amd.error()
        res = (ns_amd.ServeResult * 2)()
        assert L.ns_serve_poll(srv, res, 2) == 1
    finally:
        L.ns_serve_destroy(srv)
