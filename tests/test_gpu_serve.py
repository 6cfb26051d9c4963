"""Serving loop: episodes admitted by free pool capacity run concurrently;
each episode's result is bit-identical alone, in the mix, and to a CPU
replay through the oracle (brain) plus the task restated in float32
(oracle/tasks.py)."""
import ctypes as ct
import time

import numpy as np
import pytest

import ns_amd
import ns_workloads as W
from tasks import replay_linear

pytestmark = pytest.mark.gpu

STEPS = 40


def config(**kw):
    c = ns_amd.ServeConfig(neuron_capacity=kw.get('neurons', 200000),
                           synapse_capacity=kw.get('synapses', 1000000),
                           episode_capacity=kw.get('episodes', 4096),
                           steps_per_episode=STEPS, ticks_per_step=16,
                           steps_per_launch=kw.get('per_launch', 5), noise_sigma=0.01,
                           input_gain=230.0)
    return c


def genomes(n, seed=4):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        size = int(np.exp(rng.uniform(np.log(16), np.log(300))))
        nr, sr, mo = W.episode_genome(rng, size)
        out.append((nr, sr, mo, 1000 + i, 77 + i))
    return out


def run(eps, cfg, gap=0.0):
    L = ns_amd.lib()
    srv = L.ns_serve_create(ct.byref(cfg))
    assert srv, ns_amd.error()
    f = ct.POINTER(ct.c_float)
    ids = {}
    try:
        for k, (nr, sr, mo, seed, tseed) in enumerate(eps):
            i = L.ns_serve_submit(srv, len(nr), len(sr), nr.ctypes.data_as(f), sr.ctypes.data_as(f),
                                  mo.ctypes.data_as(ct.POINTER(ct.c_int)), seed, tseed)
            assert i >= 0, ns_amd.error()
            ids[i] = k
            if gap:
                time.sleep(gap)
        assert L.ns_serve_drain(srv) == 0, ns_amd.error()
        res = (ns_amd.ServeResult * len(eps))()
        n = L.ns_serve_poll(srv, res, len(eps))
        assert n == len(eps)
        stats = ns_amd.ServeStats()
        L.ns_serve_stats_get(srv, ct.byref(stats))
        out = {}
        for r in res[:n]:
            out[ids[r.id]] = (np.float32(r.episode_return), np.array(r.final_state[:], np.float32),
                              tuple(r.motor_spikes))
        return out, stats
    finally:
        L.ns_serve_destroy(srv)


def test_serving_mix_equals_alone_equals_cpu():
    eps = genomes(24)
    print(f'[serve] {len(eps)} episodes x {STEPS} steps', flush=True)
    mixed, stats = run(eps, config(), gap=0.002)
    assert stats.finished == len(eps) and stats.launches > 0
    for k in (0, 5, 17):
        alone, _ = run([eps[k]], config())
        assert alone[0][0] == mixed[k][0] and np.array_equal(alone[0][1], mixed[k][1])
        assert alone[0][2] == mixed[k][2]
    for k in (0, 9):
        nr, sr, mo, seed, tseed = eps[k]
        ret, x, m = replay_linear(nr, sr, mo, seed, tseed, STEPS, 0.01)
        assert ret == mixed[k][0], (ret, mixed[k][0])
        assert np.array_equal(x, mixed[k][1]) and m == mixed[k][2]


def test_admission_waits_for_free_capacity():
    eps = genomes(10, seed=8)
    need = max(len(e[0]) for e in eps)
    cfg = config(neurons=need + 5, per_launch=10)
    res, stats = run(eps, cfg)
    assert len(res) == len(eps) and stats.admitted == len(eps) and stats.finished == len(eps)
    alone, _ = run([eps[3]], config())
    assert alone[0][0] == res[3][0] and np.array_equal(alone[0][1], res[3][1])


def test_largest_episode_classes_match_cpu():
    """Episodes of the largest slot classes (2,600 and 4,096 neurons: one
    workgroup each, state in HBM) beside small ones; each result equals the
    request alone and the CPU replay."""
    rng = np.random.default_rng(12)
    eps = []
    for i, size in enumerate((4096, 40, 2600, 150)):
        nr, sr, mo = W.episode_genome(rng, size)
        eps.append((nr, sr, mo, 3000 + i, 90 + i))
    print(f'[serve] sizes {[len(e[0]) for e in eps]} x {STEPS} steps', flush=True)
    cfg = config(neurons=400000, synapses=2000000)
    mixed, stats = run(eps, cfg)
    assert stats.finished == len(eps)
    alone, _ = run([eps[0]], config(neurons=400000, synapses=2000000))
    assert alone[0][0] == mixed[0][0] and np.array_equal(alone[0][1], mixed[0][1])
    for k in (0, 2):
        nr, sr, mo, seed, tseed = eps[k]
        ret, x, m = replay_linear(nr, sr, mo, seed, tseed, STEPS, 0.01)
        assert ret == mixed[k][0], (k, ret, mixed[k][0])
        assert np.array_equal(x, mixed[k][1]) and m == mixed[k][2]
