"""Serving loop on the maze task: requests run several mazes each with the
brain's state kept; every result (return, mazes solved, steps, motor totals)
equals the request run alone and a CPU replay through the oracle brain plus
the restated maze task.  Goal-seeking genomes on seeds whose first maze they
solve exercise the reward path (reward channel into a dopamine neuron, D1
plasticity)."""
import ctypes as ct

import numpy as np
import pytest

import ns_amd
import ns_workloads as W
import maze_replay as MR

pytestmark = pytest.mark.gpu

MAZES, HORIZON = 3, 16


def config(per_launch=4):
    return ns_amd.ServeConfig(neuron_capacity=100000, synapse_capacity=800000, episode_capacity=1024,
                              steps_per_episode=1, ticks_per_step=16, steps_per_launch=per_launch,
                              noise_sigma=0.0, input_gain=230.0, max_episode_neurons=512,
                              task=ns_amd.TASK_MAZE, maze_episodes=MAZES, maze_horizon=HORIZON)


def requests():
    out = []
    nr, sr, mo = MR.goal_seeker()
    for k, tseed in enumerate((126, 29, 136, 193)):
        out.append((nr, sr, mo, 1234 + k, tseed))
    rng = np.random.default_rng(9)
    for i in range(8):
        size = int(np.exp(rng.uniform(np.log(16), np.log(300))))
        nr, sr, mo = W.maze_genome(rng, size)
        out.append((nr, sr, mo, 500 + i, 4000 + i))
    return out


def run(reqs, cfg):
    L = ns_amd.lib()
    srv = L.ns_serve_create(ct.byref(cfg))
    assert srv, ns_amd.error()
    f = ct.POINTER(ct.c_float)
    ids = {}
    try:
        for k, (nr, sr, mo, seed, tseed) in enumerate(reqs):
            i = L.ns_serve_submit(srv, len(nr), len(sr), nr.ctypes.data_as(f), sr.ctypes.data_as(f),
                                  mo.ctypes.data_as(ct.POINTER(ct.c_int)), seed, tseed)
            assert i >= 0, ns_amd.error()
            ids[i] = k
        assert L.ns_serve_drain(srv) == 0, ns_amd.error()
        res = (ns_amd.ServeResult * len(reqs))()
        n = L.ns_serve_poll(srv, res, len(reqs))
        assert n == len(reqs)
        return {ids[r.id]: (np.float32(r.episode_return), np.array(r.final_state[:], np.float32),
                            tuple(r.motor_spikes)) for r in res[:n]}
    finally:
        L.ns_serve_destroy(srv)


def same(a, b):
    return a[0] == b[0] and np.array_equal(a[1], b[1]) and a[2] == b[2]


def test_maze_requests_mix_equals_alone_equals_cpu():
    reqs = requests()
    print(f'[maze] {len(reqs)} requests x {MAZES} mazes x <= {HORIZON} steps', flush=True)
    mixed = run(reqs, config())
    assert mixed[0][1][0] >= 1, 'the goal seeker on seed 126 solves its first maze'
    for k in (0, 6):
        alone = run([reqs[k]], config(per_launch=3))
        assert same(alone[0], mixed[k]), (k, alone[0], mixed[k])
    for k in (0, 1, 2, 3, 4, 9):
        nr, sr, mo, seed, tseed = reqs[k]
        cpu = MR.replay(nr, sr, mo, seed, tseed, MAZES, HORIZON)
        assert same(cpu, mixed[k]), (k, cpu, mixed[k])


def test_largest_maze_requests_match_cpu():
    """Maze requests of 2,048 and 1,400 neurons (the largest C5 sizes) beside
    a small one; each equals the CPU replay."""
    rng = np.random.default_rng(19)
    reqs = []
    for i, size in enumerate((2048, 60, 1400)):
        nr, sr, mo = W.maze_genome(rng, size)
        reqs.append((nr, sr, mo, 700 + i, 5000 + i))
    cfg = ns_amd.ServeConfig(neuron_capacity=200000, synapse_capacity=1600000, episode_capacity=1024,
                             steps_per_episode=1, ticks_per_step=16, steps_per_launch=4,
                             noise_sigma=0.0, input_gain=230.0, max_episode_neurons=2048,
                             task=ns_amd.TASK_MAZE, maze_episodes=MAZES, maze_horizon=HORIZON)
    print(f'[maze] sizes {[len(r[0]) for r in reqs]}', flush=True)
    mixed = run(reqs, cfg)
    for k in (0, 2):
        nr, sr, mo, seed, tseed = reqs[k]
        cpu = MR.replay(nr, sr, mo, seed, tseed, MAZES, HORIZON)
        assert same(cpu, mixed[k]), (k, cpu, mixed[k])
