"""The serving tasks' two CPU restatements agree: the Python one
(oracle/tasks.py, which the GPU serving tests compare against the device) and
the C++ one in the oracle library (ora_linear_episode, ora_maze_request,
ora_maze_layout, which the serving CPU baseline runs)."""
import numpy as np

import ns_workloads as W
import tasks
import maze_replay as MR


def same(a, b):
    return a[0] == b[0] and np.array_equal(a[1], b[1]) and a[2] == b[2]


def test_maze_layouts_agree():
    for seed in (0, 5, 126, 2**32 - 1):
        for index in range(3):
            assert tasks.maze_layout_native(seed, index) == tasks.make_maze(seed, index)


def test_maze_requests_agree():
    nr, sr, mo = MR.goal_seeker()
    for tseed in (126, 29, 5):          # the first two solve a maze
        py = tasks.replay_maze(nr, sr, mo, 1234, tseed, 3, 16)
        nat, steps = tasks.maze_request_native(nr, sr, mo, 1234, tseed, 3, 16)
        assert same(py, nat) and steps == int(py[1][1])
    rng = np.random.default_rng(3)
    for i in range(3):
        g = W.maze_genome(rng, 60)
        py = tasks.replay_maze(*g, 77 + i, 900 + i, 2, 20)
        nat, _ = tasks.maze_request_native(*g, 77 + i, 900 + i, 2, 20)
        assert same(py, nat)


def test_linear_episodes_agree():
    rng = np.random.default_rng(17)
    for i in range(4):
        nr, sr, mo = W.episode_genome(rng, int(rng.integers(16, 200)))
        py = tasks.replay_linear(nr, sr, mo, 1000 + i, 5000 + i, 30, 0.01)
        nat, steps = tasks.linear_episode_native(nr, sr, mo, 1000 + i, 5000 + i, 30, 0.01)
        assert same(py, nat) and steps == 30
