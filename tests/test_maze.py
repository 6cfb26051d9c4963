"""Maze task of the serving loop on the CPU: the Python restatement of the
maze carve equals the library's own (host build of the function the device
runs), every maze is a perfect maze (a spanning tree of its 225 rooms), and
the replay's reward follows its closed form."""
import ctypes as ct
from collections import deque

import numpy as np

import ns_amd
import maze_replay as MR


def library_maze(seed, index):
    rows = (ct.c_uint32 * 31)()
    goal = (ct.c_uint32 * 2)()
    ns_amd.lib().ns_maze_layout(seed, index, rows, goal)
    return list(rows), tuple(goal)


def test_carve_matches_library():
    for seed in (0, 1, 7, 126, 99991, 2**31, 2**32 - 1):
        for index in range(5):
            assert library_maze(seed, index) == MR.make_maze(seed, index), (seed, index)


def test_mazes_are_perfect():
    for seed in range(12):
        for index in range(3):
            rows, (gx, gy) = MR.make_maze(seed, index)
            open_cells = {(x, y) for y in range(31) for x in range(31) if MR.is_open(rows, x, y)}
            # border closed; rooms all open; 224 passages, no other open cell
            assert all(not MR.is_open(rows, x, y) for x in range(31) for y in (0, 30))
            assert all(not MR.is_open(rows, x, y) for y in range(31) for x in (0, 30))
            rooms = {(1 + 2 * i, 1 + 2 * j) for i in range(15) for j in range(15)}
            assert rooms <= open_cells and len(open_cells) == 225 + 224
            # connected: a tree on 225 rooms with 224 passages
            seen, todo = {(1, 1)}, deque([(1, 1)])
            while todo:
                x, y = todo.popleft()
                for dx, dy in ((0, -1), (1, 0), (0, 1), (-1, 0)):
                    c = (x + dx, y + dy)
                    if c in open_cells and c not in seen:
                        seen.add(c)
                        todo.append(c)
            assert seen == open_cells
            assert (gx, gy) in rooms and (gx, gy) != (1, 1)


def test_goal_seeker_reward():
    """Seed 126's first maze has its goal two steps from the start along the
    goal sensors; the hand-wired brain reaches it in 2 steps:
    reward 1 - 0.9 * 2 / 16."""
    nr, sr, mo = MR.goal_seeker()
    ret, state, _ = MR.replay(nr, sr, mo, 1234, 126, mazes=1, horizon=16)
    assert state[0] == 1 and state[1] == 2
    assert ret == np.float32(1.0 - 0.9 * 2 / 16)


def test_unsolved_maze_runs_to_horizon():
    t = MR.MazeTask(5, 2, 7)
    while not t.done:
        t.step(np.zeros(4, np.float32))       # no spikes: the agent stays
    ret, state, motors = t.result()
    assert ret == 0 and list(state) == [0, 14, 2, 0] and motors == (0, 0)
