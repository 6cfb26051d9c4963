"""Maze-task helpers for the tests: the CPU restatement lives in
oracle/tasks.py; this adds a hand-wired goal-seeking genome."""
import numpy as np

from tasks import SIDE, ROOMS, is_open, make_maze, MazeTask, replay_maze as replay  # noqa: F401


def goal_seeker(extra=0, weight=20.0):
    """A hand-wired maze genome: 9 sensor neurons, `extra` idle neurons and 4
    motor neurons; goal-east/west/south/north sensors drive the east/west/
    south/north motors, the reward sensor is a dopamine modulator onto the
    motors, motor synapses carry D1 plasticity."""
    n = 9 + extra + 4
    nr = np.zeros((n, 10), np.float32)
    nr[:, 0] = 4                       # PN relay AdEx
    nr[:, 1] = -1
    nr[:9, 1] = np.arange(9)
    nr[:, 2] = 1
    nr[8, 4] = 3                       # reward sensor: dopamine modulator
    nr[:, 5] = 1.0
    nr[:, 6] = 1.0
    nr[:, 7] = -50.0
    motor = {0: n - 4, 1: n - 3, 2: n - 2, 3: n - 1}     # N, E, S, W
    pairs = [(4, motor[1]), (5, motor[3]), (6, motor[2]), (7, motor[0])]
    rows = []
    for pre, post in pairs:
        for _ in range(3):
            rows.append([pre, post, weight, 1, 1, 0, 0.55, 450.0, 0.0, 4, 4])
    for post in motor.values():
        rows.append([8, post, 1.0, 1, 1, 0, 0.55, 450.0, 0.0, 0, 0])
    sr = np.array(rows, np.float32)
    mo = np.array([motor[0], motor[1], motor[2], motor[3]], np.int32)
    return nr, sr, mo
