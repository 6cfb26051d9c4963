"""Maze task of the serving loop, restated in Python for the tests: the
depth-first maze carve, the 9-channel observation, the move rule and the
reward, step for step, plus a CPU replay of a whole maze request through the
oracle brain.  Test infrastructure only."""
import numpy as np

from oracle import OracleBrain

SIDE, ROOMS = 31, 15
MASK64 = (1 << 64) - 1


def splitmix64(state):
    """One SplitMix64 step: (new state, output)."""
    z = (state + 0x9e3779b97f4a7c15) & MASK64
    r = z
    r = ((r ^ (r >> 30)) * 0xbf58476d1ce4e5b9) & MASK64
    r = ((r ^ (r >> 27)) * 0x94d049bb133111eb) & MASK64
    return z, r ^ (r >> 31)


def make_maze(seed, index):
    """Open-cell rows (bit x of rows[y]) and the goal cell of maze `index` of
    a request seeded `seed`: recursive backtracker from room 0 at (1, 1),
    neighbours in the order north, east, south, west, one drawn as
    SplitMix64 mod count; then the goal room among the other 224."""
    z = ((0x9e3779b97f4a7c15 * (seed + 1)) & MASK64) ^ ((index * 0xd1b54a32d192ed03) & MASK64)
    rows = [0] * SIDE
    visited = [False] * (ROOMS * ROOMS)
    rows[1] |= 1 << 1
    visited[0] = True
    stack = [0]
    while stack:
        r = stack[-1]
        rx, ry = r % ROOMS, r // ROOMS
        cand = []
        if ry > 0 and not visited[r - ROOMS]:
            cand.append(r - ROOMS)
        if rx < ROOMS - 1 and not visited[r + 1]:
            cand.append(r + 1)
        if ry < ROOMS - 1 and not visited[r + ROOMS]:
            cand.append(r + ROOMS)
        if rx > 0 and not visited[r - 1]:
            cand.append(r - 1)
        if not cand:
            stack.pop()
            continue
        z, v = splitmix64(z)
        q = cand[v % len(cand)]
        ax, ay = 1 + 2 * rx, 1 + 2 * ry
        bx, by = 1 + 2 * (q % ROOMS), 1 + 2 * (q // ROOMS)
        rows[(ay + by) // 2] |= 1 << ((ax + bx) // 2)
        rows[by] |= 1 << bx
        visited[q] = True
        stack.append(q)
    z, v = splitmix64(z)
    g = 1 + v % (ROOMS * ROOMS - 1)
    return rows, (1 + 2 * (g % ROOMS), 1 + 2 * (g // ROOMS))


def is_open(rows, x, y):
    return 0 <= x < SIDE and 0 <= y < SIDE and (rows[y] >> x) & 1 == 1


class MazeTask:
    """One maze request: `mazes` mazes of at most `horizon` steps each."""

    def __init__(self, seed, mazes, horizon):
        self.seed, self.mazes, self.horizon = seed, mazes, horizon
        self.index = 0
        self.rows, self.goal = make_maze(seed, 0)
        self.x, self.y = 1, 1
        self.steps = self.total = self.solved = 0
        self.pending = 0
        self.ret = np.float32(0)
        self.motors = [np.float32(0)] * 4

    @property
    def done(self):
        return self.index >= self.mazes

    def observe(self):
        x, y, (gx, gy), rows = self.x, self.y, self.goal, self.rows
        f = np.float32
        return np.array([f(0) if is_open(rows, x, y - 1) else f(1),
                         f(0) if is_open(rows, x + 1, y) else f(1),
                         f(0) if is_open(rows, x, y + 1) else f(1),
                         f(0) if is_open(rows, x - 1, y) else f(1),
                         f(gx > x), f(gx < x), f(gy > y), f(gy < y), f(self.pending)], np.float32)

    def step(self, out):
        if self.done:
            return
        self.pending = 0
        best, top = -1, np.float32(0)
        for m in range(4):
            if out[m] > top:
                top, best = out[m], m
            self.motors[m] = np.float32(self.motors[m] + np.float32(out[m]))
        if best >= 0:
            nx = self.x + (best == 1) - (best == 3)
            ny = self.y + (best == 2) - (best == 0)
            if is_open(self.rows, nx, ny):
                self.x, self.y = nx, ny
        self.steps += 1
        self.total += 1
        nxt = False
        if (self.x, self.y) == self.goal:
            reward = np.float32(1.0 - 0.9 * (float(self.steps) / float(self.horizon)))
            self.ret = np.float32(self.ret + reward)
            self.solved += 1
            self.pending = 1
            nxt = True
        elif self.steps >= self.horizon:
            nxt = True
        if not nxt:
            return
        self.index += 1
        if self.done:
            return
        self.rows, self.goal = make_maze(self.seed, self.index)
        self.x, self.y, self.steps = 1, 1, 0

    def result(self):
        """(return, final_state, motor_spikes) as ns_serve_result holds them."""
        return (self.ret, np.array([self.solved, self.total, self.index, 0], np.float32),
                (int(self.motors[0]), int(self.motors[1])))


def replay(nr, sr, mo, seed, task_seed, mazes, horizon, ticks=16, gain=230.0):
    """A maze request on the CPU: oracle brain plus MazeTask."""
    b = OracleBrain(nr, sr, mo, 9)
    b.reset(seed)
    t = MazeTask(task_seed, mazes, horizon)
    while not t.done:
        out, _ = b.step(t.observe(), ticks, gain)
        t.step(out)
    return t.result()


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
