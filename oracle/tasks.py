"""CPU restatements of the serving loop's two tasks, for the parity tests and
the serving CPU baseline of bench.py.  Test infrastructure only: the product
runs these tasks on the device (neural-selection_amd/csrc/ns_serve.hip).

* Linear task: the seeded 4-d linear system x' = A x + B u + sigma * n
  observed through an 8-channel opponent rate code, 2 motors; float32
  arithmetic in the device's evaluation order.
* Maze task: the depth-first 31 x 31 maze carve, the 9-channel observation,
  the move rule and the reward, step for step.

A replay runs a whole request through the oracle brain (OracleBrain, the C++
restatement of the reference's tick path) and the task.
"""
import ctypes as ct

import numpy as np

from oracle import OracleBrain, oracle_lib, ptr, _f, _i32, _u32

MASK64 = (1 << 64) - 1


def splitmix64(state):
    """One SplitMix64 step: (new state, output)."""
    z = (state + 0x9e3779b97f4a7c15) & MASK64
    r = z
    r = ((r ^ (r >> 30)) * 0xbf58476d1ce4e5b9) & MASK64
    r = ((r ^ (r >> 27)) * 0x94d049bb133111eb) & MASK64
    return z, r ^ (r >> 31)


# ---------------------------------------------------------------------------
# Linear task
# ---------------------------------------------------------------------------
def make_linear_task(seed):
    """x0, A, B of a linear request (SplitMix64 -> float32)."""
    z = (0x9e3779b97f4a7c15 * (seed + 1)) & MASK64

    def nxt():
        nonlocal z
        z, r = splitmix64(z)
        return np.float32((r >> 11) * (1.0 / 9007199254740992.0))
    f = np.float32
    x = np.array([f(2) * nxt() - f(1) for _ in range(4)], np.float32)
    A = np.zeros((4, 4), np.float32)
    for i in range(4):
        row = [f(2) * nxt() - f(1) for _ in range(4)]
        norm = f(0)
        for v in row:
            norm = f(norm + abs(v))
        for j in range(4):
            A[i, j] = f(f(0.9) * row[j]) / norm
    B = np.array([nxt() - f(0.5) for _ in range(8)], np.float32).reshape(4, 2)
    return x, A, B


class LinearTask:
    """One linear request of `steps` steps with noise `sigma`."""

    def __init__(self, seed, steps, sigma, ticks=16):
        self.x, self.A, self.B = make_linear_task(seed)
        self.s = ((seed * 2654435761 + 12345) & 0xffffffff) or 1
        self.steps, self.sigma, self.ticks = steps, np.float32(sigma), np.float32(ticks)
        self.done_steps = 0
        self.ret = np.float32(0)
        self.m = [np.float32(0), np.float32(0)]

    @property
    def done(self):
        return self.done_steps >= self.steps

    def _uniform(self):
        s = self.s
        s ^= (s << 13) & 0xffffffff
        s ^= s >> 17
        s ^= (s << 5) & 0xffffffff
        self.s = s
        return np.float32((s >> 8) * (1.0 / 16777216.0))

    def observe(self):
        f = np.float32
        inp = np.zeros(8, np.float32)
        for d in range(4):
            inp[2 * d] = min(f(1), max(f(0), self.x[d]))
            inp[2 * d + 1] = min(f(1), max(f(0), -self.x[d]))
        return inp

    def step(self, out):
        f = np.float32
        u0, u1 = f(out[0]) / self.ticks, f(out[1]) / self.ticks
        nx = np.zeros(4, np.float32)
        for i in range(4):
            acc = f(0)
            for j in range(4):
                acc = f(acc + f(self.A[i, j] * self.x[j]))
            acc = f(acc + f(self.B[i, 0] * u0))
            acc = f(acc + f(self.B[i, 1] * u1))
            n = f(-6)
            for _ in range(12):
                n = f(n + self._uniform())
            nx[i] = f(acc + f(self.sigma * n))
        cost = f(0)
        for i in range(4):
            cost = f(cost + f(nx[i] * nx[i]))
        self.x = nx
        self.ret = f(self.ret - cost)
        self.m = [f(self.m[0] + out[0]), f(self.m[1] + out[1])]
        self.done_steps += 1

    def result(self):
        """(return, final_state, motor_spikes) as ns_serve_result holds them."""
        return self.ret, self.x.copy(), (int(self.m[0]), int(self.m[1]))


def replay_linear(nr, sr, mo, seed, task_seed, steps, sigma, ticks=16, gain=230.0):
    """A linear request on the CPU: oracle brain plus LinearTask."""
    b = OracleBrain(nr, sr, mo, 8)
    b.reset(seed)
    t = LinearTask(task_seed, steps, sigma, ticks)
    while not t.done:
        out, _ = b.step(t.observe(), ticks, gain)
        t.step(out)
    return t.result()


# ---------------------------------------------------------------------------
# Maze task
# ---------------------------------------------------------------------------
SIDE, ROOMS = 31, 15


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


def replay_maze(nr, sr, mo, seed, task_seed, mazes, horizon, ticks=16, gain=230.0):
    """A maze request on the CPU: oracle brain plus MazeTask."""
    b = OracleBrain(nr, sr, mo, 9)
    b.reset(seed)
    t = MazeTask(task_seed, mazes, horizon)
    while not t.done:
        out, _ = b.step(t.observe(), ticks, gain)
        t.step(out)
    return t.result()


# ---------------------------------------------------------------------------
# The same two tasks run whole in the oracle library (ora_linear_episode,
# ora_maze_request): a second restatement, and one that runs without the
# interpreter, for the serving CPU baseline.
# ---------------------------------------------------------------------------
def _native(brain, fn, *args):
    ret = np.zeros(1, np.float32)
    state = np.zeros(4, np.float32)
    motors = np.zeros(2, np.int32)
    steps = fn(brain.ptr, *args, ptr(ret, _f), ptr(state, _f), ptr(motors, _i32))
    return (ret[0], state, (int(motors[0]), int(motors[1]))), steps


def linear_episode_native(nr, sr, mo, seed, task_seed, steps, sigma, ticks=16, gain=230.0):
    """replay_linear in C++: ((return, final x, motors), steps)."""
    b = OracleBrain(nr, sr, mo, 8)
    b.reset(seed)
    return _native(b, oracle_lib().ora_linear_episode, task_seed, steps, sigma, ticks, gain)


def maze_request_native(nr, sr, mo, seed, task_seed, mazes, horizon, ticks=16, gain=230.0):
    """replay_maze in C++: ((return, state, motors), steps)."""
    b = OracleBrain(nr, sr, mo, 9)
    b.reset(seed)
    return _native(b, oracle_lib().ora_maze_request, task_seed, mazes, horizon, ticks, gain)


def maze_layout_native(seed, index):
    rows = np.zeros(31, np.uint32)
    goal = np.zeros(2, np.uint32)
    oracle_lib().ora_maze_layout(seed, index, ptr(rows, _u32), ptr(goal, _u32))
    return [int(v) for v in rows], (int(goal[0]), int(goal[1]))
