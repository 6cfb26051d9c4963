"""Seeded synthetic genomes for the per-tick brain update workloads.

Genomes use the batched-episode encoding of gym_adapter/native.cpp:46-82:
neuron rows of 10 floats {profile, sensory channel, active, inhibitory,
modulator, capacitance x, leak x, threshold mV, receptor gain Hz, receptor
event nS} and synapse rows of 11 floats {pre, post, weight nS, active, delay,
branch, U, recovery ms, facilitation ms, plasticity, modulator receptors}.

Every generator is a pure function of its seed (numpy PCG64), so the device
path, the CPU oracle and the CPU baseline all consume identical inputs.

Workloads (sizes are the defaults; each takes overrides):
  c1  64 brains x 32 neurons, profiles {0,1,4}, 20% inhibitory, fan-out 8,
      delays 1..8, lognormal(-1, 0.5) weights, plasticity off, 6 sensors.
  c2  2048 brains, sizes lognormal(ln 96, 0.75) clipped 16..1024, 20
      synapses per neuron, delays 1..16, STP on, plasticity off.
  c3  32768 brains, Pareto(1.5, 32) sizes truncated at 4096 and rescaled to
      ~4,194,304 neurons, 2 synapses per neuron (~8,388,608), STDP profiles
      1/2 with some 3/4/5, modulator neurons.
  c4  one episode genome: log-uniform 16..4096 neurons, 8 sensors, 2 motors.
"""
from __future__ import annotations

import numpy as np

# Default threshold per profile (src/physiology.cuh:82-163): the encoding
# overrides the profile threshold, so the profile's own value is written.
PROFILE_THRESHOLD = {0: -52.0, 1: -42.0, 2: -42.0, 3: -42.0, 4: -40.0, 5: -40.0, 6: -51.8}


def _neuron_rows(rng, n, types, inhib_p, sensors, channels, modulator_p=0.0,
                 active_p=1.0):
    rows = np.zeros((n, 10), np.float32)
    rows[:, 0] = types
    rows[:, 1] = -1
    k = min(sensors, n)
    rows[:k, 1] = np.arange(k) % channels
    rows[:, 2] = (rng.random(n) < active_p).astype(np.float32)
    rows[:k, 2] = 1
    inhib = rng.random(n) < inhib_p
    rows[:, 3] = inhib
    if modulator_p > 0:
        u = rng.random(n)
        mod = np.where(u < modulator_p / 3, 1, np.where(u < 2 * modulator_p / 3, 2,
                       np.where(u < modulator_p, 3, 0)))
        mod[:k] = 0
        rows[:, 4] = mod
    rows[:, 5] = 1.0
    rows[:, 6] = 1.0
    rows[:, 7] = [PROFILE_THRESHOLD[int(t)] for t in types]
    is_orn = rows[:, 0] == 5
    rows[:, 8] = np.where(is_orn, 43.0, 0.0)
    rows[:, 9] = np.where(is_orn, 2.0, 0.0)
    return rows


def _synapse_rows(rng, n, m, max_delay, weight_mu, weight_sigma, plasticity=None,
                  self_loops=False):
    rows = np.zeros((m, 11), np.float32)
    pre = rng.integers(0, n, m)
    post = rng.integers(0, n, m)
    if not self_loops and n > 1:
        same = post == pre
        post[same] = (post[same] + 1 + rng.integers(0, n - 1, same.sum())) % n
    rows[:, 0] = pre
    rows[:, 1] = post
    rows[:, 2] = rng.lognormal(weight_mu, weight_sigma, m)
    rows[:, 3] = 1
    rows[:, 4] = rng.integers(1, max_delay + 1, m)
    rows[:, 5] = rng.integers(0, 4, m)
    rows[:, 6] = 0.55          # src/main.cu:498-501 STP defaults
    rows[:, 7] = 450.0
    rows[:, 8] = 0.0
    if plasticity is not None:
        rows[:, 9] = plasticity
        rows[:, 10] = np.where(plasticity == 3, 3, np.where(plasticity >= 4, 4, 0))
    return rows


class Workload:
    """A batch of genomes plus the per-call input generator."""

    def __init__(self, name, neuron_counts, edge_counts, neurons, synapses, motors,
                 channels, outputs, max_delay, seed, input_p, gain, ticks, calls):
        self.name = name
        self.neuron_counts = np.asarray(neuron_counts, np.int32)
        self.edge_counts = np.asarray(edge_counts, np.int32)
        self.neurons = np.ascontiguousarray(neurons, np.float32)
        self.synapses = np.ascontiguousarray(synapses, np.float32)
        self.motors = np.ascontiguousarray(motors, np.int32)
        self.channels = channels
        self.outputs = outputs
        self.max_delay = max_delay
        self.seed = seed
        self.input_p = input_p
        self.gain = gain
        self.ticks = ticks
        self.calls = calls

    @property
    def count(self):
        return len(self.neuron_counts)

    @property
    def total_neurons(self):
        return int(self.neuron_counts.sum())

    @property
    def total_synapses(self):
        return int(self.edge_counts.sum())

    def brain(self, i):
        """Slices (neuron rows, synapse rows, motors) of brain i."""
        n0 = int(self.neuron_counts[:i].sum())
        s0 = int(self.edge_counts[:i].sum())
        n, s = int(self.neuron_counts[i]), int(self.edge_counts[i])
        return (self.neurons[n0:n0 + n], self.synapses[s0:s0 + s],
                self.motors[i * self.outputs:(i + 1) * self.outputs])

    def inputs(self, call):
        """Encoded input for call `call`: Bernoulli(input_p) levels of 1.0."""
        rng = np.random.default_rng([self.seed, 7, call])
        return (rng.random((self.count, self.channels)) < self.input_p).astype(np.float32)

    def seeds(self):
        return (np.arange(self.count, dtype=np.uint64) * 2654435761 + self.seed).astype(np.uint32)


def _assemble(name, rng, sizes, fanout, types_fn, inhib_p, sensors, channels, outputs,
              max_delay, weight_mu, weight_sigma, plast_fn, modulator_p, seed, input_p,
              gain, ticks, calls, active_p=1.0):
    nrows, srows, motors, ecounts = [], [], [], []
    for n in sizes:
        n = int(n)
        types = types_fn(rng, n)
        nrows.append(_neuron_rows(rng, n, types, inhib_p, sensors, channels, modulator_p,
                                  active_p))
        m = int(round(n * fanout))
        plast = plast_fn(rng, m) if plast_fn else None
        srows.append(_synapse_rows(rng, n, m, max_delay, weight_mu, weight_sigma, plast))
        ecounts.append(m)
        motors.append(np.arange(n - outputs, n, dtype=np.int32))
    return Workload(name, sizes, ecounts, np.concatenate(nrows), np.concatenate(srows),
                    np.concatenate(motors), channels, outputs, max_delay, seed, input_p,
                    gain, ticks, calls)


def _uniform_types(choices):
    choices = np.asarray(choices)
    return lambda rng, n: choices[rng.integers(0, len(choices), n)]


POISSON_20HZ_16MS = float(1.0 - np.exp(-20.0 * 0.016))


def c1(seed=2, brains=64, neurons=32, calls=200):
    rng = np.random.default_rng(seed)
    return _assemble('c1', rng, [neurons] * brains, 8, _uniform_types([0, 1, 4]), 0.2, 6, 6, 2,
                     8, -1.0, 0.5, None, 0.0, seed, POISSON_20HZ_16MS, 230.0, 16, calls)


POISSON_20HZ_1MS = float(1.0 - np.exp(-20.0 * 0.001))


def c1_per_tick(seed=2, brains=64, neurons=32, ticks=3200):
    """C1 with its Poisson 20 Hz sensory input redrawn every 1 ms tick: the
    same genomes, one tick per call (the batched ABI takes one input per
    call), each sensor on with probability 1 - e^-0.02 per tick; 3,200
    ticks are C1's 200 steps of 16.  Homeostasis then runs after every tick,
    as the reference's gym adapter does after every call."""
    w = c1(seed, brains, neurons, calls=ticks)
    w.name = 'c1-per-tick'
    w.input_p = POISSON_20HZ_1MS
    w.ticks = 1
    return w


def _stdp_mix(rng, m):
    u = rng.random(m)
    return np.where(u < 0.05, 5, np.where(u < 0.10, 4, np.where(u < 0.20, 3,
                    np.where(u < 0.60, 2, 1)))).astype(np.float32)


def c2(seed=5, brains=2048, sigma=0.75, fanout=20, max_delay=16, calls=1000, plasticity=False):
    rng = np.random.default_rng(seed)
    sizes = np.clip(np.round(rng.lognormal(np.log(96.0), sigma, brains)), 16, 1024).astype(int)
    return _assemble('c2', rng, sizes, fanout, _uniform_types([0, 1, 4, 5]), 0.2, 8, 19, 5,
                     max_delay, -1.0, 0.5, _stdp_mix if plasticity else None, 0.0, seed,
                     POISSON_20HZ_16MS, 230.0, 16, calls)


def c3_sizes(rng, brains=32768, total=4194304, alpha=1.5, xm=32, cap=4096):
    raw = xm * (1.0 - rng.random(brains)) ** (-1.0 / alpha)
    raw = np.minimum(raw, cap)
    scaled = raw * (total / raw.sum())
    sizes = np.clip(np.floor(scaled), xm, cap).astype(np.int64)
    deficit = total - int(sizes.sum())
    if deficit > 0:                      # spread the remainder over the smallest brains
        order = np.argsort(sizes)
        sizes[order[:deficit % brains]] += 1
        sizes += deficit // brains
    return sizes


def c3(seed=3, brains=32768, total_neurons=4194304, fanout=2.0, calls=100, cap=4096):
    rng = np.random.default_rng(seed)
    sizes = c3_sizes(rng, brains, total_neurons, cap=cap)
    return _assemble('c3', rng, sizes, fanout, _uniform_types([0, 1, 4, 5]), 0.2, 8, 19, 5, 16,
                     -1.0, 0.5, _stdp_mix, 0.06, seed, POISSON_20HZ_16MS, 230.0, 16, calls)


def c3_newborns(seed=33, count=2048):
    """Newborn genomes for C3's births: the same generator and size
    distribution as c3(), for `count` brains."""
    rng = np.random.default_rng(seed)
    sizes = c3_sizes(rng, count, 128 * count)
    return _assemble('c3-births', rng, sizes, 2.0, _uniform_types([0, 1, 4, 5]), 0.2, 8, 19, 5, 16,
                     -1.0, 0.5, _stdp_mix, 0.06, seed, POISSON_20HZ_16MS, 230.0, 16, 1)


def small_mix(seed=11, brains=8, calls=40, plasticity=True, max_delay=8, types=(0, 1, 4, 5, 6),
              modulator_p=0.1, fanout=4.0, lo=8, hi=48, active_p=0.9):
    """Small ragged brains with every tick-path feature on (parity cases)."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(lo, hi + 1, brains)
    return _assemble('mix', rng, sizes, fanout, _uniform_types(list(types)), 0.25, 4, 4, 2,
                     max_delay, -0.5, 0.6, _stdp_mix if plasticity else None, modulator_p, seed,
                     0.4, 230.0, 16, calls, active_p)


def plastic_pair_counts(w):
    """Per brain, the synapses that run plasticity every tick on the gym rules:
    valid (synapse and both cells active) excitatory pairs (pre and post of an
    excitatory type 0/4/5, neither inhibitory, pre not a modulator) with a
    plasticity profile (gym_adapter/core.cuh:60-66, physiology.cuh:459-470).
    The others cost a static synapse-update."""
    n0 = np.concatenate([[0], np.cumsum(w.neuron_counts, dtype=np.int64)[:-1]])
    brain = np.repeat(np.arange(w.count), w.edge_counts)
    syn = w.synapses
    pre = syn[:, 0].astype(np.int64) + n0[brain]
    post = syn[:, 1].astype(np.int64) + n0[brain]
    nr = w.neurons
    active = nr[:, 2] > 0
    exc = np.isin(nr[:, 0], (0, 4, 5)) & (nr[:, 3] == 0)
    pair = (syn[:, 3] > 0) & active[pre] & active[post] & exc[pre] & (nr[pre, 4] == 0) & exc[post]
    return np.bincount(brain, weights=pair & (syn[:, 9] > 0), minlength=w.count).astype(np.int64)


def maze_genome(rng, neurons, fanout=4.0):
    """One maze-task genome: 9 sensor neurons on channels 0-8 (walls, goal
    direction, reward), the reward sensor a dopamine modulator, 4 motor
    neurons last; synapses with pair STDP (profiles 1/2) and D1 dopamine
    plasticity (profile 4)."""
    neurons = max(neurons, 13)
    types = _uniform_types([0, 1, 4])(rng, neurons)
    nrows = _neuron_rows(rng, neurons, types, 0.2, 9, 9)
    nrows[8, 3] = 0
    nrows[8, 4] = 3
    m = int(round(neurons * fanout))
    u = rng.random(m)
    plast = np.where(u < 0.3, 4, np.where(u < 0.55, 2, np.where(u < 0.8, 1, 0))).astype(np.float32)
    srows = _synapse_rows(rng, neurons, m, 8, -1.0, 0.5, plast)
    motors = np.arange(neurons - 4, neurons, dtype=np.int32)
    return nrows, srows, motors


def episode_genome(rng, neurons, fanout=4.0, channels=8, outputs=2):
    """One serving-episode genome (log-uniform size chosen by the caller)."""
    types = _uniform_types([0, 1, 4])(rng, neurons)
    nrows = _neuron_rows(rng, neurons, types, 0.2, channels, channels)
    m = int(round(neurons * fanout))
    srows = _synapse_rows(rng, neurons, m, 8, -1.0, 0.5)
    motors = np.arange(neurons - outputs, neurons, dtype=np.int32)
    return nrows, srows, motors
