"""Pins of the oracle's world-rule harness (src/main.cu's neural step) on the
CPU.

* Reduction: with every world-only rule neutral (no cortical type-0 cells,
  so the gym path's per-call homeostasis does nothing; every cell active;
  receptor sensors on channels below 6), a brain built from its gene records
  and ports under the world rules (ora_create_world) must follow the
  reference's own gym path byte for byte: against the committed fixture
  written from oracle/_ref (tests/golden/world_reduction.npz), and live
  against oracle/_ref where it is built.
* Closed forms for what the reduction leaves neutral: an inactive cell keeps
  its ring slot and state (main.cu:717-718), a type-5 cell on a channel of 6
  or more takes 230 pA per level and leaves its receptor generator alone
  (main.cu:733-736), a frozen animal's synapses learn nothing while its cells
  run (main.cu:142-146, 750), a modulator release at tick t stamps t + 1 in
  its parity slot (main.cu:669-673), and world homeostasis measures the rate
  over max(16, 16 age) ms (main.cu:1287-1300).
"""
import hashlib
from pathlib import Path

import numpy as np
import pytest

import ns_workloads as W
import oracle as O

GOLDEN = Path(__file__).resolve().parent / 'golden' / 'world_reduction.npz'
SENSES, MOTOR_PORT = 19, 19
# The cases of tests/golden/make_golden.py's WORLD_CASES.
CASES = {
    'adex': dict(seed=81, brains=8, calls=30, types=(1, 4, 5, 6), active_p=1.0, modulator_p=0.15,
                 lo=12, hi=60),
    'slow': dict(seed=82, brains=3, calls=8, types=(1, 2, 3, 4, 5), active_p=1.0, modulator_p=0.15,
                 lo=8, hi=24),
}


class WorldBrain:
    """One animal's brain under the world rules: gene records from the gym
    parse of (neuron rows, synapse rows, motors), the row's sensory channel
    in the gene, motor j on port 19 + j."""

    def __init__(self, nrows, srows, motors, seed, max_delay=8, channels=None):
        self.lib = O.oracle_lib()
        self.gym = O.OracleBrain(nrows, srows, motors, SENSES, max_delay=max_delay)
        self.gym.reset(seed)
        ng, sg = self.gym.genes()
        ng['sensory_channel'] = np.asarray(nrows, np.float32)[:, 1].astype(np.int8)
        ports = np.full(len(nrows), -1, np.int32)
        for j, m in enumerate(motors):
            ports[m] = MOTOR_PORT + j
        self.ng, self.sg, self.ports = ng, sg, ports
        self.count, self.edges = len(ng), len(sg)
        self.h = self.lib.ora_create_world(len(ng), len(sg), ng.ctypes.data, sg.ctypes.data,
                                           ports.ctypes.data_as(O._i32), max_delay)
        ns, ss, _, _ = self.gym.export()
        self.lib.ora_import(self.h, ns.ctypes.data, ss.ctypes.data, None, 0)

    def step(self, levels, ticks=16):
        senses = np.zeros(SENSES, np.float32)
        senses[:len(levels)] = levels
        out = np.zeros(5, np.float32)
        rc = self.lib.ora_step(self.h, senses.ctypes.data_as(O._f), ticks, 0.0,
                               out.ctypes.data_as(O._f), None)
        return out, rc

    def export(self, ring=False):
        ns = np.zeros(self.count, O.NEURON_STATE)
        ss = np.zeros(self.edges, O.SYNAPSE_STATE)
        h9 = np.zeros((self.count, 9), np.uint8)
        tick = np.zeros(1, np.int64)
        stamps = np.zeros(self.count * 6, np.uint64)
        R = self.lib.ora_ring(self.h)
        raw = np.zeros((self.count, R), np.uint8)
        self.lib.ora_export(self.h, ns.ctypes.data, ss.ctypes.data, h9.ctypes.data_as(O._u8),
                            raw.ctypes.data_as(O._u8), tick.ctypes.data_as(O._i64),
                            stamps.ctypes.data_as(O._u64))
        return ns, ss, (raw if ring else h9), int(tick[0]), stamps.reshape(-1, 6)

    def import_state(self, ns, ss, ring, tick):
        self.lib.ora_import(self.h, ns.ctypes.data, ss.ctypes.data,
                            None if ring is None else ring.ctypes.data_as(O._u8), tick)

    def close(self):
        if self.h:
            self.lib.ora_destroy(self.h)
            self.h = None
        self.gym.close()


def _digest(arr, mask=None):
    b = O.state_bytes(arr, mask) if mask is not None else np.ascontiguousarray(arr).view(np.uint8)
    return np.frombuffer(hashlib.sha256(b.tobytes()).digest(), np.uint8)


def _world_brains(w):
    seeds = w.seeds()
    return [WorldBrain(*w.brain(i), seeds[i]) for i in range(w.count)]


@pytest.mark.parametrize('case', list(CASES))
def test_world_harness_reduces_to_the_reference_gym_path(case):
    kw = CASES[case]
    fx = np.load(GOLDEN)
    w = W.small_mix(**kw)
    assert not (w.neurons[:, 0] == 0).any() and (w.neurons[:, 2] == 1).all()
    assert (w.neurons[:, 1] < 6).all()
    brains = _world_brains(w)
    try:
        for c in range(kw['calls']):
            x = w.inputs(c)
            for i, b in enumerate(brains):
                out, rc = b.step(x[i])
                assert rc == 0
                assert np.array_equal(out[:w.outputs], fx[f'{case}_outputs'][c, i]), (c, i)
                assert not out[w.outputs:].any()
            parts = [b.export() for b in brains]
            ns = O.cat_records([p[0] for p in parts], O.NEURON_STATE)
            ss = O.cat_records([p[1] for p in parts], O.SYNAPSE_STATE)
            h9 = np.concatenate([p[2] for p in parts])
            h = fx[f'{case}_hashes'][c]
            assert np.array_equal(_digest(ns, O.NEURON_STATE_BYTES), h[0]), f'call {c}: neurons'
            assert np.array_equal(_digest(ss, O.SYNAPSE_STATE_BYTES), h[1]), f'call {c}: synapses'
            assert np.array_equal(_digest(h9), h[2]), f'call {c}: rings'
        assert np.array_equal(O.state_bytes(ns, O.NEURON_STATE_BYTES),
                              O.state_bytes(fx[f'{case}_final_neurons'].view(O.NEURON_STATE),
                                            O.NEURON_STATE_BYTES))
        assert np.array_equal(h9, fx[f'{case}_final_history'])
    finally:
        for b in brains:
            b.close()


@pytest.mark.skipif(not O.ref_available(), reason='oracle/_ref (reference checkout) not built')
def test_world_harness_matches_reference_live():
    """Fresh seeds and genomes through oracle/_ref's ns_step and the world
    harness side by side, every state byte every call."""
    for seed in (83, 84):
        w = W.small_mix(seed=seed, brains=6, calls=12, types=(1, 4, 5, 6), active_p=1.0,
                        modulator_p=0.2, lo=10, hi=50)
        seeds = w.seeds()
        world = _world_brains(w)
        ref = []
        for i in range(w.count):
            r = O.RefBrain(*w.brain(i), w.channels)
            r.reset(seeds[i])
            ref.append(r)
        try:
            for c in range(w.calls):
                x = w.inputs(c)
                for i in range(w.count):
                    o_ref, _ = ref[i].step(x[i], 16, 230.0)
                    o_w, _ = world[i].step(x[i])
                    assert np.array_equal(o_ref, o_w[:w.outputs]), (seed, c, i)
                    a, b = ref[i].export(), world[i].export()
                    assert np.array_equal(O.state_bytes(a[0], O.NEURON_STATE_BYTES),
                                          O.state_bytes(b[0], O.NEURON_STATE_BYTES)), (seed, c, i)
                    assert np.array_equal(O.state_bytes(a[1], O.SYNAPSE_STATE_BYTES),
                                          O.state_bytes(b[1], O.SYNAPSE_STATE_BYTES)), (seed, c, i)
                    assert np.array_equal(a[2], b[2]) and a[3] == b[3], (seed, c, i)
        finally:
            for b in world + ref:
                b.close()


def _rows(types, channels, active=None, inhibitory=None, modulator=None):
    n = len(types)
    rows = np.zeros((n, 10), np.float32)
    rows[:, 0] = types
    rows[:, 1] = channels
    rows[:, 2] = 1 if active is None else active
    rows[:, 3] = 0 if inhibitory is None else inhibitory
    rows[:, 4] = 0 if modulator is None else modulator
    rows[:, 5:7] = 1
    rows[:, 7] = [W.PROFILE_THRESHOLD[int(t)] for t in types]
    rows[:, 8] = np.where(rows[:, 0] == 5, 43.0, 0.0)
    rows[:, 9] = np.where(rows[:, 0] == 5, 2.0, 0.0)
    return rows


def _syn(pre, post, w, delay, plasticity=0, u=0.55):
    return [pre, post, w, 1, delay, 0, u, 450.0, 0.0, plasticity, 0]


def test_inactive_cell_keeps_its_ring_slot_and_state():
    """main.cu:717-718: an inactive cell returns before touching its state or
    its ring slot, so both survive every tick; the gym path (core.cuh:56-57)
    clears its spike and writes 0 into its slot instead."""
    rows = _rows([4, 4, 4], [0, -1, -1], active=[1, 0, 1])
    syn = np.array([_syn(0, 2, 2.0, 2), _syn(1, 2, 2.0, 3)], np.float32)
    b = WorldBrain(rows, syn, [2], seed=5)
    try:
        ns, ss, ring, _, _ = b.export(ring=True)
        ring[1] = [1, 0, 1, 1, 0, 0, 1, 0, 1]
        ns['spike'][1] = 1
        ns['f'][1, 0] = -61.25
        b.import_state(ns, ss, ring, 40)
        before = (O.state_bytes(ns[1:2], O.NEURON_STATE_BYTES), ring[1].copy())
        for _ in range(3):
            b.step(np.array([1.0], np.float32), ticks=6)
        ns2, _, ring2, tick, _ = b.export(ring=True)
        assert tick == 58
        assert np.array_equal(O.state_bytes(ns2[1:2], O.NEURON_STATE_BYTES), before[0])
        assert np.array_equal(ring2[1], before[1])
        assert ns2['total_spikes'][0] > 0
        # the gym rules on the same cell: spike cleared, zeros written
        g = O.OracleBrain(rows, syn, [2], SENSES)
        g.reset(5)
        g.lib.ora_import(g.ptr, ns.ctypes.data, ss.ctypes.data, ring.ctypes.data_as(O._u8), 40)
        x = np.zeros(SENSES, np.float32)
        x[0] = 1
        g.step(x, 18, 230.0)
        gn, _, gh9, _ = g.export()
        assert gn['spike'][1] == 0 and not gh9[1].any()
        g.close()
    finally:
        b.close()


def test_receptor_path_only_below_channel_6():
    """main.cu:733-736: a type-5 cell on channel 7 takes level x 230 pA (its
    receptor generator does not move); on channel 3 it takes receptor_current."""
    level = np.float32(0.37)
    lib = O.oracle_lib()
    for ch, receptor in ((7, False), (3, True)):
        rows = _rows([5, 4], [ch, -1])
        b = WorldBrain(rows, np.zeros((0, 11), np.float32), [1], seed=9)
        try:
            ns0, _, _, _, _ = b.export()
            levels = np.zeros(8, np.float32)
            levels[ch] = level
            b.step(levels, ticks=1)
            ns1, _, _, _, _ = b.export()
            g = np.ascontiguousarray(b.ng[0:1])
            s = np.ascontiguousarray(ns0[0:1].copy())
            if receptor:
                cur = lib.ora_fn_receptor_current(g.ctypes.data, s.ctypes.data, level)
            else:
                cur = np.float32(level * np.float32(230.0))
            exc = np.zeros(4, np.float32)
            lib.ora_fn_neuron_advance(g.ctypes.data, s.ctypes.data, cur, exc.ctypes.data_as(O._f), 0.0)
            assert np.array_equal(O.state_bytes(ns1[0:1], O.NEURON_STATE_BYTES),
                                  O.state_bytes(s, O.NEURON_STATE_BYTES)), ch
            moved = ns1['receptor_rng'][0] != ns0['receptor_rng'][0]
            assert moved == receptor, ch
        finally:
            b.close()


def test_frozen_animal_learns_nothing_while_its_cells_run():
    """main.cu:142-146, 750: plasticity of a frozen animal returns at once;
    delivery (and its short-term release state) and the cells still run."""
    types = [4, 4, 4, 4, 0, 0]
    rows = _rows(types, [0, 1, -1, -1, -1, -1])
    rng = np.random.default_rng(3)
    syn = np.array([_syn(int(a), int(c), 3.0, int(d), plasticity=int(p))
                    for a, c, d, p in zip(rng.integers(0, 6, 30), rng.integers(0, 6, 30),
                                          rng.integers(1, 8, 30), rng.integers(1, 3, 30)) if a != c],
                   np.float32)
    results = {}
    for frozen in (0, 1):
        b = WorldBrain(rows, syn, [5], seed=2)
        try:
            b.lib.ora_set_frozen(b.h, frozen)
            ns0, ss0, _, _, _ = b.export()
            for _ in range(4):
                b.step(np.array([1.0, 1.0], np.float32))
            ns1, ss1, _, _, _ = b.export()
            results[frozen] = (ns0, ss0, ns1, ss1)
        finally:
            b.close()
    # every state float but running_utilization (field 1, short-term release):
    # weight, traces, eligibility, dopamine and structural ages, spine, ages
    learned = [0, 2, 3, 4, 5, 6, 7, 8, 9, 10]
    ns0, ss0, ns1, ss1 = results[1]
    assert np.array_equal(ss1['f'][:, learned], ss0['f'][:, learned])
    assert (ss1['last_arrival_tick'] > 0).any()                 # deliveries happened
    assert (ns1['total_spikes'] > 0).any() and not np.array_equal(ns1['f'], ns0['f'])
    _, uss0, _, uss1 = results[0]
    assert not np.array_equal(uss1['f'][:, learned], uss0['f'][:, learned])   # unfrozen: they move


def test_modulator_release_stamps_the_next_tick_in_its_parity_slot():
    """main.cu:669-673: a modulator release at tick t raises
    mod_event_tick[post][(m - 1) * 2 + (t & 1)] to t + 1.  U = 1 with every
    release site ready makes the first release certain."""
    rows = _rows([4, 4, 4, 4], [-1, -1, -1, -1], modulator=[3, 1, 0, 0])
    syn = np.array([_syn(0, 2, 1.5, 3, u=1.0), _syn(1, 2, 1.5, 2, u=1.0), _syn(1, 3, 1.5, 5, u=1.0)],
                   np.float32)
    b = WorldBrain(rows, syn, [3], seed=4)
    try:
        ns, ss, ring, _, _ = b.export(ring=True)
        ring[0, 99 % 9] = 1          # DA cell spiked at tick 99: arrival at 102 on cell 2
        ring[1, 99 % 9] = 1          # NE cell spiked at tick 99: arrivals at 101 (cell 2), 104 (cell 3)
        b.import_state(ns, ss, ring, 100)
        b.step(np.zeros(1, np.float32), ticks=5)
        _, ss1, _, _, st = b.export()
        expect = np.zeros((4, 6), np.uint64)
        expect[2, (3 - 1) * 2 + (102 & 1)] = 103
        expect[2, (1 - 1) * 2 + (101 & 1)] = 102
        expect[3, (1 - 1) * 2 + (104 & 1)] = 105
        assert np.array_equal(st, expect), st
        assert list(ss1['last_arrival_tick']) == [102, 101, 104]
    finally:
        b.close()


def test_world_homeostasis_rate_over_sixteen_ms_per_step_of_age():
    """main.cu:1287-1300: elapsed = max(16, 16 age) ms, one 16 ms advance,
    skipped for frozen animals and inactive cells."""
    lib = O.oracle_lib()
    rows = _rows([0, 0, 4], [-1, -1, -1], active=[1, 0, 1])
    for age in (0, 1, 7):
        for frozen in (0, 1):
            b = WorldBrain(rows, np.zeros((0, 11), np.float32), [2], seed=1)
            try:
                ns, ss, ring, _, _ = b.export(ring=True)
                ns['total_spikes'] = [5, 5, 5]
                ns['f'][:, 5] = [0.1, 0.1, 0.1]          # log AMPA scale
                b.import_state(ns, ss, ring, 0)
                b.lib.ora_set_frozen(b.h, frozen)
                lib.ora_world_homeostasis(b.h, age)
                got, _, _, _, _ = b.export()
                s = np.ascontiguousarray(ns[0:1].copy())
                g = np.ascontiguousarray(b.ng[0:1])
                elapsed = np.float32(max(16.0, age * 16.0))
                lib.ora_fn_homeostasis(g.ctypes.data, s.ctypes.data, 16.0,
                                       np.float32(np.float32(5 * 1000.0) / elapsed))
                want = ns['f'][0, 5] if frozen else s['f'][0, 5]
                assert got['f'][0, 5] == want, (age, frozen)
                assert got['f'][1, 5] == np.float32(0.1)            # inactive
                assert frozen or got['f'][0, 5] != np.float32(0.1)
            finally:
                b.close()
