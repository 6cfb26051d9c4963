"""The benched workloads against the oracle on the GPU, bit for bit.

* Delays up to 16 ms: the synapse arrival test reads bits 8-15 of the spike
  register, which a 9-slot ring cannot show, so spike registers are compared
  with the oracle's 17-slot rings directly.
* The benchmark's own tiles, (1024,2,0,1,4224) and (1024,1,0,1,4448), with
  brains too large for a pack (one workgroup each, state in HBM).
* The full 32,768-brain Pareto pool (4,194,304 neurons, 8,388,608 synapses,
  plasticity and modulators on) stepped as the benchmark steps it: graph
  replays on a stream, 1% of the brains replaced by newborns between steps.
  Brains are independent, so a sample (the largest, a random set, and every
  newborn) is checked against oracle brains stepped on their own.
* The full 2,048-brain lognormal pool with 20 synapses per neuron.
"""
import numpy as np
import pytest

import ns_amd
import ns_workloads as W
from oracle import OracleBrain
from parity_util import (Hip, assert_bits_equal, assert_state_equal, oracle_brains,
                         oracle_export_all, oracle_step_all, oracle_step_threaded)

pytestmark = pytest.mark.gpu

PROD_TILES = [(1024, 2, 0, 1, 4224), (1024, 1, 0, 1, 4448)]


def _compare_all(batch, brains, where):
    assert_state_equal(batch.export(), oracle_export_all(brains), where)
    assert_bits_equal(batch.spike_bits(), brains, where)


def _run(w, calls, form=ns_amd.FORM_RESIDENT, tiles=None, check_every=4):
    print(f'[benched] {w.name} brains={w.count} neurons={w.total_neurons} '
          f'synapses={w.total_synapses} max_delay={w.max_delay} form={form} tiles={tiles}',
          flush=True)
    batch = ns_amd.GpuBatch(workload=w)
    try:
        if tiles:
            batch.set_tiles(*tiles)
        batch.set_form(form)
        batch.reset(w.seeds())
        brains = oracle_brains(w)
        for call in range(calls):
            x = w.inputs(call)
            out, status = batch.step(x, w.ticks, w.gain)
            oo, orc = oracle_step_all(brains, x, w.ticks, w.gain)
            assert np.array_equal(out, oo), f'call {call}: motor outputs differ'
            assert np.array_equal(status != 0, orc != 0), f'call {call}: status differs'
            if call % check_every == check_every - 1 or call == calls - 1:
                _compare_all(batch, brains, f'call {call}')
        return batch.info()
    finally:
        batch.close()


@pytest.mark.parametrize('form,tiles', [(ns_amd.FORM_PER_TICK, None), (ns_amd.FORM_RESIDENT, None)]
                         + [(ns_amd.FORM_RESIDENT, t) for t in PROD_TILES])
def test_delays_to_16_bit_exact(form, tiles):
    w = W.small_mix(seed=51, brains=12, calls=24, max_delay=16, lo=16, hi=160, fanout=3.0)
    info = _run(w, 24, form, tiles)
    assert info['events'] > 0


def test_history9_refused_past_delay_8():
    w = W.small_mix(seed=52, brains=3, calls=1, max_delay=16)
    batch = ns_amd.GpuBatch(workload=w)
    try:
        n = w.total_neurons
        h9 = np.zeros((n, 9), np.uint8)
        rc = batch.lib.ns_gpu_import(batch.ptr, None, None, h9.ctypes.data_as(ns_amd._u8), None)
        assert rc == -1 and 'spike bit registers' in ns_amd.error()
        rc = batch.lib.ns_gpu_export(batch.ptr, None, None, h9.ctypes.data_as(ns_amd._u8), None)
        assert rc == -1 and 'spike bit registers' in ns_amd.error()
    finally:
        batch.close()


def test_spike_bits_round_trip_at_delay_16():
    """Export the whole 32-bit registers mid-run, run on, import them into a
    second batch with the other state, and both continue identically."""
    w = W.small_mix(seed=53, brains=8, calls=12, max_delay=16, lo=20, hi=90)
    a, b = ns_amd.GpuBatch(workload=w), ns_amd.GpuBatch(workload=w)
    try:
        a.reset(w.seeds())
        b.reset(w.seeds() + 1)
        for call in range(6):
            a.step(w.inputs(call), w.ticks, w.gain)
        ns, ss, h9, ticks = a.export()
        assert h9 is None
        b.import_state(ns, ss, None, ticks)
        b.import_spike_bits(a.spike_bits())
        for call in range(6, 12):
            oa, _ = a.step(w.inputs(call), w.ticks, w.gain)
            ob, _ = b.step(w.inputs(call), w.ticks, w.gain)
            assert np.array_equal(oa, ob), f'call {call}'
        ea, eb = a.export(), b.export()
        for x, y in zip(ea, eb):
            if x is not None:
                assert np.array_equal(x.view(np.uint8), y.view(np.uint8))
        assert np.array_equal(a.spike_bits(), b.spike_bits())
    finally:
        a.close()
        b.close()


def _big_mix(seed=61):
    """Brains beyond a pack of either production tile (over 2,048 or 1,024
    neurons, or over 4,224 synapses) beside small ones, plastic, modulated,
    delays up to 16."""
    rng = np.random.default_rng(seed)
    sizes = [2300, 40, 3100, 120, 1500, 64, 4096, 200, 1100, 33, 700, 2049]
    return W._assemble('big', rng, sizes, 3.0, W._uniform_types([0, 1, 4, 5, 6]), 0.2, 8, 8, 3, 16,
                       -0.7, 0.5, W._stdp_mix, 0.06, seed, 0.4, 230.0, 16, 12)


@pytest.mark.parametrize('tiles', PROD_TILES)
def test_big_brains_on_benched_tiles(tiles):
    w = _big_mix()
    info = _run(w, 12, ns_amd.FORM_RESIDENT, tiles)
    assert info['big_brains'] >= 4 and info['resident_packs'] >= 1
    assert info['events'] > 0


class _DeviceIO:
    """Per-call inputs and one output array in device memory, and a stream,
    as the benchmark holds them."""

    def __init__(self, w, patterns):
        self.hip = Hip()
        self.w = w
        self.nin = w.count * w.channels
        self.d_in = self.hip.alloc(patterns * self.nin * 4)
        self.hip.put(self.d_in, np.ascontiguousarray(
            np.stack([w.inputs(c) for c in range(patterns)]), np.float32))
        self.d_out = self.hip.alloc(w.count * w.outputs * 4)
        self.stream = self.hip.stream()

    def input(self, k):
        return self.d_in + k * self.nin * 4

    def output(self):
        self.hip.sync(self.stream)
        return self.hip.get(self.d_out, np.zeros((self.w.count, self.w.outputs), np.float32))

    def close(self):
        self.hip.sync(self.stream)
        self.hip.free(self.d_in)
        self.hip.free(self.d_out)
        assert self.hip.h.hipStreamDestroy(self.stream) == 0


def test_c3_full_pool_with_births_sampled():
    """C3 as benched: production tile, slot reserve, 1% newborns per step,
    graph replays on a stream; 3 world steps."""
    steps, tile = 3, (1024, 2, 0, 1, 4224)
    w = W.c3()
    print(f'[benched] c3 brains={w.count} neurons={w.total_neurons} synapses={w.total_synapses}',
          flush=True)
    io = _DeviceIO(w, steps)
    batch = ns_amd.GpuBatch(workload=w)
    births = None
    try:
        batch.set_tiles(*tile)
        batch.reserve(0.05, 0.03)
        batch.reset(w.seeds())
        info0 = batch.info()
        print(f'[benched] c3 packs={info0["resident_packs"]} big={info0["big_brains"]}', flush=True)
        assert info0['big_brains'] > 100
        newborn = W.c3_newborns(count=256)
        births = ns_amd.GenomeHandles([newborn.brain(i) for i in range(newborn.count)], w.channels,
                                      max_delay=w.max_delay)
        rng = np.random.default_rng(78)
        per_step = int(round(0.01 * w.count))
        order = np.argsort(w.neuron_counts)
        sample = set(order[-20:].tolist()) | set(rng.choice(w.count, 300, replace=False).tolist())
        seeds = w.seeds()
        ora = {}
        for i in sorted(sample):
            n, s, m = w.brain(i)
            ora[i] = OracleBrain(n, s, m, w.channels, max_delay=w.max_delay)
            ora[i].reset(seeds[i])
        genome = {i: w.brain(i) for i in sample}
        batch.set_graph(True)

        def births_for(k):
            idx = np.ascontiguousarray(rng.choice(w.count, per_step, replace=False), np.int32)
            which = rng.integers(0, newborn.count, per_step)
            sd = np.ascontiguousarray(rng.integers(1, 2**32 - 1, per_step), np.uint32)
            batch.replace(idx, births, which, sd, stream=io.stream)
            for i, j, s in zip(idx.tolist(), which.tolist(), sd.tolist()):
                genome[i] = newborn.brain(j)
                n, syn, m = genome[i]
                if i in ora:
                    ora[i].close()
                ora[i] = OracleBrain(n, syn, m, w.channels, max_delay=w.max_delay)
                ora[i].reset(s)

        births_for(0)
        for k in range(steps):
            batch.step_device(io.input(k), w.ticks, w.gain, io.d_out, stream=io.stream)
            dev_out = io.output()
            x = w.inputs(k)
            for i, b in ora.items():
                o, _ = b.step(x[i], w.ticks, w.gain)
                assert np.array_equal(dev_out[i], o), f'step {k}: brain {i} motor outputs differ'
            print(f'[benched] c3 step {k}: {len(ora)} sampled brains agree', flush=True)
            if k + 1 < steps:
                births_for(k + 1)
        info = batch.info()
        assert info['placements'] + info['relayouts'] > 0
        ns, ss, _, ticks = batch.export()
        bits = batch.spike_bits()
        n_off = np.concatenate([[0], np.cumsum(batch.neuron_counts, dtype=np.int64)])
        s_off = np.concatenate([[0], np.cumsum(batch.edge_counts, dtype=np.int64)])
        for i, b in sorted(ora.items()):
            dn = ns[n_off[i]:n_off[i + 1]]
            ds = ss[s_off[i]:s_off[i + 1]]
            assert_state_equal((dn, ds, None, ticks[i:i + 1]),
                               tuple(x for x in oracle_export_all([b])), f'brain {i}')
            assert_bits_equal(bits[n_off[i]:n_off[i + 1]], [b], f'brain {i}')
        print(f'[benched] c3 final state of {len(ora)} brains agrees', flush=True)
    finally:
        batch.close()
        io.close()
        if births is not None:
            births.close()


@pytest.mark.parametrize('tiles', PROD_TILES)
def test_c2_full_pool(tiles):
    """C2 as benched (2,048 brains, 20 synapses per neuron, delays 1-16, STP
    on), 6 graph-replayed calls against the threaded oracle over every brain."""
    calls = 6
    w = W.c2(calls=calls)
    print(f'[benched] c2 brains={w.count} neurons={w.total_neurons} synapses={w.total_synapses} '
          f'tiles={tiles}', flush=True)
    io = _DeviceIO(w, calls)
    batch = ns_amd.GpuBatch(workload=w)
    try:
        batch.set_tiles(*tiles)
        batch.reset(w.seeds())
        batch.set_graph(True)
        brains = oracle_brains(w)
        for call in range(calls):
            batch.step_device(io.input(call), w.ticks, w.gain, io.d_out, stream=io.stream)
            oo, _ = oracle_step_threaded(brains, w.inputs(call), w.ticks, w.gain)
            assert np.array_equal(io.output(), oo), f'call {call}: motor outputs differ'
        info = batch.info()
        assert info['big_brains'] > 0 and info['events'] > 0
        _compare_all(batch, brains, 'c2 end')
    finally:
        batch.close()
        io.close()
