"""The world step on the device (ns_ecology_*) against the oracle's
restatement of it, step by step and bit-exact: the animals' bodies, the food
grid, the senses, the motor counts and, at the end, every brain's state.
The worlds are small and crowded, so that harvests contend for one item,
bites land and animals die; the last case is the full 256 x 256 torus with
2048 animals of the lognormal-size population."""
import numpy as np
import pytest

import ns_amd
import ns_workloads as W
import oracle as O
from parity_util import assert_state_equal
from test_ecology import world_genomes

pytestmark = pytest.mark.gpu


def oracle_brain_states(e, w):
    parts = []
    for i, h in enumerate(e.brains):
        n, s, _ = w.brain(i)
        ns = np.zeros(len(n), O.NEURON_STATE)
        ss = np.zeros(len(s), O.SYNAPSE_STATE)
        h9 = np.zeros((len(n), 9), np.uint8)
        tick = np.zeros(1, np.int64)
        e.lib.ora_export(h, ns.ctypes.data, ss.ctypes.data, O.ptr(h9, O._u8), None,
                         O.ptr(tick, O._i64), None)
        parts.append((ns, ss, h9, int(tick[0])))
    ns = O.cat_records([p[0] for p in parts], O.NEURON_STATE)
    ss = O.cat_records([p[1] for p in parts], O.SYNAPSE_STATE)
    h9 = np.concatenate([p[2] for p in parts]) if w.max_delay <= 8 else None
    return ns, ss, h9, np.array([p[3] for p in parts], np.int64)


def assert_world_equal(dev, ora, step):
    names = ('animals', 'food', 'kinds', 'senses', 'motors')
    for name, d, o in zip(names, dev, ora):
        a = np.ascontiguousarray(d).view(np.uint8)
        b = np.ascontiguousarray(o).view(np.uint8)
        if not np.array_equal(a, b):
            if name == 'animals':
                bad = np.flatnonzero((a.reshape(len(d), -1) != b.reshape(len(o), -1)).any(1))
                raise AssertionError(f'step {step}: animals {bad[:8]} differ: {d[bad[:2]]} vs {o[bad[:2]]}')
            raise AssertionError(f'step {step}: {name} differ at {np.flatnonzero(a != b)[:8]}')


def run_world(w, width, height, torus, density, regrowth, seed, steps, instant=False):
    batch = ns_amd.GpuBatch(workload=w, world_rules=True)
    batch.reset(w.seeds())
    eco = ns_amd.Ecology(batch, width, height, torus, density, regrowth, seed, instant)
    ora = O.OracleEcology(w, width, height, torus, density, regrowth, seed, instant)
    try:
        assert_world_equal(eco.export(), ora.export(), 0)
        seen = {'alive': [], 'eaten': 0, 'pain': 0.0}
        for step in range(1, steps + 1):
            before = ora.export()[1].sum()
            eco.step(step)
            ora.step(step)
            dev, oe = eco.export(), ora.export()
            assert_world_equal(dev, oe, step)
            seen['alive'].append(int((oe[0]['alive'] == 1).sum()))
            seen['eaten'] += int(before > oe[1].sum())
            seen['pain'] = max(seen['pain'], float(oe[0]['pain'].max()))
        assert_state_equal(batch.export(), oracle_brain_states(ora, w), f'after {steps} steps')
        return seen
    finally:
        eco.close()
        ora.close()
        batch.close()


@pytest.mark.parametrize('torus,instant', [(True, False), (False, True)])
def test_crowded_world_steps_bit_exact(torus, instant):
    w = world_genomes(seed=3, animals=160)
    seen = run_world(w, 16, 16, torus, 0.1, 0.02, 9, 40, instant)
    assert seen['eaten'] > 0 and seen['pain'] > 0


def test_predation_and_deaths_bit_exact():
    w = world_genomes(seed=5, animals=200, weight=3.0)
    seen = run_world(w, 10, 10, True, 0.1, 0.02, 11, 40)
    assert seen['alive'][-1] < 150 and seen['eaten'] > 0


def test_c2_torus_world_steps_bit_exact():
    w = W.c2(brains=2048)
    seen = run_world(w, 256, 256, True, 0.05, 1e-3, 1, 3)
    assert seen['alive'][-1] == 2048
