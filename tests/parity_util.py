"""Shared helpers for the parity tests: run a workload through the device
batch and through the CPU oracle on the same seeded inputs and compare."""
import numpy as np

import oracle as O


def oracle_brains(w):
    brains = []
    seeds = w.seeds()
    for i in range(w.count):
        n, s, m = w.brain(i)
        b = O.OracleBrain(n, s, m, w.channels, max_delay=max(8, w.max_delay))
        b.reset(seeds[i])
        brains.append(b)
    return brains


def oracle_step_all(brains, x, ticks, gain, mask=None):
    outs, rcs = [], []
    for i, b in enumerate(brains):
        if mask is not None and not mask[i]:
            outs.append(np.zeros(b.outputs, np.float32))
            rcs.append(0)
            continue
        o, rc = b.step(x[i], ticks, gain)
        outs.append(o)
        rcs.append(rc)
    return np.stack(outs), np.array(rcs)


def oracle_export_all(brains):
    parts = [b.export() for b in brains]
    ns = O.cat_records([p[0] for p in parts], O.NEURON_STATE)
    ss = O.cat_records([p[1] for p in parts], O.SYNAPSE_STATE) if sum(len(p[1]) for p in parts) else np.zeros(0, O.SYNAPSE_STATE)
    h9 = np.concatenate([p[2] for p in parts])
    ticks = np.array([p[3] for p in parts], np.int64)
    return ns, ss, h9, ticks


def assert_state_equal(dev, ora, where=''):
    dns, dss, dh9, dticks = dev
    ons, oss, oh9, oticks = ora
    assert np.array_equal(dticks, oticks), f'{where}: ticks differ'
    a = O.state_bytes(dns.view(O.NEURON_STATE), O.NEURON_STATE_BYTES)
    b = O.state_bytes(ons, O.NEURON_STATE_BYTES)
    if not np.array_equal(a, b):
        bad = np.flatnonzero((a != b).any(axis=1))
        raise AssertionError(f'{where}: {len(bad)} neuron states differ, first {bad[:8]}')
    if len(oss):
        a = O.state_bytes(dss.view(O.SYNAPSE_STATE), O.SYNAPSE_STATE_BYTES)
        b = O.state_bytes(oss, O.SYNAPSE_STATE_BYTES)
        if not np.array_equal(a, b):
            bad = np.flatnonzero((a != b).any(axis=1))
            raise AssertionError(f'{where}: {len(bad)} synapse states differ, first {bad[:8]}')
    assert np.array_equal(dh9, oh9), f'{where}: history rings differ'
