"""The world step's CPU restatement (oracle ecology): its sines and cosines
against the mathematical values, and the invariants of its rules, on the CPU.
The reference computes these with CUDA fast-math intrinsics (the world build's
--use_fast_math), which this build does not reproduce: there the trigonometry
is parity unpinned, and the device is pinned to this restatement instead
(tests/test_gpu_ecology.py)."""
import numpy as np
import pytest

import ns_workloads as W
import oracle as O


def _ulps(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, -(a & 0x7fffffff), a)
    b = np.where(b < 0, -(b & 0x7fffffff), b)
    return np.abs(a - b)


def test_world_sincos_within_one_ulp():
    lib = O.oracle_lib()
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(-50, 50, 40000), rng.uniform(-1e5, 1e5, 10000),
                        np.float32([0, np.pi / 2, -np.pi / 2, np.pi, 2 * np.pi, 1e-8, -1e-30])])
    x = x.astype(np.float32)
    s = np.array([lib.ora_world_sin(float(v)) for v in x], np.float32)
    c = np.array([lib.ora_world_cos(float(v)) for v in x], np.float32)
    want_s = np.sin(x.astype(np.float64)).astype(np.float32)
    want_c = np.cos(x.astype(np.float64)).astype(np.float32)
    # away from zeros of the function a float ulp is the right unit
    big_s, big_c = np.abs(want_s) > 1e-3, np.abs(want_c) > 1e-3
    assert _ulps(s[big_s], want_s[big_s]).max() <= 1
    assert _ulps(c[big_c], want_c[big_c]).max() <= 1
    assert np.abs(s - want_s).max() < 1e-6 and np.abs(c - want_c).max() < 1e-6


def world_genomes(seed=3, animals=64, lo=20, hi=60, max_delay=8, weight=0.8):
    """Brains for the world: 19 sensory channels (every sense drives a cell),
    5 motors, weights of mean `weight` (0.8: motors fire now and then; 3.0:
    often enough that animals bite, eat poison and die within tens of steps)."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(lo, hi + 1, animals)
    return W._assemble('world', rng, sizes, 4.0, W._uniform_types([0, 1, 4, 5]), 0.2, 19, 19, 5,
                       max_delay, weight, 0.4, W._stdp_mix, 0.05, seed, 0.5, 230.0, 16, 1)


@pytest.mark.parametrize('torus', [True, False])
def test_oracle_world_step_invariants(torus):
    w = world_genomes(animals=48)
    e = O.OracleEcology(w, 20, 20, torus, 0.1, 0.02, 9)
    e2 = O.OracleEcology(w, 20, 20, torus, 0.1, 0.02, 9)
    try:
        an0, food0, kind0, _, _ = e.export()
        assert set(np.unique(food0)) <= {0, 1} and set(np.unique(kind0)) <= {0, 1, 2}
        assert (food0 > 0).mean() > 0.15 and (an0['alive'] == 1).all()
        assert (an0['energy'] == 14).all() and (an0['health'] == 10).all()
        moved, spikes = False, 0.0
        for step in range(1, 16):
            prev = e.export()[0]
            e.step(step)
            e2.step(step)
            an, food, kind, senses, motors = e.export()
            assert np.array_equal(an.view(np.uint8), e2.export()[0].view(np.uint8))
            assert np.array_equal(kind, kind0)                        # kinds never change
            assert set(np.unique(food)) <= {0, 1}
            live = prev['alive'] == 1
            assert np.array_equal(an['age'][live], prev['age'][live] + 1)
            assert (an['alive'][~live] == 0).all()                    # the dead stay dead
            assert (an['x'] >= 0).all() and (an['x'] < 20).all() and (an['y'] < 20).all()
            assert (senses >= 0).all() and (senses <= 1).all()
            moved |= not np.array_equal(an['x'], prev['x'])
            spikes += motors.sum()
            assert (an['energy'] <= 30).all()
        assert moved and spikes > 0
    finally:
        e.close()
        e2.close()
