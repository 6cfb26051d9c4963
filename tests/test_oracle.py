"""The CPU oracle is pinned against the reference's own CPU path:
(1) golden vectors written by tests/golden/make_golden.py from the
reference's host-compiled sources, (2) the reference library itself when
oracle/_ref is present, (3) closed forms the mathematics fixes."""
import hashlib
from pathlib import Path

import numpy as np
import pytest

import oracle as O
import ns_workloads as W
from parity_util import oracle_brains, oracle_export_all, oracle_step_all

GOLDEN = Path(__file__).resolve().parent / 'golden'


def _digest(arr, mask=None):
    b = O.state_bytes(arr, mask) if mask is not None else np.ascontiguousarray(arr).view(np.uint8)
    return np.frombuffer(hashlib.sha256(b.tobytes()).digest(), np.uint8)


def _check_against_fixture(w, fx, calls):
    brains = oracle_brains(w)
    for c in range(calls):
        out, rc = oracle_step_all(brains, w.inputs(c), w.ticks, w.gain)
        assert np.array_equal(out.astype(np.uint8), fx['outputs'][c]), f'call {c}: outputs'
        ns, ss, h9, _ = oracle_export_all(brains)
        assert np.array_equal(_digest(ns, O.NEURON_STATE_BYTES), fx['hashes'][c, 0]), f'call {c}: neurons'
        assert np.array_equal(_digest(ss, O.SYNAPSE_STATE_BYTES), fx['hashes'][c, 1]), f'call {c}: synapses'
        assert np.array_equal(_digest(h9), fx['hashes'][c, 2]), f'call {c}: history'
    return brains


def test_oracle_matches_c1_golden():
    fx = np.load(GOLDEN / 'c1_reference.npz')
    brains = _check_against_fixture(W.c1(calls=200), fx, 200)
    k = len(fx['final_ticks'])
    ns, ss, h9, ticks = oracle_export_all(brains[:k])
    assert np.array_equal(O.state_bytes(ns, O.NEURON_STATE_BYTES),
                          O.state_bytes(fx['final_neurons'].view(O.NEURON_STATE), O.NEURON_STATE_BYTES))
    assert np.array_equal(h9, fx['final_history']) and np.array_equal(ticks, fx['final_ticks'])


def test_oracle_matches_mix_golden():
    fx = np.load(GOLDEN / 'mix_reference.npz')
    _check_against_fixture(W.small_mix(seed=12, brains=8, calls=40, types=(0, 1, 2, 3, 4, 5, 6)), fx, 40)


def test_oracle_matches_per_tick_raster():
    fx = np.load(GOLDEN / 'c1_ticks.npz')
    raster = np.unpackbits(fx['raster'], axis=1)
    w = W.c1(calls=320, brains=8)
    brains = oracle_brains(w)
    for t in range(320):
        x = w.inputs(t)
        row = []
        for i, b in enumerate(brains):
            _, rc, r = b.step(x[i], 1, w.gain, raster=True)
            row.append(r[0])
        assert np.array_equal(np.concatenate(row), raster[t, :w.total_neurons]), f'tick {t}'


def _neuron_records(raw):
    return raw.view(O.NEURON_STATE)


def test_oracle_functions_match_golden_vectors():
    fx = np.load(GOLDEN / 'functions.npz')
    lib = O.oracle_lib()
    genes = fx['profiles'].view(O.NEURON_GENE)
    for t in range(7):
        g = np.zeros(1, O.NEURON_GENE)
        lib.ora_fn_profile(t, g.ctypes.data)
        assert np.array_equal(O.state_bytes(g, O.NEURON_GENE_BYTES),
                              O.state_bytes(genes[t:t + 1], O.NEURON_GENE_BYTES)), f'profile {t}'
    init = fx['init_neurons'].view(O.NEURON_STATE)
    for k, seed in enumerate(fx['init_neuron_seeds']):
        g = genes[k // 8:k // 8 + 1].copy()
        g['sensory_channel'] = 0
        s = np.zeros(1, O.NEURON_STATE)
        lib.ora_fn_initial_neuron(g.ctypes.data, int(seed), s.ctypes.data)
        assert np.array_equal(O.state_bytes(s, O.NEURON_STATE_BYTES),
                              O.state_bytes(init[k:k + 1], O.NEURON_STATE_BYTES)), f'init neuron {k}'
    before = fx['adv_before'].view(O.NEURON_STATE)
    after = fx['adv_after'].view(O.NEURON_STATE)
    for k, t in enumerate(fx['adv_type']):
        g = genes[t:t + 1].copy()
        g['sensory_channel'] = 0
        s = before[k:k + 1].copy()
        exc = np.ascontiguousarray(fx['adv_exc'][k])
        lib.ora_fn_neuron_advance(g.ctypes.data, s.ctypes.data, float(fx['adv_current'][k]),
                                  exc.ctypes.data_as(O._f), float(fx['adv_inh'][k]))
        assert np.array_equal(O.state_bytes(s, O.NEURON_STATE_BYTES),
                              O.state_bytes(after[k:k + 1], O.NEURON_STATE_BYTES)), f'advance {k} type {t}'
    g5 = fx['rc_gene'].view(O.NEURON_GENE).copy()
    rb, ra = fx['rc_before'].view(O.NEURON_STATE), fx['rc_after'].view(O.NEURON_STATE)
    for k in range(len(rb)):
        s = rb[k:k + 1].copy()
        v = lib.ora_fn_receptor_current(g5.ctypes.data, s.ctypes.data, float(fx['rc_level'][k]))
        assert np.float32(v) == fx['rc_value'][k]
        assert np.array_equal(O.state_bytes(s, O.NEURON_STATE_BYTES), O.state_bytes(ra[k:k + 1], O.NEURON_STATE_BYTES))
    sg = fx['rel_genes'].view(O.SYNAPSE_GENE)
    rb, ra = fx['rel_before'].view(O.SYNAPSE_STATE), fx['rel_after'].view(O.SYNAPSE_STATE)
    for k in range(len(sg)):
        s = rb[k:k + 1].copy()
        v = lib.ora_fn_synapse_release(sg[k:k + 1].ctypes.data, s.ctypes.data, int(fx['rel_tick'][k]))
        assert np.float32(v) == fx['rel_value'][k], f'release {k}'
        assert np.array_equal(O.state_bytes(s, O.SYNAPSE_STATE_BYTES), O.state_bytes(ra[k:k + 1], O.SYNAPSE_STATE_BYTES))
    pg = fx['pl_genes'].view(O.SYNAPSE_GENE)
    pb, pa = fx['pl_before'].view(O.SYNAPSE_STATE), fx['pl_after'].view(O.SYNAPSE_STATE)
    for k in range(len(pg)):
        s = pb[k:k + 1].copy()
        f = fx['pl_flags'][k]
        lib.ora_fn_plasticity(pg[k:k + 1].ctypes.data, s.ctypes.data, int(f[0]), int(f[1]), int(f[2]),
                              float(fx['pl_depol'][k]), int(f[3]))
        assert np.array_equal(O.state_bytes(s, O.SYNAPSE_STATE_BYTES),
                              O.state_bytes(pa[k:k + 1], O.SYNAPSE_STATE_BYTES)), f'plasticity {k}'
    hb, ha = fx['ho_before'].view(O.NEURON_STATE), fx['ho_after'].view(O.NEURON_STATE)
    for k in range(len(hb)):
        s = hb[k:k + 1].copy()
        t = int(fx['ho_type'][k])
        lib.ora_fn_homeostasis(genes[t:t + 1].ctypes.data, s.ctypes.data, float(fx['ho_elapsed'][k]),
                               float(fx['ho_hz'][k]))
        assert np.array_equal(O.state_bytes(s, O.NEURON_STATE_BYTES), O.state_bytes(ha[k:k + 1], O.NEURON_STATE_BYTES))
    si = fx['init_synapses'].view(O.SYNAPSE_STATE)
    for k, seed in enumerate(fx['init_syn_seeds']):
        s = np.zeros(1, O.SYNAPSE_STATE)
        lib.ora_fn_initial_synapse(sg[k:k + 1].ctypes.data, int(seed), s.ctypes.data)
        assert np.array_equal(O.state_bytes(s, O.SYNAPSE_STATE_BYTES), O.state_bytes(si[k:k + 1], O.SYNAPSE_STATE_BYTES))


@pytest.mark.skipif(not O.ref_available(), reason='oracle/_ref not built (needs the reference sources)')
@pytest.mark.parametrize('seed,types', [(41, (0, 1, 4, 5, 6)), (42, (0, 1, 2, 3, 4, 5, 6))])
def test_oracle_matches_reference_library_live(seed, types):
    w = W.small_mix(seed=seed, brains=6, calls=25, types=types, modulator_p=0.2)
    seeds = w.seeds()
    ora, ref = oracle_brains(w), []
    for i in range(w.count):
        n, s, m = w.brain(i)
        b = O.RefBrain(n, s, m, w.channels)
        b.reset(seeds[i])
        ref.append(b)
    for c in range(w.calls):
        x = w.inputs(c)
        for i in range(w.count):
            o1, r1 = ora[i].step(x[i], w.ticks, w.gain)
            o2, r2 = ref[i].step(x[i], w.ticks, w.gain)
            assert r1 == r2 and np.array_equal(o1, o2)
            a, b = ora[i].export(), ref[i].export()
            assert np.array_equal(O.state_bytes(a[0], O.NEURON_STATE_BYTES), O.state_bytes(b[0], O.NEURON_STATE_BYTES))
            assert np.array_equal(O.state_bytes(a[1], O.SYNAPSE_STATE_BYTES), O.state_bytes(b[1], O.SYNAPSE_STATE_BYTES))
            assert np.array_equal(a[2], b[2]) and a[3] == b[3]


# ---- closed forms ----------------------------------------------------------

def test_xorshift_seed_mix_closed_form():
    """initial_synapse_state's rng = the 32-bit hash of physiology.cuh:395-399."""
    lib = O.oracle_lib()
    g = np.zeros(1, O.SYNAPSE_GENE)
    g['release_sites'] = 14
    for seed in [0, 1, 2, 12345, 2**31 + 7, 2**32 - 1]:
        s = np.zeros(1, O.SYNAPSE_STATE)
        lib.ora_fn_initial_synapse(g.ctypes.data, seed, s.ctypes.data)
        x = seed or 1
        x ^= x >> 16; x = (x * 0x7feb352d) & 0xffffffff
        x ^= x >> 15; x = (x * 0x846ca68b) & 0xffffffff
        x ^= x >> 16
        assert int(s['rng_state'][0]) == (x or 1)
        assert int(s['ready_mask'][0]) == (1 << 14) - 1


@pytest.mark.parametrize('sites,mask', [(0, (1 << 14) - 1), (1, 1), (63, (1 << 63) - 1),
                                        (64, 2**64 - 1), (200, 2**64 - 1)])
def test_ready_mask_site_count(sites, mask):
    lib = O.oracle_lib()
    g = np.zeros(1, O.SYNAPSE_GENE)
    g['release_sites'] = min(sites, 255)
    s = np.zeros(1, O.SYNAPSE_STATE)
    lib.ora_fn_initial_synapse(g.ctypes.data, 5, s.ctypes.data)
    assert int(s['ready_mask'][0]) == mask


def test_release_with_all_sites_ready_and_certain_release():
    """U = 1 and all sites ready: every site releases, output = w * slow."""
    lib = O.oracle_lib()
    g = np.zeros(1, O.SYNAPSE_GENE)
    g['initial_weight_ns'] = 1.5; g['utilization'] = 1.0; g['recovery_ms'] = 100.0
    g['release_sites'] = 10
    s = np.zeros(1, O.SYNAPSE_STATE)
    lib.ora_fn_initial_synapse(g.ctypes.data, 9, s.ctypes.data)
    out = lib.ora_fn_synapse_release(g.ctypes.data, s.ctypes.data, 50)
    assert np.float32(out) == np.float32(1.5) and int(s['ready_mask'][0]) == 0
    assert int(s['last_arrival_tick'][0]) == 50


def test_pre_trace_decays_geometrically():
    """Profile 1 with no events: pre_trace(k) = pre_trace(0) * c^k, c = expf(-1/17)."""
    lib = O.oracle_lib()
    g = np.zeros(1, O.SYNAPSE_GENE)
    g['plasticity'] = 1; g['initial_weight_ns'] = 1.0
    s = np.zeros(1, O.SYNAPSE_STATE)
    lib.ora_fn_initial_synapse(g.ctypes.data, 3, s.ctypes.data)
    s['f'][0, 2] = 1.0
    c = np.float32(lib.ora_fn_expf(np.float32(-1.0) / np.float32(17.0)))
    want = np.float32(1.0)
    for k in range(50):
        lib.ora_fn_plasticity(g.ctypes.data, s.ctypes.data, 0, 0, 1, 0.0, 0)
        want = np.float32(want * c)
        assert s['f'][0, 2] == want
    assert s['f'][0, 9] == np.float32(1000050)


def test_weights_stay_in_bounds_under_pairing():
    """Weight bound [0, 2 w0] (physiology.cuh:506,534,554) for every profile."""
    lib = O.oracle_lib()
    rng = np.random.default_rng(5)
    for prof in range(1, 6):
        g = np.zeros(1, O.SYNAPSE_GENE)
        g['plasticity'] = prof; g['initial_weight_ns'] = 0.8; g['mod_receptors'] = 7
        s = np.zeros(1, O.SYNAPSE_STATE)
        lib.ora_fn_initial_synapse(g.ctypes.data, 3, s.ctypes.data)
        for _ in range(3000):
            lib.ora_fn_plasticity(g.ctypes.data, s.ctypes.data, int(rng.random() < .3),
                                  int(rng.random() < .3), 1, 3.0, int(rng.integers(0, 8)))
            assert 0.0 <= s['f'][0, 0] <= np.float32(1.6)


def test_delay_ring_delivers_exactly_delay_ticks_later():
    """A spike at tick t arrives at its synapse exactly `delay` ticks later,
    for delays up to the extended ring (32)."""
    for d in (1, 5, 8, 9, 16, 32):
        neurons = np.array([[4, 0, 1, 0, 0, 1, 1, -40, 0, 0], [4, -1, 1, 0, 0, 1, 1, -40, 0, 0]], np.float32)
        syn = np.array([[0, 1, 50.0, 1, d, 0, 1.0, 450, 0, 0, 0]], np.float32)
        syn[0, 6] = 1.0
        b = O.OracleBrain(neurons, syn, [1], 1, max_delay=max(8, d))
        b.reset(1)
        first_spike = first_arrival = None
        for t in range(80):
            _, rc, r = b.step(np.array([1.0 if t < 30 else 0.0], np.float32), 1, 3000.0, raster=True)
            ss = b.export()[1]
            if r[0, 0] and first_spike is None:
                first_spike = t
            if ss['arrived'][0] and first_arrival is None:
                first_arrival = t
        assert first_spike is not None and first_arrival == first_spike + d
