"""Device tick path vs the CPU oracle, bit-exact, through the C ABI."""
import numpy as np
import pytest

import ns_amd
import ns_workloads as W
from parity_util import assert_state_equal, oracle_brains, oracle_export_all, oracle_step_all

pytestmark = pytest.mark.gpu


def _run_compare(w, form, calls, check_every=10, tiles=None, mask_fn=None):
    print(f'[parity] workload={w.name} brains={w.count} neurons={w.total_neurons} '
          f'synapses={w.total_synapses} form={form} calls={calls} tiles={tiles}', flush=True)
    batch = ns_amd.GpuBatch(workload=w)
    if tiles:
        batch.set_tiles(*tiles)
    batch.set_form(form)
    batch.reset(w.seeds())
    brains = oracle_brains(w)
    for call in range(calls):
        x = w.inputs(call)
        mask = mask_fn(call) if mask_fn else None
        out, status = batch.step(x, w.ticks, w.gain, mask)
        oo, orc = oracle_step_all(brains, x, w.ticks, w.gain, mask)
        assert np.array_equal(out, oo), f'call {call}: motor outputs differ'
        assert np.array_equal(status != 0, orc != 0), f'call {call}: status differs'
        if call % check_every == check_every - 1 or call == calls - 1:
            assert_state_equal(batch.export(), oracle_export_all(brains), f'call {call}')
    info = batch.info()
    batch.close()
    return info


@pytest.mark.parametrize('form', [ns_amd.FORM_RESIDENT, ns_amd.FORM_PER_TICK])
def test_c1_plain_tick_bit_exact(form):
    w = W.c1(calls=200)
    info = _run_compare(w, form, calls=200, check_every=50)
    assert info['events'] > 0


def test_c1_per_tick_poisson_bit_exact():
    """C1's genomes with the sensory input redrawn every 1 ms tick (Poisson
    20 Hz per sensor), one tick per call: 800 ticks of 16 brains."""
    w = W.c1_per_tick(brains=16, ticks=800)
    info = _run_compare(w, ns_amd.FORM_RESIDENT, calls=800, check_every=200)
    assert info['events'] > 0


@pytest.mark.parametrize('form', [ns_amd.FORM_RESIDENT, ns_amd.FORM_PER_TICK])
@pytest.mark.parametrize('seed,types', [(11, (0, 1, 4, 5, 6)), (12, (0, 1, 2, 3, 4, 5, 6))])
def test_mixed_plastic_brains_bit_exact(form, seed, types):
    w = W.small_mix(seed=seed, brains=12, calls=40, types=types)
    _run_compare(w, form, calls=40, check_every=5)


def test_masks_and_resets_match():
    w = W.small_mix(seed=21, brains=6, calls=30, types=(0, 1, 4, 5))
    _run_compare(w, ns_amd.FORM_RESIDENT, calls=30, check_every=3,
                 mask_fn=lambda c: np.array([1, c % 2, 1, c % 3 == 0, 0, 1], np.uint8))


@pytest.mark.parametrize('tiles', [(256, 1, 4), (512, 2, 4), (1024, 1, 8), (128, 2, 8),
                                   (256, 1, 4, 3), (1024, 1, 0, 1, 4448), (256, 1, 0, 1, 896),
                                   (512, 2, 0, 1, 4448), (1024, 2, 0, 1, 4224)])
def test_tile_shapes_agree(tiles):
    w = W.small_mix(seed=31, brains=10, calls=12, types=(0, 1, 4, 5, 6), lo=20, hi=120)
    _run_compare(w, ns_amd.FORM_RESIDENT, calls=12, check_every=4, tiles=tiles)


def test_two_neurons_per_thread_packs():
    """The 1024-thread tile with two neuron slots per thread: packs of more
    than 1,024 neurons, so every thread also runs a second neuron."""
    w = W.small_mix(seed=37, brains=24, calls=12, types=(0, 1, 4, 5, 6), lo=60, hi=120,
                    fanout=1.5)
    assert w.total_neurons > 1400
    info = _run_compare(w, ns_amd.FORM_RESIDENT, calls=12, check_every=4, tiles=(1024, 2, 0, 1, 4224))
    assert info['resident_packs'] <= 2 and info['big_brains'] == 0


def test_full_packs_reach_every_index_bit():
    """256 brains of 16 neurons on the two-neuron tile: two packs of 128
    brains and 2,048 neurons each, so a pack's brain index reaches 127 and its
    neuron index 2,047, the tops of their bit fields in the synapse's LDS
    endpoint word."""
    w = W.small_mix(seed=41, brains=256, calls=6, types=(0, 1, 4, 5, 6), lo=16, hi=16, fanout=1.5)
    assert w.total_neurons == 2 * 2048
    info = _run_compare(w, ns_amd.FORM_RESIDENT, calls=6, check_every=3, tiles=(1024, 2, 0, 1, 4224))
    assert info['resident_packs'] == 2 and info['big_brains'] == 0


@pytest.mark.parametrize('tiles', [(1024, 2, 0, 1, 4224), (1024, 1, 0, 1, 4448)])
def test_release_bursts_past_the_queue(tiles):
    """Every neuron driven, fan-out 48, 1 ms delays, strong weights: about
    2,000 short-term arrivals per brain per tick, so most of a pack's queued
    releases go past the LDS release array and are released by their post
    neuron's thread during the input sum."""
    rng = np.random.default_rng(71)
    w = W._assemble('burst', rng, [64] * 6, 48, W._uniform_types([0, 4]), 0.0, 64, 8, 2, 1,
                    1.5, 0.2, W._stdp_mix, 0.0, 71, 1.0, 230.0, 16, 6)
    info = _run_compare(w, ns_amd.FORM_RESIDENT, calls=6, check_every=2, tiles=tiles)
    assert info['events'] > 6 * 16 * 6 * 256 and info['big_brains'] == 0


def test_big_brains_bit_exact():
    """Brains larger than a pack (here > 128 neurons or > 512 synapses) run in
    a workgroup of their own inside the resident launch."""
    w = W.small_mix(seed=33, brains=14, calls=20, types=(0, 1, 4, 5, 6), lo=16, hi=400,
                    fanout=3.0)
    info = _run_compare(w, ns_amd.FORM_RESIDENT, calls=20, check_every=5, tiles=(128, 1, 4, 4))
    assert info['big_brains'] > 0 and info['resident_packs'] > 0
    info = _run_compare(w, ns_amd.FORM_RESIDENT, calls=20, check_every=5, tiles=(256, 1, 0, 1, 896))
    assert info['big_brains'] > 0 and info['resident_packs'] > 0 and info['lds_synapses'] > 0


def test_plasticity_free_pools_use_the_static_lds_kernel():
    w = W.small_mix(seed=35, brains=10, calls=10, types=(0, 1, 4, 5), plasticity=False)
    info = _run_compare(w, ns_amd.FORM_RESIDENT, calls=10, check_every=5, tiles=(256, 1, 0, 1, 896))
    assert info['lds_synapses'] == 1408


@pytest.mark.parametrize('tiles,capacity', [((1024, 2, 0, 1, 4224), 10272), ((1024, 1, 0, 1, 4448), 10784)])
def test_static_packs_fill_the_lds(tiles, capacity):
    """Without plastic synapses a pack's LDS holds no traces and no initial
    weights, so the full-LDS tiles take 10,272 / 10,784 synapses per pack:
    here two packs of about 8,700 synapses each, past what 18 bytes per
    synapse would allow (8,096)."""
    w = W.small_mix(seed=36, brains=14, calls=8, types=(0, 1, 4, 5), plasticity=False, lo=60, hi=110,
                    fanout=14.0)
    assert w.total_synapses > 2 * 8096
    info = _run_compare(w, ns_amd.FORM_RESIDENT, calls=8, check_every=4, tiles=tiles)
    assert info['lds_synapses'] == capacity
    assert info['resident_packs'] == 2 and info['big_brains'] == 0


@pytest.mark.parametrize('tiles', [(1024, 2, 0, 1, 4224), (1024, 1, 0, 1, 4448)])
def test_long_quiet_intervals_replay_exactly(tiles):
    """Calls of 300 ticks with sparse input: pair synapses defer their quiet
    plasticity steps for long stretches (up to the 127-step cap that forces a
    replay) and replay them at arrivals, post spikes, modulator events and
    write-back, bit-identical to the oracle's step of every tick."""
    w = W.small_mix(seed=57, brains=10, calls=4, types=(0, 1, 4, 5), lo=20, hi=90, fanout=3.0,
                    modulator_p=0.08)
    w.input_p = 0.15
    batch = ns_amd.GpuBatch(workload=w)
    try:
        batch.set_tiles(*tiles)
        batch.reset(w.seeds())
        brains = oracle_brains(w)
        for call, ticks in enumerate((300, 17, 300, 5)):
            x = w.inputs(call)
            out, status = batch.step(x, ticks, w.gain)
            oo, orc = oracle_step_all(brains, x, ticks, w.gain)
            assert np.array_equal(out, oo), f'call {call}: motor outputs differ'
            assert_state_equal(batch.export(), oracle_export_all(brains), f'call {call}')
        assert batch.info()['events'] > 0
    finally:
        batch.close()


@pytest.mark.parametrize('tiles,big', [((1024, 2, 0, 1, 4224), False), ((1024, 1, 0, 1, 4448), False),
                                       ((512, 2, 0, 1, 4448), False), ((256, 1, 0, 1, 896), True)])
def test_identity_ticks_skipped_bit_exact(tiles, big):
    """Brains of 100-400 neurons with sparse wiring: most neurons away from the
    sensors settle to a state their tick no longer changes, so the resident
    kernel skips their ticks (until a synapse onto them arrives): in packs,
    which also move them to the end of the pack's slot order after each
    launch, and (on the 256-thread tile) in brains too large for a pack,
    which also defer their quiet plasticity steps.  Every state byte and
    every output must still equal the oracle's, which runs every tick."""
    w = W.small_mix(seed=43, brains=12, calls=40, types=(0, 1, 4, 5), lo=100, hi=400, fanout=1.0,
                    modulator_p=0.05)
    info = _run_compare(w, ns_amd.FORM_RESIDENT, calls=40, check_every=5, tiles=tiles)
    print(f"[parity] skipped {info['skipped_neuron_ticks']} of {w.total_neurons * w.ticks * 40} neuron ticks, "
          f"big brains {info['big_brains']}", flush=True)
    assert (info['big_brains'] > 0) == big
    assert info['skipped_neuron_ticks'] > 0.2 * w.total_neurons * w.ticks * 40


def test_settled_marks_end_with_an_imported_state():
    """A cell without input that ends a call settled skips the first tick of
    the next call too.  Importing a state replaces what that mark was about:
    after 20 calls (cells settled, marks set) the state right after reset is
    imported, and the next calls must equal the oracle's run from that state,
    which computes every tick."""
    w = W.small_mix(seed=43, brains=12, calls=40, types=(0, 1, 4, 5), lo=100, hi=400, fanout=1.0,
                    modulator_p=0.05)
    print(f'[parity] settled marks: brains={w.count} neurons={w.total_neurons}', flush=True)
    batch = ns_amd.GpuBatch(workload=w)
    try:
        batch.set_tiles(1024, 2, 0, 1, 4224)
        batch.reset(w.seeds())
        start = batch.export()
        for call in range(20):
            batch.step(w.inputs(call), w.ticks, w.gain)
        skipped = batch.info()['skipped_neuron_ticks']
        assert skipped > 0.2 * w.total_neurons * w.ticks * 20
        batch.import_state(*start)
        brains = oracle_brains(w)
        for call in range(6):
            x = w.inputs(call)
            out, status = batch.step(x, w.ticks, w.gain)
            oo, orc = oracle_step_all(brains, x, w.ticks, w.gain)
            assert np.array_equal(out, oo), f'call {call} after the import: motor outputs differ'
            assert_state_equal(batch.export(), oracle_export_all(brains), f'call {call} after the import')
    finally:
        batch.close()
