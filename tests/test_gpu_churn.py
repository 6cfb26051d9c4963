"""Brain turnover (births and deaths) through ns_gpu_replace: every result
after a replacement equals that of oracle brains where the replaced ones were
created fresh at the same point and reset with the same seeds.  Covers
placement into spare slots, into the slot a brain vacated, into slots freed
by earlier rounds, new slots appended past the last one when no free slot
fits (in a pack of their own, or a workgroup of their own when larger than a
pack), and the rebuild taken when a newborn is larger than the room kept for
appended slots or has LP/Prinz cells."""
import numpy as np
import pytest

import ns_amd
import ns_workloads as W
from oracle import OracleBrain
from parity_util import assert_state_equal, oracle_brains, oracle_export_all, oracle_step_all

pytestmark = pytest.mark.gpu


def _run(form, tiles, types, rounds):
    w = W.small_mix(seed=41, brains=16, calls=40, types=types, lo=8, hi=120)
    print(f'[churn] form={form} tiles={tiles} types={types} brains={w.count}', flush=True)
    batch = ns_amd.GpuBatch(workload=w)
    if tiles:
        batch.set_tiles(*tiles)
    batch.set_form(form)
    batch.reserve(0.19, 0.25)
    batch.reset(w.seeds())
    brains = oracle_brains(w)
    md = max(8, w.max_delay)
    handle_sets = []
    try:
        for call in range(w.calls):
            if call in rounds:
                seed, idx, lo, hi, ntypes = rounds[call]
                w2 = W.small_mix(seed=seed, brains=len(idx), calls=1, types=ntypes, lo=lo, hi=hi)
                genomes = [w2.brain(i) for i in range(len(idx))]
                hs = ns_amd.GenomeHandles(genomes, w.channels, max_delay=md)
                handle_sets.append(hs)
                seeds = [7000 + 13 * call + j for j in range(len(idx))]
                batch.replace(idx, hs, range(len(idx)), seeds)
                for j, i in enumerate(idx):
                    n, s, m = genomes[j]
                    b = OracleBrain(n, s, m, w.channels, max_delay=md)
                    b.reset(seeds[j])
                    brains[i] = b
                assert_state_equal(batch.export(), oracle_export_all(brains), f'after replace {call}')
            x = w.inputs(call)
            out, status = batch.step(x, w.ticks, w.gain)
            oo, orc = oracle_step_all(brains, x, w.ticks, w.gain)
            assert np.array_equal(out, oo), f'call {call}: motor outputs differ'
            assert np.array_equal(status != 0, orc != 0)
            if call % 5 == 4:
                assert_state_equal(batch.export(), oracle_export_all(brains), f'call {call}')
        info = batch.info()
    finally:
        for hs in handle_sets:
            hs.close()
        batch.close()
    return info


ROUNDS = {
    3: (101, [1, 7, 3], 8, 60, (0, 1, 4, 5, 6)),          # small newborns: spares / own slots
    9: (102, [2, 11, 5, 14], 8, 120, (0, 1, 4, 5, 6)),   # mixed sizes, earlier vacancies
    16: (103, [7], 400, 420, (0, 1, 4, 5)),               # larger than any free slot: appended
    24: (104, [0, 9, 12], 8, 100, (0, 1, 4, 5)),
    30: (107, [4], 2400, 2500, (0, 1, 4, 5)),             # larger than the room left: rebuild
    34: (108, [13], 3000, 3100, (0, 1, 4, 5)),            # past the new largest slot: appended,
}                                                         # a workgroup of its own


@pytest.mark.parametrize('form', [ns_amd.FORM_RESIDENT, ns_amd.FORM_PER_TICK])
def test_replacements_match_fresh_brains(form):
    info = _run(form, None, (0, 1, 4, 5, 6), ROUNDS)
    assert info['placements'] >= 10 and info['appended_slots'] >= 2 and info['relayouts'] >= 1


@pytest.mark.parametrize('tiles', [(1024, 1, 0, 1, 4448), (1024, 2, 0, 1, 4224), (256, 2, 4)])
def test_replacements_on_lds_and_register_tiles(tiles):
    info = _run(ns_amd.FORM_RESIDENT, tiles, (0, 1, 4, 5, 6), ROUNDS)
    assert info['placements'] >= 10 and info['appended_slots'] >= 2 and info['relayouts'] >= 1


def test_slow_cell_newborn_rebuilds():
    rounds = {4: (105, [3, 8], 8, 40, (0, 1, 2, 3, 4)), 12: (106, [3, 10], 8, 60, (0, 1, 4, 5))}
    info = _run(ns_amd.FORM_RESIDENT, None, (0, 1, 4, 5), rounds)
    assert info['relayouts'] >= 1
