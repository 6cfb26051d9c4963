"""World-pool surface (checkpoint records in, 16-tick world steps,
homeostasis, records out) against the oracle's world rules, bit-exact."""
import ctypes as ct

import numpy as np
import pytest

import ns_amd
import ns_workloads as W
import oracle as O
from parity_util import Hip

pytestmark = pytest.mark.gpu

NEURON_RECORD = np.dtype({
    'names': ['slot', 'owner', 'next', 'port', 'gene', 'state', 'pending_exc', 'pending_inh',
              'mod_event_tick', 'spike_history'],
    'formats': ['<i4', '<i4', '<i4', '<i4', O.NEURON_GENE, O.NEURON_STATE, ('<f4', 4), '<f4',
                ('<u8', 6), ('u1', 9)],
    'offsets': [0, 4, 8, 12, 16, 140, 260, 276, 280, 328], 'itemsize': 344})
# the 17-slot layout (delays 1-16): the same fields, a longer ring
NEURON_RECORD17 = np.dtype({
    'names': list(NEURON_RECORD.names),
    'formats': ['<i4', '<i4', '<i4', '<i4', O.NEURON_GENE, O.NEURON_STATE, ('<f4', 4), '<f4',
                ('<u8', 6), ('u1', 17)],
    'offsets': [0, 4, 8, 12, 16, 140, 260, 276, 280, 328], 'itemsize': 352})
SYNAPSE_RECORD = np.dtype({
    'names': ['slot', 'owner', 'next', 'gene', 'state'],
    'formats': ['<i4', '<i4', '<i4', O.SYNAPSE_GENE, O.SYNAPSE_STATE],
    'offsets': [0, 4, 8, 12, 56], 'itemsize': 136})


def build_world(seed=7, animals=10, lo=30, hi=60, types=(0, 1, 4, 5, 6), max_delay=8,
                record=NEURON_RECORD):
    """Records for `animals` animals with interleaved slots; genes from the
    small mixed workload, ports 0..18 sensory, 19..23 motors; with the
    17-slot record, random ring contents."""
    lib = O.oracle_lib()
    w = W.small_mix(seed=seed, brains=animals, types=types, lo=lo, hi=hi,
                    modulator_p=0.15, max_delay=max_delay)
    rng = np.random.default_rng(seed)
    neurons, synapses = [], []
    slot_n, slot_s = 0, 0
    brains = []
    for a in range(animals):
        nrow, srow, mot = w.brain(a)
        n = len(nrow)
        slots = slot_n + 2 * np.arange(n) + (a % 2)       # interleave two animals
        slot_n += 2 * n
        genes = np.zeros(n, O.NEURON_GENE)
        for i in range(n):
            lib.ora_fn_profile(int(nrow[i, 0]), genes[i:i + 1].ctypes.data)
        genes['active'] = nrow[:, 2].astype(np.uint8)
        genes['inhibitory'] = nrow[:, 3].astype(np.uint8)
        genes['modulator'] = nrow[:, 4].astype(np.uint8)
        ch = np.full(n, -1, np.int8)
        ch[:8] = rng.integers(0, 19, 8)
        ch[:3] = [0, 2, 7]                                # receptor path below 6, not above
        genes['sensory_channel'] = ch
        port = np.full(n, -1, np.int32)
        port[n - 5:] = 19 + np.arange(5)
        for i in range(n):
            st = np.zeros(1, O.NEURON_STATE)
            lib.ora_fn_initial_neuron(genes[i:i + 1].ctypes.data, int(seed * 1000 + a * 97 + i), st.ctypes.data)
            r = np.zeros(1, record)
            r['slot'] = slots[i]; r['owner'] = a; r['next'] = -1; r['port'] = port[i]
            r['gene'] = genes[i]; r['state'] = st[0]
            if record is NEURON_RECORD17:
                r['spike_history'] = (rng.random(17) < 0.2).astype(np.uint8)
            neurons.append(r)
        sg = np.zeros(len(srow), O.SYNAPSE_GENE)
        sg['pre'] = slots[srow[:, 0].astype(int)]
        sg['post'] = slots[srow[:, 1].astype(int)]
        sg['initial_weight_ns'] = srow[:, 2] * 3
        sg['active'] = srow[:, 3].astype(np.uint8)
        sg['delay'] = srow[:, 4].astype(np.uint8)
        sg['branch'] = srow[:, 5].astype(np.uint8)
        sg['utilization'] = srow[:, 6]
        sg['recovery_ms'] = srow[:, 7]
        sg['plasticity'] = srow[:, 9].astype(np.uint8)
        sg['mod_receptors'] = srow[:, 10].astype(np.uint8)
        sg['release_sites'] = 14
        for k in range(len(srow)):
            st = np.zeros(1, O.SYNAPSE_STATE)
            lib.ora_fn_initial_synapse(sg[k:k + 1].ctypes.data, int(seed * 7 + a * 131 + k), st.ctypes.data)
            r = np.zeros(1, SYNAPSE_RECORD)
            r['slot'] = slot_s + k; r['owner'] = a; r['next'] = -1
            r['gene'] = sg[k]; r['state'] = st[0]
            synapses.append(r)
        slot_s += len(srow)
        brains.append((genes, sg, slots, ch, port))
    nr = O.cat_records(neurons, record)
    nr = O.cat_records([nr[np.argsort(nr['slot'])]], record)
    return nr, O.cat_records(synapses, SYNAPSE_RECORD), brains


def oracle_world(nr, sr, brains, max_delay=8):
    lib = O.oracle_lib()
    handles = []
    for a, (genes, sg, slots, ch, port) in enumerate(brains):
        mine = nr[nr['owner'] == a]
        mine = mine[np.argsort(mine['slot'])]
        local = {int(s): i for i, s in enumerate(mine['slot'])}
        g = np.ascontiguousarray(mine['gene'])
        syn = sr[sr['owner'] == a]
        sgl = np.ascontiguousarray(syn['gene']).copy()
        sgl['pre'] = [local[int(x)] for x in syn['gene']['pre']]
        sgl['post'] = [local[int(x)] for x in syn['gene']['post']]
        ports = np.ascontiguousarray(mine['port'], np.int32)
        h = lib.ora_create_world(len(g), len(sgl), g.ctypes.data, sgl.ctypes.data,
                                 ports.ctypes.data_as(O._i32), max_delay)
        ns = np.ascontiguousarray(mine['state'])
        ss = np.ascontiguousarray(syn['state'])
        ring = (np.ascontiguousarray(mine['spike_history']) if max_delay > 8 else None)
        lib.ora_import(h, ns.ctypes.data, ss.ctypes.data, O.ptr(ring, O._u8), 0)
        handles.append((h, mine['slot'], syn['slot']))
    return handles


@pytest.mark.parametrize('types', [(0, 1, 4, 5, 6), (0, 1, 2, 3, 5)], ids=['adex-lfs', 'lp-prinz'])
def test_world_steps_match_oracle_and_round_trip_records(types):
    L = ns_amd.lib()
    lib = O.oracle_lib()
    nr, sr, brains = build_world(types=types)
    P = len(brains)
    print(f'[world] animals={P} neurons={len(nr)} synapses={len(sr)}', flush=True)
    pool = L.ns_pool_create(P, int(nr['slot'].max()) + 1, int(sr['slot'].max()) + 1)
    assert pool, ns_amd.error()
    try:
        assert L.ns_pool_import_records(pool, nr.ctypes.data, len(nr), sr.ctypes.data, len(sr)) == 0, ns_amd.error()
        ora = oracle_world(nr, sr, brains)
        rng = np.random.default_rng(3)
        motors = np.zeros(P * 5, np.int32)
        ora_motors = np.zeros(P * 5, np.float32)
        frozen = np.zeros(P, np.uint8)
        frozen[3] = 1
        first = 5000
        for step in range(12):
            senses = (rng.random(P * 19) < 0.5).astype(np.float32)
            tick = first + 16 * step
            assert L.ns_world_neural_step(pool, tick, senses.ctypes.data_as(O._f),
                                          motors.ctypes.data_as(O._i32), frozen.ctypes.data_as(O._u8)) == 0, ns_amd.error()
            age = (np.arange(P) * 3 + step).astype(np.int32)
            assert L.ns_world_homeostasis(pool, age.ctypes.data_as(O._i32), frozen.ctypes.data_as(O._u8)) == 0
            for a, (h, _, _) in enumerate(ora):
                lib.ora_set_tick(h, tick)
                lib.ora_set_frozen(h, int(frozen[a]))
                out = np.zeros(5, np.float32)
                mine = np.ascontiguousarray(senses[a * 19:(a + 1) * 19])
                lib.ora_step(h, mine.ctypes.data_as(O._f), 16, 230.0, out.ctypes.data_as(O._f), None)
                ora_motors[a * 5:(a + 1) * 5] += out
                lib.ora_world_homeostasis(h, int(age[a]))
        assert np.array_equal(motors, ora_motors.astype(np.int32))
        out_n = np.zeros(len(nr), NEURON_RECORD)
        out_s = np.zeros(len(sr), SYNAPSE_RECORD)
        cn, cs = ct.c_int(), ct.c_int()
        assert L.ns_pool_export_records(pool, out_n.ctypes.data, len(nr), out_s.ctypes.data, len(sr),
                                        ct.byref(cn), ct.byref(cs)) == 0
        assert cn.value == len(nr) and cs.value == len(sr)
        assert np.array_equal(out_n['slot'], nr['slot']) and np.array_equal(out_s['slot'], sr['slot'])
        for a, (h, nslots, sslots) in enumerate(ora):
            n = len(nslots)
            ns = np.zeros(n, O.NEURON_STATE)
            ss = np.zeros(len(sslots), O.SYNAPSE_STATE)
            h9 = np.zeros((n, 9), np.uint8)
            stamps = np.zeros(n * 6, np.uint64)
            lib.ora_export(h, ns.ctypes.data, ss.ctypes.data, h9.ctypes.data_as(O._u8), None, None,
                           stamps.ctypes.data_as(O._u64))
            dn = out_n[np.isin(out_n['slot'], nslots)]
            ds = out_s[np.isin(out_s['slot'], sslots)]
            assert np.array_equal(O.state_bytes(np.ascontiguousarray(dn['state']), O.NEURON_STATE_BYTES),
                                  O.state_bytes(ns, O.NEURON_STATE_BYTES)), f'animal {a} neurons'
            assert np.array_equal(O.state_bytes(np.ascontiguousarray(ds['state']), O.SYNAPSE_STATE_BYTES),
                                  O.state_bytes(ss, O.SYNAPSE_STATE_BYTES)), f'animal {a} synapses'
            assert np.array_equal(dn['spike_history'], h9), f'animal {a} rings'
            assert np.array_equal(dn['mod_event_tick'].reshape(-1), stamps), f'animal {a} stamps'
        # re-import the exported records into a fresh pool and continue: same as continuing
        pool2 = L.ns_pool_create(P, int(nr['slot'].max()) + 1, int(sr['slot'].max()) + 1)
        assert L.ns_pool_import_records(pool2, out_n.ctypes.data, len(nr), out_s.ctypes.data, len(sr)) == 0
        senses = (rng.random(P * 19) < 0.5).astype(np.float32)
        tick = first + 16 * 12
        m1 = np.zeros(P * 5, np.int32)
        m2 = np.zeros(P * 5, np.int32)
        for pl, m in ((pool, m1), (pool2, m2)):
            assert L.ns_world_neural_step(pl, tick, senses.ctypes.data_as(O._f), m.ctypes.data_as(O._i32),
                                          frozen.ctypes.data_as(O._u8)) == 0
        assert np.array_equal(m1, m2)
        e1 = np.zeros(len(nr), NEURON_RECORD)
        e2 = np.zeros(len(nr), NEURON_RECORD)
        f1 = np.zeros(len(sr), SYNAPSE_RECORD)
        f2 = np.zeros(len(sr), SYNAPSE_RECORD)
        L.ns_pool_export_records(pool, e1.ctypes.data, len(nr), f1.ctypes.data, len(sr), ct.byref(cn), ct.byref(cs))
        L.ns_pool_export_records(pool2, e2.ctypes.data, len(nr), f2.ctypes.data, len(sr), ct.byref(cn), ct.byref(cs))
        assert e1.tobytes() == e2.tobytes() and f1.tobytes() == f2.tobytes()
        L.ns_pool_destroy(pool2)
    finally:
        L.ns_pool_destroy(pool)


def _field_bytes(dtype, base=0, out=None):
    """Byte mask of the bytes a (nested) structured dtype's fields cover:
    numpy leaves the holes between fields unset when it copies records
    field by field (boolean indexing), so record comparisons skip them."""
    out = np.zeros(dtype.itemsize, bool) if out is None else out
    for name in dtype.names:
        sub, off = dtype.fields[name][:2]
        if sub.names:
            _field_bytes(sub, base + off, out)
        else:
            out[base + off:base + off + sub.itemsize] = True
    return out


def _same_records(x, y):
    if len(x) != len(y):
        return False
    m = _field_bytes(x.dtype)
    rx = np.ascontiguousarray(x).view(np.uint8).reshape(len(x), -1)[:, m]
    ry = np.ascontiguousarray(y).view(np.uint8).reshape(len(y), -1)[:, m]
    return np.array_equal(rx, ry)


def _export(L, pool, ncap, scap):
    n = np.zeros(ncap, NEURON_RECORD)
    s = np.zeros(scap, SYNAPSE_RECORD)
    cn, cs = ct.c_int(), ct.c_int()
    assert L.ns_pool_export_records(pool, n.ctypes.data, ncap, s.ctypes.data, scap,
                                    ct.byref(cn), ct.byref(cs)) == 0, ns_amd.error()
    return n[:cn.value].copy(), s[:cs.value].copy()


def _newborn(dn, ds, donor, owner, nslots, sslots):
    """Donor animal's records relabelled to `owner` on the given free slots
    (its synapses' endpoints follow their neurons)."""
    mine = dn[dn['owner'] == donor].copy()
    syn = ds[ds['owner'] == donor].copy()
    where = {int(old): int(new) for old, new in zip(mine['slot'], nslots)}
    mine['slot'] = nslots
    mine['owner'] = owner
    g = syn['gene'].copy()
    g['pre'] = [where[int(x)] for x in g['pre']]
    g['post'] = [where[int(x)] for x in g['post']]
    syn['gene'] = g
    syn['slot'] = sslots
    syn['owner'] = owner
    return mine, syn


def test_births_and_deaths_equal_export_merge_reimport():
    """ns_pool_replace_records with a slot reserve (placements into free
    slots, one forced rebuild, births into empty animal indices, deaths
    without births, before the first step and between steps, then a skipped
    tick) equals exporting the records, applying the same change and
    importing them, bit for bit, after every change and every step."""
    L = ns_amd.lib()
    nr, sr, _ = build_world(seed=11, animals=12)
    dn, ds, _ = build_world(seed=12, animals=8)             # donor genomes for newborns
    bn, bs, _ = build_world(seed=13, animals=1, lo=150, hi=160)   # larger than every slot
    ln, ls, _ = build_world(seed=14, animals=1, types=(0, 2, 3, 4))   # LP/Prinz cells: a rebuild
    P, ncap, scap = 16, 4000, 16000
    print(f'[world-births] animals={P} neurons={len(nr)} synapses={len(sr)}', flush=True)
    A = L.ns_pool_create(P, ncap, scap)
    B = L.ns_pool_create(P, ncap, scap)
    assert A and B, ns_amd.error()
    try:
        assert L.ns_pool_reserve(A, 0.19, 0.25) == 0, ns_amd.error()
        for pl in (A, B):
            assert L.ns_pool_import_records(pl, nr.ctypes.data, len(nr), sr.ctypes.data, len(sr)) == 0
        # (dead, [(donor records, donor, owner)]) per change; step index -> change
        changes = {
            0: ([3, 12], [(dn, ds, 0, 3), (dn, ds, 1, 12)]),            # before the first step
            2: ([0, 5, 13], [(dn, ds, 2, 0), (dn, ds, 3, 13)]),          # 5 just dies
            4: ([7], [(bn, bs, 0, 7)]),                                  # no slot fits: rebuild
            5: ([1, 2, 5], [(dn, ds, 4, 2), (dn, ds, 5, 5), (dn, ds, 6, 1)]),
            6: ([9], [(ln, ls, 0, 9)]),
        }
        rng = np.random.default_rng(5)
        motors_a = np.zeros(P * 5, np.int32)
        motors_b = np.zeros(P * 5, np.int32)
        frozen = np.zeros(P, np.uint8)
        frozen[4] = 1
        tick = 3000
        for step in range(8):
            if step in changes:
                dead, born = changes[step]
                en, es = _export(L, B, ncap, scap)
                keep_n = en[~np.isin(en['owner'], dead)]
                keep_s = es[~np.isin(es['owner'], dead)]
                free_n = np.setdiff1d(np.arange(ncap), keep_n['slot'])
                free_s = np.setdiff1d(np.arange(scap), keep_s['slot'])
                new_n, new_s = [], []
                for src_n, src_s, donor, owner in born:
                    k = int((src_n['owner'] == donor).sum())
                    m = int((src_s['owner'] == donor).sum())
                    a, b = _newborn(src_n, src_s, donor, owner, free_n[:k], free_s[:m])
                    free_n, free_s = free_n[k:], free_s[m:]
                    new_n.append(a)
                    new_s.append(b)
                new_n = O.cat_records(new_n, NEURON_RECORD)             # keeps the record padding
                new_s = O.cat_records(new_s, SYNAPSE_RECORD)
                new_n = O.cat_records([new_n[np.argsort(new_n['slot'])]], NEURON_RECORD)
                new_s = O.cat_records([new_s[np.argsort(new_s['slot'])]], SYNAPSE_RECORD)
                d = np.array(dead, np.int32)
                assert L.ns_pool_replace_records(A, len(d), d.ctypes.data_as(O._i32), new_n.ctypes.data,
                                                 len(new_n), new_s.ctypes.data, len(new_s)) == 0, ns_amd.error()
                mn = O.cat_records([keep_n, new_n], NEURON_RECORD)
                ms = O.cat_records([keep_s, new_s], SYNAPSE_RECORD)
                mn = O.cat_records([mn[np.argsort(mn['slot'])]], NEURON_RECORD)
                ms = O.cat_records([ms[np.argsort(ms['slot'])]], SYNAPSE_RECORD)
                assert L.ns_pool_import_records(B, mn.ctypes.data, len(mn), ms.ctypes.data, len(ms)) == 0
                xa, xb = _export(L, A, ncap, scap), _export(L, B, ncap, scap)
                assert _same_records(xa[0], xb[0]), f'neurons after change {step}'
                assert _same_records(xa[1], xb[1]), f'synapses after change {step}'
            senses = (rng.random(P * 19) < 0.5).astype(np.float32)
            tick += 48 if step == 7 else 16                          # the last step skips two
            for pl, m in ((A, motors_a), (B, motors_b)):
                assert L.ns_world_neural_step(pl, tick, senses.ctypes.data_as(O._f), m.ctypes.data_as(O._i32),
                                              frozen.ctypes.data_as(O._u8)) == 0, ns_amd.error()
                age = (np.arange(P) * 2 + step).astype(np.int32)
                assert L.ns_world_homeostasis(pl, age.ctypes.data_as(O._i32),
                                              frozen.ctypes.data_as(O._u8)) == 0, ns_amd.error()
            assert np.array_equal(motors_a, motors_b), f'motors at step {step}'
            xa, xb = _export(L, A, ncap, scap), _export(L, B, ncap, scap)
            assert _same_records(xa[0], xb[0]), f'neurons at step {step}'
            assert _same_records(xa[1], xb[1]), f'synapses at step {step}'
        assert motors_a.any()
        info = ns_amd.PoolStats()
        assert L.ns_pool_info(A, ct.byref(info)) == 0
        assert info.placements == 7 and info.relayouts == 2, (info.placements, info.relayouts)
        assert info.animals == 14 and info.free_slots > 0
        # rejected changes leave the pool as it was
        before = _export(L, A, ncap, scap)
        en, es = before
        clash_n = en[en['owner'] == 6][:1].copy()                # a survivor's slot
        clash_n['owner'] = 9
        d = np.array([9], np.int32)
        assert L.ns_pool_replace_records(A, 1, d.ctypes.data_as(O._i32), clash_n.ctypes.data, 1, None, 0) == -1
        assert 'slot in use' in ns_amd.error()
        d = np.array([8], np.int32)
        assert L.ns_pool_replace_records(A, 1, d.ctypes.data_as(O._i32), clash_n.ctypes.data, 1, None, 0) == -1
        assert 'owner not replaced' in ns_amd.error()
        d = np.array([8, 8], np.int32)
        assert L.ns_pool_replace_records(A, 2, d.ctypes.data_as(O._i32), None, 0, None, 0) == -1
        assert 'Invalid animal index' in ns_amd.error()
        after = _export(L, A, ncap, scap)
        assert _same_records(before[0], after[0]) and _same_records(before[1], after[1])
    finally:
        L.ns_pool_destroy(A)
        L.ns_pool_destroy(B)


def test_checkpoint_resume_equals_continuous_run(tmp_path):
    """The protocol of the reference's tests/checkpoint_roundtrip.py on the
    world pool: 12 world steps in one run equal 6 steps, an NSCHKP01
    checkpoint, a resume into a fresh pool and 6 more steps (motors every
    step, records at the end, bit for bit); a truncated checkpoint is
    refused and leaves the pool as it was."""
    import checkpoint as C
    L = ns_amd.lib()
    nr, sr, _ = build_world(seed=21, animals=12)
    counts = {a: (int((nr['owner'] == a).sum()), int((sr['owner'] == a).sum())) for a in range(12)}
    world = C.world_block(np.random.default_rng(21), counts)
    head = ns_amd.CheckpointHeader.from_buffer_copy(
        C.header(L.ns_checkpoint_layout(), step=6, initial=12, flags=1))
    P = C.P
    print(f'[world-checkpoint] animals=12 of {P} neurons={len(nr)} synapses={len(sr)}', flush=True)
    A, B, R = (L.ns_pool_create(P, C.NEURON_CAPACITY, C.SYNAPSE_CAPACITY) for _ in range(3))
    assert A and B and R, ns_amd.error()
    try:
        for pl in (A, B):
            assert L.ns_pool_import_records(pl, nr.ctypes.data, len(nr), sr.ctypes.data, len(sr)) == 0
        frozen = np.zeros(P, np.uint8)
        frozen[2] = 1

        def step(pl, k):
            senses = (np.random.default_rng(100 + k).random(P * 19) < 0.5).astype(np.float32)
            m = np.zeros(P * 5, np.int32)
            assert L.ns_world_neural_step(pl, 9000 + 16 * k, senses.ctypes.data_as(O._f), m.ctypes.data_as(O._i32),
                                          frozen.ctypes.data_as(O._u8)) == 0, ns_amd.error()
            age = (np.arange(P) * 3 + k).astype(np.int32)
            assert L.ns_world_homeostasis(pl, age.ctypes.data_as(O._i32), frozen.ctypes.data_as(O._u8)) == 0
            return m

        for k in range(6):
            assert np.array_equal(step(A, k), step(B, k))
        path = tmp_path / 'world.checkpoint.gz'
        assert L.ns_pool_save_checkpoint(B, str(path).encode(), ct.byref(head), world, len(world)) == 0, \
            ns_amd.error()
        got = ns_amd.CheckpointHeader()
        wb = ct.c_size_t()
        assert L.ns_pool_load_checkpoint(R, str(path).encode(), ct.byref(got), None, 0, ct.byref(wb)) == 0, \
            ns_amd.error()
        assert wb.value == len(world) and got.neuron_count == len(nr) and got.synapse_count == len(sr)
        back = ct.create_string_buffer(len(world))
        assert L.ns_pool_load_checkpoint(R, str(path).encode(), None, back, len(world), None) == 0
        assert back.raw == world
        for k in range(6, 12):
            ma, mr = step(A, k), step(R, k)
            assert np.array_equal(ma, mr), f'motors at step {k}'
        xa, xr = _export(L, A, len(nr), len(sr)), _export(L, R, len(nr), len(sr))
        assert _same_records(xa[0], xr[0]) and _same_records(xa[1], xr[1])
        assert xa[0]['state']['total_spikes'].sum() > 0      # the brains were active
        bad = tmp_path / 'corrupt.gz'
        payload = path.read_bytes()
        bad.write_bytes(payload[:len(payload) // 2])
        assert L.ns_pool_load_checkpoint(R, str(bad).encode(), None, None, 0, None) == -1
        assert 'incomplete or corrupt' in ns_amd.error()
        after = _export(L, R, len(nr), len(sr))
        assert _same_records(after[0], xr[0]) and _same_records(after[1], xr[1])
    finally:
        for pl in (A, B, R):
            L.ns_pool_destroy(pl)


def test_replace_into_empty_pool_and_after_every_death():
    """Edges of ns_pool_replace_records: births into a pool that holds no
    records (the first build), every animal dying (steps then do nothing and
    export nothing), and births after that; each state equals importing the
    same records."""
    L = ns_amd.lib()
    nr, sr, _ = build_world(seed=31, animals=4)
    P, ncap, scap = 6, int(nr['slot'].max()) + 1, int(sr['slot'].max()) + 1
    A = L.ns_pool_create(P, ncap, scap)
    B = L.ns_pool_create(P, ncap, scap)
    assert A and B, ns_amd.error()
    print(f'[world-edges] animals={P} neurons={len(nr)} synapses={len(sr)}', flush=True)
    senses = (np.random.default_rng(4).random(P * 19) < 0.5).astype(np.float32)
    frozen = np.zeros(P, np.uint8)

    def step(pl, tick):
        m = np.zeros(P * 5, np.int32)
        assert L.ns_world_neural_step(pl, tick, senses.ctypes.data_as(O._f), m.ctypes.data_as(O._i32),
                                      frozen.ctypes.data_as(O._u8)) == 0, ns_amd.error()
        return m

    try:
        all4 = np.arange(4, dtype=np.int32)
        # births into an empty pool: the first build
        assert L.ns_pool_replace_records(A, 4, all4.ctypes.data_as(O._i32), nr.ctypes.data, len(nr),
                                         sr.ctypes.data, len(sr)) == 0, ns_amd.error()
        assert L.ns_pool_import_records(B, nr.ctypes.data, len(nr), sr.ctypes.data, len(sr)) == 0
        for k in range(3):
            assert np.array_equal(step(A, 100 + 16 * k), step(B, 100 + 16 * k))
        xa, xb = _export(L, A, len(nr), len(sr)), _export(L, B, len(nr), len(sr))
        assert _same_records(xa[0], xb[0]) and _same_records(xa[1], xb[1])
        # every animal dies: nothing runs, nothing is exported
        assert L.ns_pool_replace_records(A, 4, all4.ctypes.data_as(O._i32), None, 0, None, 0) == 0
        info = ns_amd.PoolStats()
        assert L.ns_pool_info(A, ct.byref(info)) == 0
        assert info.animals == 0 and info.neuron_records == 0 and info.free_slots == info.slots
        assert not step(A, 148).any()
        n0, s0 = _export(L, A, 1, 1)
        assert len(n0) == 0 and len(s0) == 0
        # and births again, into the freed slots, from the exported state
        assert L.ns_pool_replace_records(A, 4, all4.ctypes.data_as(O._i32), xa[0].ctypes.data, len(xa[0]),
                                         xa[1].ctypes.data, len(xa[1])) == 0, ns_amd.error()
        assert L.ns_pool_import_records(B, xa[0].ctypes.data, len(xa[0]), xa[1].ctypes.data, len(xa[1])) == 0
        for k in range(3):
            assert np.array_equal(step(A, 148 + 16 * k), step(B, 148 + 16 * k))
        xa, xb = _export(L, A, len(nr), len(sr)), _export(L, B, len(nr), len(sr))
        assert _same_records(xa[0], xb[0]) and _same_records(xa[1], xb[1])
        assert xa[0]['state']['total_spikes'].sum() > 0
    finally:
        L.ns_pool_destroy(A)
        L.ns_pool_destroy(B)


def test_device_step_equals_host_step():
    """ns_world_neural_step_device (senses and motors in device memory,
    stream-ordered) equals ns_world_neural_step: motors every step and
    records at the end, across homeostasis, a replacement, a change of the
    freeze flags and a skipped tick."""
    L = ns_amd.lib()
    hip = Hip()
    nr, sr, _ = build_world(seed=41, animals=8)
    dn, ds, _ = build_world(seed=42, animals=2)
    P, ncap, scap = 10, 4000, 16000
    A = L.ns_pool_create(P, ncap, scap)
    B = L.ns_pool_create(P, ncap, scap)
    assert A and B, ns_amd.error()
    print(f'[world-device] animals={P} neurons={len(nr)} synapses={len(sr)}', flush=True)
    stream = ct.c_void_p()
    assert hip.h.hipStreamCreate(ct.byref(stream)) == 0
    d_senses = hip.alloc(P * 19 * 4)
    d_motors = hip.alloc(P * 5 * 4)
    frozen = np.zeros(P, np.uint8)
    frozen[1] = 1
    try:
        for pl in (A, B):
            assert L.ns_pool_reserve(pl, 0.19, 0.25) == 0
            assert L.ns_pool_import_records(pl, nr.ctypes.data, len(nr), sr.ctypes.data, len(sr)) == 0
        motors_h = np.zeros(P * 5, np.int32)
        hip.put(d_motors, np.zeros(P * 5, np.int32))
        tick = 700
        for step in range(8):
            senses = (np.random.default_rng(50 + step).random(P * 19) < 0.5).astype(np.float32)
            assert hip.h.hipStreamSynchronize(stream) == 0
            hip.put(d_senses, senses)
            tick += 32 if step == 5 else 16
            if step == 4:
                frozen[2] = 1                                        # flags change: uploaded again
            assert L.ns_world_neural_step(A, tick, senses.ctypes.data_as(O._f), motors_h.ctypes.data_as(O._i32),
                                          frozen.ctypes.data_as(O._u8)) == 0, ns_amd.error()
            assert L.ns_world_neural_step_device(B, tick, d_senses, d_motors, frozen.ctypes.data_as(O._u8),
                                                 stream) == 0, ns_amd.error()
            age = (np.arange(P) + step).astype(np.int32)
            for pl in (A, B):
                assert L.ns_world_homeostasis(pl, age.ctypes.data_as(O._i32), frozen.ctypes.data_as(O._u8)) == 0
            assert hip.h.hipStreamSynchronize(stream) == 0
            got = hip.get(d_motors, np.zeros(P * 5, np.int32))
            assert np.array_equal(motors_h, got), f'motors at step {step}'
            if step == 3:                                            # a birth in a free slot
                en, es = _export(L, A, ncap, scap)
                free_n = np.setdiff1d(np.arange(ncap), en['slot'][en['owner'] != 4])
                free_s = np.setdiff1d(np.arange(scap), es['slot'][es['owner'] != 4])
                k = int((dn['owner'] == 0).sum())
                m = int((ds['owner'] == 0).sum())
                bn, bs = _newborn(dn, ds, 0, 4, free_n[:k], free_s[:m])
                d = np.array([4], np.int32)
                for pl in (A, B):
                    assert L.ns_pool_replace_records(pl, 1, d.ctypes.data_as(O._i32), bn.ctypes.data, len(bn),
                                                     bs.ctypes.data, len(bs)) == 0, ns_amd.error()
        xa, xb = _export(L, A, ncap, scap), _export(L, B, ncap, scap)
        assert _same_records(xa[0], xb[0]) and _same_records(xa[1], xb[1])
        assert xa[0]['state']['total_spikes'].sum() > 0
    finally:
        L.ns_pool_destroy(A)
        L.ns_pool_destroy(B)
        hip.h.hipFree(d_senses)
        hip.h.hipFree(d_motors)
        hip.h.hipStreamDestroy(stream)


def test_17_slot_world_layout_steps_rings_and_checkpoint(tmp_path):
    """The 17-slot record layout (delays 1-16, ring slot = tick % 17):
    world steps of a ns_pool_create_ex(..., 17) pool equal the oracle's
    world brains with a 17-slot ring, random rings in, inactive neurons
    keeping their slots, rings out byte for byte; the 9-slot calls refuse the
    pool; a checkpoint of the 17-slot layout resumes into a fresh pool as the
    continuous run goes on."""
    import checkpoint as C
    L = ns_amd.lib()
    lib = O.oracle_lib()
    nr, sr, brains = build_world(seed=41, animals=10, max_delay=16, record=NEURON_RECORD17)
    assert sr['gene']['delay'].max() > 8 and (nr['gene']['active'] == 0).any()
    P = C.P
    print(f'[world17] animals=10 of {P} neurons={len(nr)} synapses={len(sr)}', flush=True)
    A = L.ns_pool_create_ex(P, C.NEURON_CAPACITY, C.SYNAPSE_CAPACITY, 17)
    R = L.ns_pool_create_ex(P, C.NEURON_CAPACITY, C.SYNAPSE_CAPACITY, 17)
    assert A and R and L.ns_pool_history(A) == 17, ns_amd.error()
    try:
        assert L.ns_pool_import_records(A, nr.ctypes.data, len(nr), sr.ctypes.data, len(sr)) == -1
        assert '17-slot' in ns_amd.error()
        twice = nr.copy()                              # two neurons of animal 0 on motor port 19
        mine = np.flatnonzero(twice['owner'] == 0)
        twice['port'][mine[0]] = 19
        assert L.ns_pool_import_records17(A, twice.ctypes.data, len(nr), sr.ctypes.data, len(sr)) == -1
        assert 'motor port' in ns_amd.error()
        assert L.ns_pool_import_records17(A, nr.ctypes.data, len(nr), sr.ctypes.data, len(sr)) == 0, \
            ns_amd.error()
        ora = oracle_world(nr, sr, brains, max_delay=16)
        frozen = np.zeros(P, np.uint8)
        frozen[4] = 1
        first = 7003                                   # not a multiple of 17: slots wrap mid-step
        ora_motors = np.zeros(P * 5, np.float32)

        def step(pl, k):
            senses = (np.random.default_rng(300 + k).random(P * 19) < 0.5).astype(np.float32)
            m = np.zeros(P * 5, np.int32)
            assert L.ns_world_neural_step(pl, first + 16 * k, senses.ctypes.data_as(O._f),
                                          m.ctypes.data_as(O._i32), frozen.ctypes.data_as(O._u8)) == 0, \
                ns_amd.error()
            age = (np.arange(P) * 3 + k).astype(np.int32)
            assert L.ns_world_homeostasis(pl, age.ctypes.data_as(O._i32), frozen.ctypes.data_as(O._u8)) == 0
            return m, senses, age

        def export17(pl):
            n = np.zeros(len(nr), NEURON_RECORD17)
            s_ = np.zeros(len(sr), SYNAPSE_RECORD)
            cn, cs = ct.c_int(), ct.c_int()
            assert L.ns_pool_export_records17(pl, n.ctypes.data, len(nr), s_.ctypes.data, len(sr),
                                              ct.byref(cn), ct.byref(cs)) == 0, ns_amd.error()
            assert cn.value == len(nr) and cs.value == len(sr)
            return n, s_

        motors = np.zeros(P * 5, np.int32)
        for k in range(8):
            m, senses, age = step(A, k)
            motors += m
            for a, (h, _, _) in enumerate(ora):
                lib.ora_set_tick(h, first + 16 * k)
                lib.ora_set_frozen(h, int(frozen[a]))
                out = np.zeros(5, np.float32)
                mine = np.ascontiguousarray(senses[a * 19:(a + 1) * 19])
                lib.ora_step(h, mine.ctypes.data_as(O._f), 16, 230.0, out.ctypes.data_as(O._f), None)
                ora_motors[a * 5:(a + 1) * 5] += out
                lib.ora_world_homeostasis(h, int(age[a]))
        assert np.array_equal(motors, ora_motors.astype(np.int32))
        out_n, out_s = export17(A)
        for a, (h, nslots, sslots) in enumerate(ora):
            n = len(nslots)
            ns = np.zeros(n, O.NEURON_STATE)
            ss = np.zeros(len(sslots), O.SYNAPSE_STATE)
            ring = np.zeros(n * 17, np.uint8)
            stamps = np.zeros(n * 6, np.uint64)
            lib.ora_export(h, ns.ctypes.data, ss.ctypes.data, None, ring.ctypes.data_as(O._u8), None,
                           stamps.ctypes.data_as(O._u64))
            dn = out_n[np.isin(out_n['slot'], nslots)]
            ds = out_s[np.isin(out_s['slot'], sslots)]
            assert np.array_equal(O.state_bytes(np.ascontiguousarray(dn['state']), O.NEURON_STATE_BYTES),
                                  O.state_bytes(ns, O.NEURON_STATE_BYTES)), f'animal {a} neurons'
            assert np.array_equal(O.state_bytes(np.ascontiguousarray(ds['state']), O.SYNAPSE_STATE_BYTES),
                                  O.state_bytes(ss, O.SYNAPSE_STATE_BYTES)), f'animal {a} synapses'
            assert np.array_equal(dn['spike_history'].reshape(-1), ring), f'animal {a} 17-slot rings'
            assert np.array_equal(dn['mod_event_tick'].reshape(-1), stamps), f'animal {a} stamps'
        # checkpoint of the 17-slot layout, resumed into a fresh pool
        counts = {a: (int((nr['owner'] == a).sum()), int((sr['owner'] == a).sum())) for a in range(10)}
        world = C.world_block(np.random.default_rng(41), counts)
        layout17 = L.ns_checkpoint_layout_ex(17)
        assert layout17 not in (0, L.ns_checkpoint_layout())
        path = tmp_path / 'world17.checkpoint.gz'
        head9 = ns_amd.CheckpointHeader.from_buffer_copy(C.header(L.ns_checkpoint_layout(), step=8, initial=12))
        assert L.ns_pool_save_checkpoint(A, str(path).encode(), ct.byref(head9), world, len(world)) == -1
        head = ns_amd.CheckpointHeader.from_buffer_copy(C.header(layout17, step=8, initial=12))
        assert L.ns_pool_save_checkpoint(A, str(path).encode(), ct.byref(head), world, len(world)) == 0, \
            ns_amd.error()
        nine = L.ns_pool_create(P, C.NEURON_CAPACITY, C.SYNAPSE_CAPACITY)
        try:
            assert L.ns_pool_load_checkpoint(nine, str(path).encode(), None, None, 0, None) == -1
            assert 'ring length' in ns_amd.error()
        finally:
            L.ns_pool_destroy(nine)
        assert L.ns_pool_load_checkpoint(R, str(path).encode(), None, None, 0, None) == 0, ns_amd.error()
        for k in range(8, 12):
            assert np.array_equal(step(A, k)[0], step(R, k)[0]), f'motors at step {k}'
        xa, xr = export17(A), export17(R)
        assert _same_records(xa[0], xr[0]) and _same_records(xa[1], xr[1])
    finally:
        L.ns_pool_destroy(A)
        L.ns_pool_destroy(R)
