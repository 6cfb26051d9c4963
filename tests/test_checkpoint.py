"""NSCHKP01 checkpoint files (src/checkpoint_impl.cuh) through libns_amd's
host-only reader and writer, against the container restated in
oracle/checkpoint.py: sizes the compiler lays out, the layout hash, files
written by one side read by the other, byte-identical rewrites, and the
reference's rejections (tests/checkpoint_roundtrip.py: a truncated file is
refused and nothing changes).  No GPU calls."""
import ctypes as ct
import gzip

import numpy as np
import pytest

import ns_amd
import oracle as O
import checkpoint as C


def _layout():
    out = np.zeros(10, np.uint64)
    O.oracle_lib().ora_world_layout(out.ctypes.data_as(O._u64))
    return [int(x) for x in out]


def test_world_struct_sizes_are_the_compilers():
    animal, alive, ncount, scount, totals, life, bite, cohort, behaviour, head = _layout()
    assert (animal, alive, ncount, scount) == (C.ANIMAL_BYTES, C.ALIVE_OFFSET, C.NEURON_COUNT_OFFSET,
                                               C.SYNAPSE_COUNT_OFFSET)
    assert (totals, life, bite, cohort) == (C.TOTALS_BYTES, C.LIFE_BYTES, C.BITE_BYTES, C.COHORT_BYTES)
    assert behaviour == 276 and head == C.HEADER.size == ct.sizeof(ns_amd.CheckpointHeader) == 608


def test_layout_hash_matches_the_restatement():
    animal, _, _, _, totals, life, bite, _, _, _ = _layout()
    values = C.layout_values(animal, O.NEURON_GENE.itemsize, O.NEURON_STATE.itemsize,
                             O.SYNAPSE_GENE.itemsize, O.SYNAPSE_STATE.itemsize, C.NEURON_RECORD.itemsize,
                             C.SYNAPSE_RECORD.itemsize, totals, life, bite)
    assert values[10:] == [472, 124, 120, 40, 80, 344, 136, 528, 504, 24]
    h = C.layout_hash(values)
    assert ns_amd.lib().ns_checkpoint_layout() == h
    # each input moves the hash (a dropped or swapped value would show)
    for i in range(len(values)):
        moved = list(values)
        moved[i] += 1
        assert C.layout_hash(moved) != h
    assert C.layout_hash(values[::-1]) != h


def test_17_slot_record_layout_and_hash():
    """The layout version of 1-16 ms delays: the compiler's CheckpointNeuron
    with HISTORY 17 is 352 bytes with the ring at the same offset; its hash
    is checkpoint_layout over HISTORY 17 and that size; other ring lengths
    have none."""
    sizes = np.zeros(3, np.uint64)
    O.oracle_lib().ora_checkpoint_neuron_layout(sizes.ctypes.data_as(O._u64))
    assert [int(x) for x in sizes] == [C.NEURON_RECORD.itemsize, C.NEURON_RECORD17.itemsize, 328]
    animal, _, _, _, totals, life, bite, _, _, _ = _layout()
    base = (animal, O.NEURON_GENE.itemsize, O.NEURON_STATE.itemsize, O.SYNAPSE_GENE.itemsize,
            O.SYNAPSE_STATE.itemsize)
    L = ns_amd.lib()
    h9 = C.layout_hash(C.layout_values(*base, 344, 136, totals, life, bite, history=9))
    h17 = C.layout_hash(C.layout_values(*base, 352, 136, totals, life, bite, history=17))
    assert L.ns_checkpoint_layout_ex(9) == L.ns_checkpoint_layout() == h9
    assert L.ns_checkpoint_layout_ex(17) == h17 != h9
    assert L.ns_checkpoint_layout_ex(8) == 0 and L.ns_checkpoint_layout_ex(33) == 0


def records(rng, counts):
    """Neuron and synapse records for animals with counts[a] = (n, s):
    slots interleaved across animals, ascending; random genes and state."""
    total_n = sum(n for n, _ in counts.values())
    total_s = sum(s for _, s in counts.values())
    nslots = np.sort(rng.choice(4 * total_n + 10, total_n, replace=False))
    sslots = np.sort(rng.choice(4 * total_s + 10, total_s, replace=False))
    nowner = np.concatenate([[a] * n for a, (n, _) in counts.items()])
    sowner = np.concatenate([[a] * s for a, (_, s) in counts.items()])
    rng.shuffle(nowner)
    rng.shuffle(sowner)
    n = np.frombuffer(rng.bytes(total_n * 344), C.NEURON_RECORD).copy()
    n['slot'] = nslots
    n['owner'] = nowner
    n['next'] = -1
    n['port'] = rng.integers(-1, 24, total_n)
    s = np.frombuffer(rng.bytes(total_s * 136), C.SYNAPSE_RECORD).copy()
    s['slot'] = sslots
    s['owner'] = sowner
    s['next'] = rng.integers(-1, 100, total_s)
    s['pre'] = rng.choice(nslots, total_s)
    s['post'] = rng.choice(nslots, total_s)
    return n, s


def case(seed=1, cohort=False):
    rng = np.random.default_rng(seed)
    counts = {3: (5, 9), 17: (8, 20), 32767: (3, 4), 900: (6, 0)}
    n, s = records(rng, counts)
    world = C.world_block(rng, counts, cohort=cohort)
    head = C.header(ns_amd.lib().ns_checkpoint_layout(), flags=65 if cohort else 1,
                    neurons=len(n), synapses=len(s), initial=1000)
    return head, world, n, s


def lib_save(path, head, world, n, s):
    h = ns_amd.CheckpointHeader.from_buffer_copy(head)
    return ns_amd.lib().ns_checkpoint_save(str(path).encode(), ct.byref(h), world, len(world),
                                           n.ctypes.data, len(n), s.ctypes.data, len(s))


def lib_open(path):
    L = ns_amd.lib()
    ck = L.ns_checkpoint_open(str(path).encode())
    if not ck:
        return None
    try:
        h = ns_amd.CheckpointHeader()
        w, wb, pn, ps = ct.c_void_p(), ct.c_size_t(), ct.c_void_p(), ct.c_void_p()
        assert L.ns_checkpoint_get(ck, ct.byref(h), ct.byref(w), ct.byref(wb), ct.byref(pn), ct.byref(ps)) == 0
        world = ct.string_at(w, wb.value)
        n = np.frombuffer(ct.string_at(pn, h.neuron_count * 344), C.NEURON_RECORD).copy() if h.neuron_count else \
            np.zeros(0, C.NEURON_RECORD)
        s = np.frombuffer(ct.string_at(ps, h.synapse_count * 136), C.SYNAPSE_RECORD).copy() if h.synapse_count else \
            np.zeros(0, C.SYNAPSE_RECORD)
        return bytes(h), world, n, s
    finally:
        L.ns_checkpoint_close(ck)


@pytest.mark.parametrize('cohort', [False, True])
def test_library_file_reads_back_in_both_readers(tmp_path, cohort):
    head, world, n, s = case(cohort=cohort)
    path = tmp_path / 'world.checkpoint.gz'
    assert lib_save(path, head, world, n, s) == 0, ns_amd.error()
    assert not (tmp_path / 'world.checkpoint.gz.tmp').exists()
    raw = gzip.decompress(path.read_bytes())
    assert raw == C.compose(head, world, n, s)                 # the restated layout, byte for byte
    h, w2, n2, s2 = C.parse(raw)
    assert w2 == world and n2.tobytes() == n.tobytes() and s2.tobytes() == s.tobytes()
    got = lib_open(path)
    assert got is not None, ns_amd.error()
    assert got[0] == head and got[1] == world
    assert got[2].tobytes() == n.tobytes() and got[3].tobytes() == s.tobytes()
    # rewriting what was read gives the same file
    again = tmp_path / 'again.gz'
    assert lib_save(again, *got) == 0
    assert again.read_bytes() == path.read_bytes()


def test_restated_writer_file_opens_in_the_library(tmp_path):
    head, world, n, s = case(seed=4)
    path = tmp_path / 'py.gz'
    C.write(path, head, world, n, s)
    got = lib_open(path)
    assert got is not None, ns_amd.error()
    assert got[0] == head and got[1] == world and got[2].tobytes() == n.tobytes()


def test_header_counts_come_from_the_records(tmp_path):
    head, world, n, s = case(seed=5)
    wrong = C.header(ns_amd.lib().ns_checkpoint_layout(), flags=1, neurons=0, synapses=999,
                     initial=1000)
    path = tmp_path / 'c.gz'
    assert lib_save(path, wrong, world, n, s) == 0, ns_amd.error()
    assert lib_open(path)[0] == head


def _expect_open_error(path, message):
    assert lib_open(path) is None
    assert message in ns_amd.error(), ns_amd.error()


def test_rejections_follow_the_reference(tmp_path):
    head, world, n, s = case(seed=6)
    good = C.compose(head, world, n, s)
    cases = {
        'truncated': (gzip.compress(good)[:len(gzip.compress(good)) // 2], 'incomplete or corrupt'),
        'short body': (gzip.compress(good[:-1]), 'incomplete or corrupt'),
        'trailing': (gzip.compress(good + b'\0'), 'Unexpected trailing checkpoint data'),
        'magic': (gzip.compress(b'NSCHKP02' + good[8:]), 'Incompatible or invalid checkpoint header'),
        'layout': (gzip.compress(C.header(123, neurons=len(n), synapses=len(s), initial=1000)
                                 + good[C.HEADER.size:]), 'Incompatible or invalid checkpoint header'),
        'flags': (gzip.compress(C.header(ns_amd.lib().ns_checkpoint_layout(), flags=128, neurons=len(n),
                                         synapses=len(s), initial=1000) + good[C.HEADER.size:]),
                  'Incompatible or invalid checkpoint header'),
    }
    bad_hist = bytearray(good)
    bad_hist[C.HEADER.size + C.TOTALS_BYTES:C.HEADER.size + C.TOTALS_BYTES + 4] = (10003).to_bytes(4, 'little')
    cases['history'] = (gzip.compress(bytes(bad_hist)), 'Invalid checkpoint behavior history')
    # a record owned by a dead animal, and counts that disagree
    n2 = n.copy()
    n2['owner'][0] = 5
    cases['dead owner'] = (gzip.compress(C.compose(head, world, n2, s)), 'Invalid checkpoint neuron record')
    s2 = s.copy()
    s2['slot'][1] = s2['slot'][0]
    cases['slot order'] = (gzip.compress(C.compose(head, world, n, s2)), 'Invalid checkpoint synapse record')
    fewer = C.header(ns_amd.lib().ns_checkpoint_layout(), neurons=len(n) - 1, synapses=len(s),
                     initial=1000)
    cases['counts'] = (gzip.compress(C.compose(fewer, world, n[:-1], s)),
                       'Checkpoint brain allocation counts disagree')
    rng = np.random.default_rng(7)
    counts = {3: (5, 9), 17: (8, 20), 32767: (3, 4), 900: (6, 0)}
    cw = bytearray(C.world_block(rng, counts, cohort=True))
    tag_at = C.observer_bytes(bytes(cw)) + C.ARRAY_BYTES + 16
    cw[tag_at:tag_at + 8] = (81).to_bytes(8, 'little')       # PlasticityCohort size
    ch = C.header(ns_amd.lib().ns_checkpoint_layout(), flags=65, neurons=len(n), synapses=len(s),
                  initial=1000)
    cases['extension'] = (gzip.compress(C.compose(ch, bytes(cw), n, s)),
                          'Incompatible plasticity intervention extension')
    for name, (payload, message) in cases.items():
        path = tmp_path / f'{name.replace(" ", "_")}.gz'
        path.write_bytes(payload)
        _expect_open_error(path, message)


def test_failed_save_writes_nothing(tmp_path):
    head, world, n, s = case(seed=8)
    path = tmp_path / 'keep.gz'
    assert lib_save(path, head, world, n, s) == 0
    before = path.read_bytes()
    n2 = n[:-1]                                  # living animals own one more neuron than given
    assert lib_save(path, head, world, n2, s) == -1
    assert 'ownership disagrees' in ns_amd.error()
    assert lib_save(path, head, world[:-1], n, s) == -1
    assert 'world sections' in ns_amd.error()
    n3 = n.copy()
    n3['slot'][[0, 1]] = n3['slot'][[1, 0]]
    assert lib_save(path, head, world, n3, s) == -1
    assert 'Invalid checkpoint neuron record' in ns_amd.error()
    assert path.read_bytes() == before and not (tmp_path / 'keep.gz.tmp').exists()


def records17(rng, counts):
    n9, s = records(rng, counts)
    n = np.zeros(len(n9), C.NEURON_RECORD17)
    raw9 = n9.view(np.uint8).reshape(len(n9), -1)
    raw = n.view(np.uint8).reshape(len(n), -1)
    raw[:, :328] = raw9[:, :328]
    raw[:, 328:345] = np.frombuffer(rng.bytes(len(n) * 17), np.uint8).reshape(len(n), 17)
    return n, s


def test_17_slot_files_in_both_readers(tmp_path):
    """A file of the 17-slot layout: written by the library as the restated
    container lays it out, read back by either reader; the 9-slot calls
    refuse it and a header of the other layout is refused on save."""
    L = ns_amd.lib()
    rng = np.random.default_rng(9)
    counts = {3: (5, 9), 17: (8, 20), 32767: (3, 4), 900: (6, 0)}
    n, s = records17(rng, counts)
    world = C.world_block(rng, counts)
    head = C.header(L.ns_checkpoint_layout_ex(17), neurons=len(n), synapses=len(s), initial=1000)
    path = tmp_path / 'w17.gz'
    h = ns_amd.CheckpointHeader.from_buffer_copy(head)
    assert L.ns_checkpoint_save17(str(path).encode(), ct.byref(h), world, len(world), n.ctypes.data, len(n),
                                  s.ctypes.data, len(s)) == 0, ns_amd.error()
    raw = gzip.decompress(path.read_bytes())
    assert raw == C.compose(head, world, n, s)
    _, w2, n2, s2 = C.parse(raw, C.NEURON_RECORD17)
    assert w2 == world and n2.tobytes() == n.tobytes() and s2.tobytes() == s.tobytes()
    ck = L.ns_checkpoint_open(str(path).encode())
    assert ck, ns_amd.error()
    try:
        assert L.ns_checkpoint_history(ck) == 17
        got = ns_amd.CheckpointHeader()
        w, wb, pn, ps = ct.c_void_p(), ct.c_size_t(), ct.c_void_p(), ct.c_void_p()
        assert L.ns_checkpoint_get(ck, ct.byref(got), ct.byref(w), ct.byref(wb), ct.byref(pn), ct.byref(ps)) == -1
        assert '17-slot' in ns_amd.error()
        assert L.ns_checkpoint_get17(ck, ct.byref(got), ct.byref(w), ct.byref(wb), ct.byref(pn),
                                     ct.byref(ps)) == 0
        assert bytes(got) == head and ct.string_at(w, wb.value) == world
        assert ct.string_at(pn, len(n) * 352) == n.tobytes() and ct.string_at(ps, len(s) * 136) == s.tobytes()
    finally:
        L.ns_checkpoint_close(ck)
    # the 9-slot calls on a 9-slot file, and each layout's header on the other's records
    head9 = C.header(L.ns_checkpoint_layout(), neurons=len(n), synapses=len(s), initial=1000)
    h9 = ns_amd.CheckpointHeader.from_buffer_copy(head9)
    assert L.ns_checkpoint_save17(str(tmp_path / 'x.gz').encode(), ct.byref(h9), world, len(world),
                                  n.ctypes.data, len(n), s.ctypes.data, len(s)) == -1
    assert 'Incompatible or invalid checkpoint header' in ns_amd.error()
    assert lib_save(tmp_path / 'y.gz', head, world, n, s) == -1          # 17-slot layout, 9-slot call
    assert 'Incompatible or invalid checkpoint header' in ns_amd.error()
    assert not (tmp_path / 'x.gz').exists() and not (tmp_path / 'y.gz').exists()


def test_failed_write_leaves_no_temporary_file(tmp_path):
    """A save whose final rename fails (the path is a directory) reports the
    error and removes path.tmp."""
    head, world, n, s = case(seed=10)
    target = tmp_path / 'busy'
    target.mkdir()
    (target / 'inside').write_text('x')
    assert lib_save(target, head, world, n, s) == -1
    assert not (tmp_path / 'busy.tmp').exists()
    assert (target / 'inside').read_text() == 'x'
