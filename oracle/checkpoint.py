"""CPU restatement of the NSCHKP01 checkpoint container (src/checkpoint_impl.cuh)
— TEST INFRASTRUCTURE.  Only tests/ import it, as the checker of the
library's checkpoint reader and writer.

File = gzip of: CheckpointHeader (checkpoint_impl.cuh:6-19, 608 B), the
observer section (:85-107), the world arrays (:173-190), the neuron and
synapse records (:191-212), the plasticity-cohort extension when flags & 64
(:213-217), nothing after (:362-364).
"""
from __future__ import annotations

import gzip
import struct

import numpy as np

# src/main.cu:33-47, src/physiology.cuh:17
P, NEURON_CAPACITY, SYNAPSE_CAPACITY = 32768, 4194304, 8388608
S, M, W, H, DENDRITES, HISTORY, TICKS = 19, 5, 128, 128, 4, 9, 16
# struct sizes and Animal field offsets (src/main.cu:66-104, 1845-1849);
# the oracle library's ora_world_layout has the compiler lay the same
# field lists out, and the tests compare.
ANIMAL_BYTES, TOTALS_BYTES, LIFE_BYTES, BITE_BYTES, COHORT_BYTES = 472, 528, 504, 24, 80
ALIVE_OFFSET, NEURON_COUNT_OFFSET, SYNAPSE_COUNT_OFFSET = 80, 92, 96
HEADER = struct.Struct('<8sIIQqqQiifIIIQQQ512s')              # 608 B
COHORT_TAG = 0x4e53504c41535431                                # checkpoint_impl.cuh:214
ARRAY_BYTES = (P * ANIMAL_BYTES + 2 * W * H * 4 + P * S * 4 + P * M * 4 + 5 * 4 + M * 4 + 4
               + 6 * M * 4)

# Record dtypes covering every byte (padding included: numpy copies
# structured records field by field and would leave holes unset).
NEURON_RECORD = np.dtype({
    'names': ['slot', 'owner', 'next', 'port', 'gene', 'state', 'pending_exc', 'pending_inh',
              'mod_event_tick', 'spike_history', 'pad'],
    'formats': ['<i4', '<i4', '<i4', '<i4', ('u1', 124), ('u1', 120), ('<f4', 4), '<f4',
                ('<u8', 6), ('u1', 9), ('u1', 7)],
    'offsets': [0, 4, 8, 12, 16, 140, 260, 276, 280, 328, 337], 'itemsize': 344})
# The record of a build with 1-16 ms delays: main.cu:46 HISTORY = D + 1 = 17
# ring slots, 345 bytes of fields padded to the 8-byte alignment of the
# mod_event_tick array: 352.
HISTORY17 = 17
NEURON_RECORD17 = np.dtype({
    'names': list(NEURON_RECORD.names),
    'formats': ['<i4', '<i4', '<i4', '<i4', ('u1', 124), ('u1', 120), ('<f4', 4), '<f4',
                ('<u8', 6), ('u1', 17), ('u1', 7)],
    'offsets': [0, 4, 8, 12, 16, 140, 260, 276, 280, 328, 345], 'itemsize': 352})
SYNAPSE_RECORD = np.dtype({
    'names': ['slot', 'owner', 'next', 'gene', 'pre', 'post', 'pad', 'state'],
    'formats': ['<i4', '<i4', '<i4', ('u1', 40), '<i4', '<i4', '<u4', ('u1', 80)],
    'offsets': [0, 4, 8, 12, 28, 32, 52, 56], 'itemsize': 136})


def layout_hash(values):
    """checkpoint_impl.cuh:21-32: FNV-1a offset basis and prime over whole
    64-bit values (not bytes)."""
    h = 1469598103934665603
    for v in values:
        h ^= int(v)
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def layout_values(animal, neuron_gene, neuron_state, synapse_gene, synapse_state, ck_neuron, ck_synapse,
                  totals, life, bite, history=HISTORY):
    return [P, NEURON_CAPACITY, SYNAPSE_CAPACITY, S, M, W, H, DENDRITES, history, TICKS, animal,
            neuron_gene, neuron_state, synapse_gene, synapse_state, ck_neuron, ck_synapse, totals, life,
            bite]


def header(layout, step=100, every=100, seed=2, initial=12, trace=-1, renewal=0.5, flags=1,
           neurons=0, synapses=0, sizes=(10, 20, 30), output=b'run.csv'):
    return HEADER.pack(b'NSCHKP01', 1, 0x01020304, layout, step, every, seed, initial, trace, renewal,
                       flags, neurons, synapses, *sizes, output)


def world_block(rng, counts, behaviour=3, bites=2, cohort=False):
    """Observer section + world arrays (+ extension) with random contents;
    animal a is alive with counts[a] = (neurons, synapses), every other
    animal is dead."""
    out = bytearray()
    out += rng.bytes(TOTALS_BYTES)
    out += struct.pack('<I', behaviour) + rng.bytes(behaviour * (8 + TOTALS_BYTES))
    out += struct.pack('<I', bites) + rng.bytes(bites * (8 + BITE_BYTES))
    out += rng.bytes(8)
    animals = np.frombuffer(rng.bytes(P * ANIMAL_BYTES), np.uint8).copy().reshape(P, ANIMAL_BYTES)
    alive = np.zeros(P, np.int32)
    nc = np.zeros(P, np.int32)
    sc = np.zeros(P, np.int32)
    for a, (n, s) in counts.items():
        alive[a], nc[a], sc[a] = 1, n, s
    animals[:, ALIVE_OFFSET:ALIVE_OFFSET + 4] = alive.view(np.uint8).reshape(P, 4)
    mask = alive == 1
    animals[mask, NEURON_COUNT_OFFSET:NEURON_COUNT_OFFSET + 4] = nc[mask].view(np.uint8).reshape(-1, 4)
    animals[mask, SYNAPSE_COUNT_OFFSET:SYNAPSE_COUNT_OFFSET + 4] = sc[mask].view(np.uint8).reshape(-1, 4)
    out += animals.tobytes()
    out += rng.bytes(ARRAY_BYTES - P * ANIMAL_BYTES)
    if cohort:
        out += struct.pack('<QQQQ', COHORT_TAG, 1, COHORT_BYTES, P) + rng.bytes(P * COHORT_BYTES)
    return bytes(out)


def observer_bytes(world):
    off = TOTALS_BYTES
    (behaviour,) = struct.unpack_from('<I', world, off)
    off += 4 + behaviour * (8 + TOTALS_BYTES)
    (bites,) = struct.unpack_from('<I', world, off)
    return off + 4 + bites * (8 + BITE_BYTES) + 8


def compose(head, world, neurons, synapses):
    """The uncompressed file: header, observer + arrays, records, extension."""
    flags = HEADER.unpack(head)[10]
    split = observer_bytes(world) + ARRAY_BYTES
    body = head + world[:split] + neurons.tobytes() + synapses.tobytes()
    if flags & 64:
        body += world[split:]
    return body


def write(path, head, world, neurons, synapses):
    with gzip.open(path, 'wb', compresslevel=1) as f:
        f.write(compose(head, world, neurons, synapses))


def parse(raw, neuron_record=NEURON_RECORD):
    """Split an uncompressed file into (header tuple, world block, neuron
    records, synapse records) in the reader's order; neuron_record is the
    layout version's record (NEURON_RECORD or NEURON_RECORD17)."""
    h = HEADER.unpack_from(raw, 0)
    off = HEADER.size
    obs = observer_bytes(raw[off:])
    world = raw[off:off + obs + ARRAY_BYTES]
    off += obs + ARRAY_BYTES
    nn, ns = h[11], h[12]
    neurons = np.frombuffer(raw, neuron_record, nn, off).copy()
    off += nn * neuron_record.itemsize
    synapses = np.frombuffer(raw, SYNAPSE_RECORD, ns, off).copy()
    off += ns * SYNAPSE_RECORD.itemsize
    if h[10] & 64:
        world += raw[off:off + 32 + P * COHORT_BYTES]
        off += 32 + P * COHORT_BYTES
    assert off == len(raw), 'trailing bytes'
    return h, world, neurons, synapses
