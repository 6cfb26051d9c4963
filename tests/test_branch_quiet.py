"""The dendritic transfer (src/physiology.cuh:226-230) returns exactly +0 for
every conductance in [0, 2^-37), which lets the tick kernels skip it for quiet
dendrites.  Checked on the oracle's restatement over a dense stride of that
range plus its end points, and against the reference's own function where
oracle/_ref is built; the first conductance above the range is shown to give a
nonzero output, so the bound is the last one this argument allows."""
import struct

import numpy as np
import pytest

import oracle as O

QUIET = 2.0 ** -37


def bits(x):
    return struct.unpack('<I', struct.pack('<f', x))[0]


def as_float(b):
    return struct.unpack('<f', struct.pack('<I', b))[0]


def test_quiet_range_gives_positive_zero():
    lib = O.oracle_lib()
    hi = bits(QUIET)
    assert lib.ora_branch_io_nonzero(0, hi, 997) == 0
    for b in (0, 1, 2, 0x007fffff, 0x00800000, hi - 2, hi - 1, hi):
        y = lib.ora_fn_branch_io(as_float(b))
        assert bits(y) == 0, f'branch_io({as_float(b)!r}) = {y!r}'
    # -0 behaves like +0
    assert bits(lib.ora_fn_branch_io(-0.0)) == 0
    # just above the half ulp of the offset the sum no longer rounds back
    assert lib.ora_fn_branch_io(as_float(hi + 1)) > 0
    assert lib.ora_fn_branch_io(2.0 ** -30) > 0


@pytest.mark.skipif(not O.ref_available(), reason='oracle/_ref not built (reference sources absent)')
def test_quiet_range_matches_reference():
    ora, ref = O.oracle_lib(), O.ref_lib()
    rng = np.random.default_rng(5)
    hi = bits(QUIET)
    sample = np.concatenate([rng.integers(0, hi, 20000), [0, 1, hi - 1, hi, hi + 1, hi + 2]])
    for b in sample:
        c = as_float(int(b))
        r = ref.ref_fn_dendritic_transfer(c)
        o = ora.ora_fn_branch_io(c)
        assert bits(r) == bits(o), f'c={c!r}: ref {r!r} oracle {o!r}'
        if int(b) <= hi:
            assert bits(r) == 0
    # and over ordinary conductances
    for c in rng.lognormal(-1, 2, 5000).astype(np.float32):
        assert bits(ref.ref_fn_dendritic_transfer(float(c))) == bits(ora.ora_fn_branch_io(float(c)))
