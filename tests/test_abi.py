"""libns_amd.so loads, exports every function include/ns_amd.h declares,
and its host-side genome parsing reports the reference's errors.  No GPU
calls (these run on the CPU)."""
import ctypes as ct
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import ns_amd

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / 'include' / 'ns_amd.h'


def declared_functions():
    text = re.sub(r'/\*.*?\*/', '', HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r'\b(ns_[a-z0-9_]+)\s*\(', text)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 30
    out = subprocess.run(['nm', '-D', '--defined-only', str(ns_amd.LIB_PATH)], check=True,
                         capture_output=True, text=True).stdout
    exported = set(re.findall(r' T (ns_[a-z0-9_]+)', out))
    missing = [n for n in names if n not in exported]
    assert not missing, f'missing exports: {missing}'
    lib = ns_amd.lib()
    for n in names:
        assert hasattr(lib, n)
    assert set(ns_amd.SIGNATURES) == set(names)


def _genome(n=4):
    neurons = np.array([[0, 0, 1, 0, 0, 1, 1, -52, 0, 0]] * n, np.float32)
    syn = np.array([[0, 1, 1.0, 1, 3, 0, .5, 450, 0, 0, 0]], np.float32)
    return neurons, syn, np.array([n - 1], np.int32)


@pytest.mark.parametrize('mutate,message', [
    (lambda n, s, m: (n, s, m, 0), 'Invalid brain dimensions'),
    (lambda n, s, m: (n.__setitem__((0, 0), 7) or n, s, m, 1), 'Invalid neuron gene'),
    (lambda n, s, m: (n.__setitem__((0, 7), -90) or n, s, m, 1), 'Invalid membrane parameters'),
    (lambda n, s, m: (n, s.__setitem__((0, 4), 9) or s, m, 1), 'Invalid synapse gene'),
    (lambda n, s, m: (n, s.__setitem__((0, 5), 4) or s, m, 1), 'Invalid synapse gene'),
    (lambda n, s, m: (n, s, np.array([9], np.int32), 1), 'Invalid motor index'),
])
def test_ns_create_validates_like_the_reference(mutate, message):
    lib = ns_amd.lib()
    n, s, m = _genome()
    n, s, m, outputs = mutate(n, s, m)
    f = ct.POINTER(ct.c_float)
    h = lib.ns_create(len(n), len(s), 1, outputs, n.ctypes.data_as(f), s.ctypes.data_as(f),
                      m.ctypes.data_as(ct.POINTER(ct.c_int)))
    assert not h
    assert ns_amd.error() == message


def test_ns_create_ex_accepts_long_delays_and_destroys():
    lib = ns_amd.lib()
    n, s, m = _genome()
    s[0, 4] = 16
    f = ct.POINTER(ct.c_float)
    assert not lib.ns_create(len(n), 1, 1, 1, n.ctypes.data_as(f), s.ctypes.data_as(f),
                             m.ctypes.data_as(ct.POINTER(ct.c_int)))
    h = lib.ns_create_ex(len(n), 1, 1, 1, n.ctypes.data_as(f), s.ctypes.data_as(f),
                         m.ctypes.data_as(ct.POINTER(ct.c_int)), 16)
    assert h
    idx = np.arange(4, dtype=np.uint32)
    lib.ns_original_indices(h, idx.ctypes.data_as(ct.POINTER(ct.c_uint32)),
                            idx.ctypes.data_as(ct.POINTER(ct.c_uint32)))
    lib.ns_destroy(h)


def test_empty_batch_is_rejected_without_a_device():
    lib = ns_amd.lib()
    assert not lib.ns_gpu_create(None, 0)
    assert ns_amd.error() == 'Empty GPU batch'
