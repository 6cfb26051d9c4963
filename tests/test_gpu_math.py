"""The device build of ns_expf / ns_logf equals this host's libm on every
input: per-chunk digests of the device results (ns_device_math_digest)
against the same digests of libm computed on the host."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import ns_amd

ROOT = Path(__file__).resolve().parents[1]
pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def checker(tmp_path_factory):
    exe = tmp_path_factory.mktemp('math') / 'math_check'
    subprocess.run(['g++', '-O2', '-std=c++17', '-I', str(ROOT / 'neural-selection_amd' / 'csrc'),
                    str(ROOT / 'tests' / 'helpers' / 'math_check.cpp'), '-o', str(exe),
                    '-lpthread'], check=True)
    return exe


@pytest.mark.parametrize('fn,name,hi', [(0, 'expf', 1 << 32), (1, 'logf', 1 << 31)])
def test_device_math_digests_match_host_libm(checker, fn, name, hi):
    chunk = 1 << 22
    print(f'[math] {name} device digests over [0, {hi:#x})', flush=True)
    dev = ns_amd.device_math_digest(fn, 0, hi, chunk)
    out = subprocess.run([str(checker), name, 'digest', '0', str(hi), str(chunk)], check=True,
                         capture_output=True, text=True).stdout.split()
    host = np.array([int(x) for x in out], np.uint64)
    bad = np.flatnonzero(dev != host)
    assert len(bad) == 0, f'{name}: {len(bad)} chunks differ, first chunk {bad[:4]} (x{chunk})'
