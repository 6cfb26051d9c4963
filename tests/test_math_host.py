"""The device expf/logf (ns_math.h) compiled for the host agree bit-for-bit
with this machine's libm: exhaustive over all 2^32 floats (expf) and all
2^31 non-negative floats (logf).  The same source runs on the device; the
GPU test tests/test_gpu_math.py checks the device build."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope='module')
def checker(tmp_path_factory):
    exe = tmp_path_factory.mktemp('math') / 'math_check'
    subprocess.run(['g++', '-O2', '-std=c++17', '-I', str(ROOT / 'neural-selection_amd' / 'csrc'),
                    str(ROOT / 'tests' / 'helpers' / 'math_check.cpp'), '-o', str(exe),
                    '-lpthread'], check=True)
    return exe


@pytest.mark.parametrize('fn', ['expf', 'logf'])
def test_device_math_matches_libm_exhaustively(checker, fn):
    out = subprocess.run([str(checker), fn], check=True, capture_output=True, text=True).stdout
    bad, first = out.split()
    assert int(bad) == 0, f'{fn}: {bad} inputs differ from libm, first {first}'
