"""The tick kernels replace float division by a per-neuron constant c with
the double product x * RN(1/c) rounded once to float (falling back to the
division for nonzero quotients below 2^-126), and 5/(1+e) by a Newton
reciprocal in double.  Both must equal IEEE float division for every float
numerator; checked here over all 2^32 bit patterns for the divisors the
profiles use (src/physiology.cuh:82-163, branch slope 0.4) and for every
denominator 1+e of the dendritic transfer."""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_division_shortcuts_are_correctly_rounded(tmp_path):
    exe = tmp_path / 'div_check'
    subprocess.run(['g++', '-O2', '-std=c++17', str(ROOT / 'tests' / 'helpers' / 'div_check.cpp'),
                    '-o', str(exe), '-lpthread'], check=True)
    divisors = ['0.4', '104', '59', '10', '8.7', '1200', '0.8', '3.0', '88', '16', '0.25', '7.3e-3',
                '3.1e5']
    out = subprocess.run([str(exe), *divisors], check=True, capture_output=True, text=True).stdout
    bad, first = out.split()
    assert int(bad) == 0, f'{bad} mismatches, first numerator {first}'
