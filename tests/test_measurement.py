"""The measurement chain on the CPU: the roofline's bytes per unit of work
come from include/ns_amd.h, bench.py refuses more than one GPU before any GPU
call, and rocprof_summary.py turns rocprofv3 CSV output into the figures
bench.py reads back (timed window, per-dispatch HBM bytes scaled by a device
copy of known size), only for the workload and tile they were taken on."""
import csv
import re
import statistics
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402

SUMMARY = ROOT / 'neural-selection_amd' / 'rocprof_summary.py'
KERNEL = 'void nsa::k_resident_lds<1024, 2, 1, true>(nsa::Pool, ...)'


def test_roofline_bytes_are_the_headers():
    text = (ROOT / 'include' / 'ns_amd.h').read_text()
    want = {m.group(1): int(m.group(2)) for m in re.finditer(r'#define NS_BYTES_PER_(\w+) (\d+)', text)}
    assert set(want) == {'NEURON_UPDATE', 'SYNAPSE_UPDATE_PLASTIC', 'SYNAPSE_UPDATE_STATIC', 'EVENT'}
    assert (bench.NEURON_UPDATE_BYTES, bench.SYNAPSE_UPDATE_BYTES_PLASTIC,
            bench.SYNAPSE_UPDATE_BYTES_STATIC, bench.EVENT_BYTES) == \
        (want['NEURON_UPDATE'], want['SYNAPSE_UPDATE_PLASTIC'], want['SYNAPSE_UPDATE_STATIC'], want['EVENT'])


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / 'bench.py'), '--gpus', '2'], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert 'one GPU only' in r.stderr and r.stdout == ''


def _write(path, rows):
    path.parent.mkdir(parents=True, exist_ok=True)
    with open(path, 'w', newline='') as f:
        w = csv.DictWriter(f, fieldnames=list(rows[0]))
        w.writeheader()
        w.writerows(rows)


def _run_summary(tmp_path, durations_ns, skip, last, fetch, write, copy_fetch, copy_write):
    trace = tmp_path / 'trace'
    start, rows = 1000, []
    for d in durations_ns:                       # written out of order: the tool sorts by start
        rows.append({'Kernel_Name': KERNEL, 'Start_Timestamp': start, 'End_Timestamp': start + d})
        start += d + 500
    rows.append({'Kernel_Name': 'other_kernel', 'Start_Timestamp': 5, 'End_Timestamp': 99999999})
    _write(trace / 'run_kernel_trace.csv', rows[::-1])
    _write(trace / 'run_kernel_stats.csv', [{'Name': KERNEL, 'Calls': len(durations_ns),
                                             'AverageNs': statistics.mean(durations_ns),
                                             'Percentage': 90.0}])
    dirs = {}
    for name, vals, copy in (('fetch', fetch, copy_fetch), ('write', write, copy_write)):
        _write(tmp_path / name / 'run_counter_collection.csv',
               [{'Kernel_Name': KERNEL, 'Counter_Value': v} for v in vals])
        _write(tmp_path / f'copy_{name}' / 'run_counter_collection.csv',
               [{'Kernel_Name': '__amd_rocclr_copyBuffer', 'Counter_Value': v} for v in copy])
        dirs[name] = tmp_path / name
    out = tmp_path / 'profiles' / 'k_resident_lds_c3.md'
    out.parent.mkdir(parents=True, exist_ok=True)
    subprocess.run([sys.executable, str(SUMMARY), '--kernel', 'k_resident_lds', '--trace', str(trace),
                    '--last', str(last), '--skip', str(skip), '--fetch', str(dirs['fetch']),
                    '--write', str(dirs['write']), '--copy-fetch', str(tmp_path / 'copy_fetch'),
                    '--copy-write', str(tmp_path / 'copy_write'), '--copy-mib', '1024',
                    '--command', 'cmd', '--pmc-command', 'pmc', '--workload', 'c3',
                    '--tile', '1024,2,0,1,4224', '--out', str(out)],
                   check=True, capture_output=True, timeout=120)
    return out.read_text()


def test_summary_window_scaling_and_read_back(tmp_path, monkeypatch):
    # 10 autotune/warmup launches, 6 timed ones (one slow), 3 of a later leg
    durs = [9_000_000] * 10 + [3_000_000, 3_100_000, 2_900_000, 3_000_000, 10_000_000, 3_000_000] \
        + [11_000_000] * 3
    text = _run_summary(tmp_path, durs, skip=3, last=6, fetch=[1000, 1200], write=[300, 500],
                        copy_fetch=[524300, 524200, 524400], copy_write=[1048576])
    timed = [d / 1e6 for d in durs[10:16]]
    assert f'average duration of the timed steps: {statistics.mean(timed):.4f} ms' in text
    assert f'median {statistics.median(timed):.4f}, min 2.9000, max 10.0000' in text
    # FETCH_SIZE and WRITE_SIZE are scaled by copy bytes / the copy's own reading
    copy = 1024 * 2**20
    expect = 1100 * copy / 524300 + 400 * copy / 1048576
    line = [x for x in text.splitlines() if x.startswith('bytes_per_launch:')][0]
    assert float(line.split(':')[1]) == pytest.approx(expect, abs=1)
    # bench.py reads it back for the same workload and tile, and not for another tile
    monkeypatch.setattr(bench, 'ROOT', tmp_path)
    assert bench.load_traffic('k_resident_lds', 'c3', (1024, 2, 0, 1, 4224)) == pytest.approx(expect, abs=1)
    assert bench.load_traffic('k_resident_lds', 'c3', (1024, 1, 0, 1, 4448)) is None
    assert bench.load_traffic('k_resident_lds', 'c2', (1024, 2, 0, 1, 4224)) is None
