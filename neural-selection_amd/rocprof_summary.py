"""Summarise rocprofv3 CSV output for one tick kernel into a short markdown
file (the figures bench.py reports beside its own HIP-event timing).

    python neural-selection_amd/rocprof_summary.py --kernel k_resident_lds \
        --trace DIR_OF_KERNEL_TRACE --last 50 \
        [--fetch DIR --write DIR --copy-fetch DIR --copy-write DIR --copy-mib 1024] \
        [--sq DIR ... --sq-command "..."] \
        --command "..." --pmc-command "..." --workload c3 --tile 1024,1,0,1,2048 \
        --out profiles/k_resident_lds_c3.md

--trace is a `rocprofv3 --kernel-trace --stats --output-format csv` output
directory; the average over the --last dispatches of the kernel that precede
the last --skip (the timed steps of bench.py, after its autotuning and warmup
and before its later legs) is what must
agree with bench.py's `launch_ms`.  The counter directories are `rocprofv3
--pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs (KiB per dispatch); each counter
is scaled by bytes / counter reading of a device-to-device copy of --copy-mib
MiB, measured the same way, so the traffic figure is in bytes per dispatch.
"""
from __future__ import annotations

import argparse
import csv
import glob
import os
import statistics


def _csv(d, suffix):
    files = glob.glob(os.path.join(d, f'*{suffix}'))
    if not files:
        raise SystemExit(f'no *{suffix} under {d}')
    with open(files[0]) as f:
        return list(csv.DictReader(f))


def _counter_per_dispatch(d, kernel):
    vals = [float(r['Counter_Value']) for r in _csv(d, 'counter_collection.csv')
            if kernel in r['Kernel_Name']]
    if not vals:
        raise SystemExit(f'kernel {kernel} not in {d}')
    return statistics.mean(vals), len(vals)


def _copy_reading(d):
    vals = [float(r['Counter_Value']) for r in _csv(d, 'counter_collection.csv')
            if 'copyBuffer' in r['Kernel_Name']]
    if not vals:
        raise SystemExit(f'no copy kernel in {d}')
    return statistics.median(vals)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--kernel', required=True)
    ap.add_argument('--trace', required=True)
    ap.add_argument('--last', type=int, default=50)
    ap.add_argument('--skip', type=int, default=0,
                    help='dispatches after the timed window (later legs of the same run)')
    ap.add_argument('--fetch')
    ap.add_argument('--write')
    ap.add_argument('--copy-fetch')
    ap.add_argument('--copy-write')
    ap.add_argument('--copy-mib', type=int, default=1024)
    ap.add_argument('--command', required=True)
    ap.add_argument('--pmc-command', default='')
    ap.add_argument('--workload', required=True)
    ap.add_argument('--tile', required=True)
    ap.add_argument('--sq', nargs='*', default=[],
                    help='`rocprofv3 --pmc SQ_...` output directories (counters per dispatch)')
    ap.add_argument('--sq-command', default='')
    ap.add_argument('--clock-ghz', type=float, default=2.4,
                    help='shader clock assumed for the SIMD utilisation figure')
    ap.add_argument('--simds', type=int, default=1024, help='SIMDs of the device (256 CUs x 4)')
    ap.add_argument('--out', required=True)
    a = ap.parse_args()

    trace = _csv(a.trace, 'kernel_trace.csv')
    rows = [r for r in trace if a.kernel in r['Kernel_Name']]
    if not rows:
        raise SystemExit(f'kernel {a.kernel} not in {a.trace}')
    rows.sort(key=lambda r: int(r['Start_Timestamp']))
    durs = [(int(r['End_Timestamp']) - int(r['Start_Timestamp'])) / 1e6 for r in rows]
    timed = durs[len(durs) - a.skip - a.last:len(durs) - a.skip]
    name = rows[-1]['Kernel_Name']
    stats = [r for r in _csv(a.trace, 'kernel_stats.csv')]
    top = sorted(stats, key=lambda r: -float(r['Percentage']))[:8]

    lines = [f'# {a.kernel}, workload {a.workload}, tile {a.tile}', '',
             f'Kernel: `{name}`', '', '## Kernel trace', '',
             f'Command: `{a.command}`', '',
             f'- dispatches of this kernel: {len(durs)} (all); the timed steps are the {len(timed)} '
             f'before the last {a.skip}' if a.skip else
             f'- dispatches of this kernel: {len(durs)} (all), last {len(timed)} = the timed steps',
             f'- average duration of the timed steps: {statistics.mean(timed):.4f} ms '
             f'(median {statistics.median(timed):.4f}, min {min(timed):.4f}, max {max(timed):.4f})',
             '', '`--stats` top kernels (all dispatches of the run, autotuning included):', '',
             '| kernel | calls | average ms | % |', '|---|---|---|---|']
    for r in top:
        lines.append(f"| `{r['Name'][:70]}` | {r['Calls']} | {float(r['AverageNs']) / 1e6:.4f} | "
                     f"{float(r['Percentage']):.2f} |")
    if a.fetch and a.write and a.copy_fetch and a.copy_write:
        copy_bytes = a.copy_mib * 2**20
        f_read, nf = _counter_per_dispatch(a.fetch, a.kernel)
        w_read, nw = _counter_per_dispatch(a.write, a.kernel)
        f_scale = copy_bytes / _copy_reading(a.copy_fetch)
        w_scale = copy_bytes / _copy_reading(a.copy_write)
        rd, wr = f_read * f_scale, w_read * w_scale
        lines += ['', '## HBM traffic (`rocprofv3 --pmc`, one counter per run)', '',
                  f'Command: `{a.pmc_command}`', '',
                  f'- FETCH_SIZE {f_read:.0f} KiB per dispatch (mean of {nf}); a {a.copy_mib} MiB '
                  f'device copy reads {_copy_reading(a.copy_fetch):.0f} KiB, i.e. {f_scale:.2f} bytes '
                  f'per counter unit -> {rd / 1e9:.3f} GB read',
                  f'- WRITE_SIZE {w_read:.0f} KiB per dispatch (mean of {nw}); the same copy writes '
                  f'{_copy_reading(a.copy_write):.0f} KiB, i.e. {w_scale:.2f} bytes per counter '
                  f'unit -> {wr / 1e9:.3f} GB written',
                  '', f'bytes_per_launch: {rd + wr:.0f}']
    if a.sq:
        sq = {}
        for d in a.sq:
            for r in _csv(d, 'counter_collection.csv'):
                if a.kernel in r['Kernel_Name']:
                    sq.setdefault(r['Counter_Name'], []).append(float(r['Counter_Value']))
        mean = {k: statistics.mean(v) for k, v in sq.items()}
        lines += ['', '## SQ counters (`rocprofv3 --pmc`, per dispatch of this kernel)', '',
                  f'Command: `{a.sq_command}`', '', '| counter | mean per dispatch | dispatches |',
                  '|---|---|---|']
        for k in sorted(mean):
            lines.append(f'| {k} | {mean[k]:.4g} | {len(sq[k])} |')
        waves = mean.get('SQ_WAVES')
        if waves:
            lines += ['', 'Per wave (counter / SQ_WAVES):', '']
            for k, label in (('SQ_INSTS_VALU', 'VALU instructions'), ('SQ_INSTS_SALU', 'SALU instructions'),
                             ('SQ_INSTS_LDS', 'LDS instructions'), ('SQ_INSTS_VMEM_RD', 'vector memory reads'),
                             ('SQ_INSTS_VMEM_WR', 'vector memory writes')):
                if k in mean:
                    lines.append(f'- {label}: {mean[k] / waves:.0f}')
        if 'SQ_INSTS_VALU' in mean and timed:
            cycles = statistics.mean(timed) * 1e-3 * a.clock_ghz * 1e9
            util = mean['SQ_INSTS_VALU'] * 4 / (a.simds * cycles)
            lines += ['', f'- VALU issue share: {util:.2f} of the SIMD cycles of the average timed launch '
                          f'(each wave64 VALU instruction holds a SIMD for 4 cycles; {a.simds} SIMDs at '
                          f'{a.clock_ghz} GHz over {statistics.mean(timed):.3f} ms)']
        if 'SQ_LDS_BANK_CONFLICT' in mean and 'SQ_LDS_IDX_ACTIVE' in mean:
            lines.append(f'- LDS bank conflict cycles / LDS active cycles: '
                         f'{mean["SQ_LDS_BANK_CONFLICT"] / mean["SQ_LDS_IDX_ACTIVE"]:.3f}')
    with open(a.out, 'w') as f:
        f.write('\n'.join(lines) + '\n')
    print('\n'.join(lines))


if __name__ == '__main__':
    main()
