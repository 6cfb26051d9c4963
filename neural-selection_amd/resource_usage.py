"""Per-kernel resource usage of the tick kernels as the compiler reports it
(hipcc -Rpass-analysis=kernel-resource-usage, gfx950, the library's flags):
VGPRs, spills, scratch, LDS and occupancy of every resident-kernel tile, as a
short markdown table.

    python neural-selection_amd/resource_usage.py --out profiles/resource_usage_tick_kernels.md
"""
from __future__ import annotations

import argparse
import re
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
FLAGS = ['-O3', '-std=c++17', '-fPIC', '--offload-arch=gfx950', '-ffp-contract=off', '-fno-fast-math',
         '-fhip-fp32-correctly-rounded-divide-sqrt', '-Xclang', '-target-feature', '-Xclang',
         '-packed-fp32-ops', '-I../include']
FIELDS = ['VGPRs', 'AGPRs', 'ScratchSize [bytes/lane]', 'Occupancy [waves/SIMD]', 'SGPRs Spill',
          'VGPRs Spill', 'LDS Size [bytes/block]']


def demangle_tile(name):
    m = re.search(r'k_resident_ldsILi(\d+)ELi(\d+)ELi(\d+)ELb(\d)E', name)
    if m:
        wg, kn, minw, plastic = m.groups()
        return f'k_resident_lds<{wg}, {kn}, {minw}, {"true" if plastic == "1" else "false"}>'
    m = re.search(r'k_residentILi(\d+)ELi(\d+)ELi(\d+)ELi(\d+)E', name)
    if m:
        return 'k_resident<%s, %s, %s, %s>' % m.groups()
    m = re.search(r'_ZN3nsa\d+(k_\w+?)E', name)
    return m.group(1) if m else name


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--hipcc', default='/opt/rocm/bin/hipcc')
    ap.add_argument('--source', default='csrc/ns_kernels.hip')
    ap.add_argument('--out', required=True)
    a = ap.parse_args()
    cmd = [a.hipcc] + FLAGS + ['--cuda-device-only', '-Rpass-analysis=kernel-resource-usage', '-c',
                               a.source, '-o', '/dev/null']
    text = subprocess.run(cmd, cwd=HERE, capture_output=True, text=True, check=True).stderr
    kernels, cur = [], None
    for line in text.splitlines():
        m = re.search(r'remark: Function Name: (\S+)', line)
        if m:
            cur = {'name': demangle_tile(m.group(1))}
            kernels.append(cur)
            continue
        m = re.search(r'remark:\s+([^:]+(?:\[[^\]]*\])?): (\S+) \[-Rpass', line)
        if m and cur is not None and m.group(1).strip() in FIELDS:
            cur[m.group(1).strip()] = m.group(2)
    keep = [k for k in kernels if k['name'].startswith(('k_resident', 'k_deliver', 'k_neuron', 'k_plasticity'))]
    rows = ['# Compiler resource usage of the tick kernels (gfx950)', '',
            f'Command (in `neural-selection_amd/`): `{" ".join(cmd[:1] + ["<library flags>"] + cmd[-7:])}`',
            f'with the library flags `{" ".join(FLAGS)}`.', '',
            'LDS Size is the static LDS of the kernel; the LDS tiles add their dynamic synapse arrays '
            '(34 bytes per synapse with plasticity, 14 without) at launch.', '',
            '| kernel | ' + ' | '.join(FIELDS) + ' |', '|---|' + '---|' * len(FIELDS)]
    for k in keep:
        rows.append(f'| `{k["name"]}` | ' + ' | '.join(k.get(f, '') for f in FIELDS) + ' |')
    Path(a.out).write_text('\n'.join(rows) + '\n')
    print('\n'.join(rows))


if __name__ == '__main__':
    main()
