"""Benchmark of the per-tick brain update on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3|c2|c1]

A step is one world step: 16 one-ms ticks of every brain in the workload
(one resident tick-kernel launch, plus the per-tick kernels for brains too
large for a workgroup), with the encoded sensory input already resident in
HBM.  Prints one JSON line: neuron-updates/s (value), synaptic events/s,
synapse-updates/s, p50/p99 world-step latency, the plain per-tick form's
throughput on the same pools, the roofline of the resident tick kernel
(HIP-event durations measured here, on the stream it runs on) and the CPU
oracle timed on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import re
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path[:0] = [str(ROOT / 'neural-selection_amd'), str(ROOT / 'oracle')]

import ns_amd  # noqa: E402
import ns_workloads as W  # noqa: E402

HBM_PEAK_GBPS = 8000.0   # MI355X HBM3E
# Algorithmic bytes per unit of work (SoA-minimal streaming form, one tick):
def _header_bytes():
    """The roofline's bytes per unit of work, from include/ns_amd.h."""
    text = (ROOT / 'include' / 'ns_amd.h').read_text()
    return {m.group(1): int(m.group(2))
            for m in re.finditer(r'#define NS_BYTES_PER_(\w+) (\d+)', text)}


_BYTES = _header_bytes()
NEURON_UPDATE_BYTES = _BYTES['NEURON_UPDATE']
SYNAPSE_UPDATE_BYTES_PLASTIC = _BYTES['SYNAPSE_UPDATE_PLASTIC']
SYNAPSE_UPDATE_BYTES_STATIC = _BYTES['SYNAPSE_UPDATE_STATIC']
EVENT_BYTES = _BYTES['EVENT']


def log(msg):
    print(f'[bench] {msg}', file=sys.stderr, flush=True)


def make_workload(name):
    if name == 'c1':
        return W.c1(), False
    if name == 'c2':
        return W.c2(), False
    if name == 'c3':
        return W.c3(), True
    raise SystemExit(f'unknown workload {name}')


def describe(name, w, churn=0.0):
    base = {'workload': name, 'brains': w.count, 'neurons': w.total_neurons,
            'synapses': w.total_synapses, 'ticks_per_step': w.ticks, 'max_delay': w.max_delay}
    if name == 'c3':
        base.update(desc='capacity scale with plasticity: 32768 brains, Pareto(1.5, 32) sizes '
                         'truncated at 4096 and rescaled to 4194304 neurons, 2 synapses/neuron, '
                         'STDP profiles 1/2 with eligibility (3), D1 dopamine (4) and spine (5) '
                         'profiles, modulator neurons; '
                         + (f'births/deaths: {churn:.0%} of brains replaced by newborn genomes '
                            'every step (best-fit free slots, fresh state), inside the timed steps'
                            if churn > 0 else 'no births/deaths'),
                    plasticity=True, churn_per_step=churn)
    elif name == 'c2':
        base.update(desc='2048 brains, lognormal(ln 96, 0.75) sizes clipped 16..1024, 20 '
                         'synapses/neuron, delays 1..16, STP on, plasticity off; the world '
                         '(grid, food) is not included', plasticity=False)
    else:
        base.update(desc='64 brains x 32 neurons, fan-out 8, delays 1..8, plasticity off',
                    plasticity=False)
    return base


def host_info():
    """The host the CPU baseline ran on: logical CPUs, the CPUs this process
    may use, the CPU model and the C library."""
    import platform
    model = ''
    try:
        for line in Path('/proc/cpuinfo').read_text().splitlines():
            if line.startswith('model name'):
                model = line.split(':', 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {'nproc': os.cpu_count(), 'usable_cpus': usable, 'model': model,
            'libc': ' '.join(platform.libc_ver())}


# Host threads for the CPU legs: the CPUs this process may use, at most the
# 16 that one GPU's share of the machine holds.
HOST_THREADS_CAP = 16


def cpu_baseline(w, seconds=10.0, share=1 / 32):
    """The CPU oracle (bit-identical to the reference's CPU path) on a sample
    of the same workload's brains: every 1/share-th brain in size order, so
    the sample spans the size distribution (its largest brains included).
    Timed twice: one thread, then threaded over brains on the host's share."""
    import ctypes as ct
    import oracle as O
    lib = O.oracle_lib()
    order = np.argsort(w.neuron_counts, kind='stable')
    step = max(1, int(round(1 / share)))
    pick = np.sort(order[step // 2::step])
    seeds = w.seeds()
    brains = []
    for i in pick:
        n, s, m = w.brain(int(i))
        b = O.OracleBrain(n, s, m, w.channels, max_delay=max(8, w.max_delay))
        b.reset(seeds[i])
        brains.append(b)
    n_b = len(brains)
    arr = (ct.c_void_p * n_b)(*[b.ptr for b in brains])
    out = np.zeros((n_b, w.outputs), np.float32)
    status = np.zeros(n_b, np.int32)
    neurons = int(w.neuron_counts[pick].sum())
    host = host_info()
    threads = max(1, min(host['usable_cpus'], HOST_THREADS_CAP))

    def leg(nthreads):
        calls, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < seconds and calls < 20000:
            x = np.ascontiguousarray(w.inputs(calls)[pick])
            lib.ora_batch_step(arr, n_b, x.ctypes.data_as(O._f), w.channels, None, w.ticks, w.gain,
                               out.ctypes.data_as(O._f), w.outputs, status.ctypes.data_as(O._i32),
                               nthreads)
            calls += 1
        dt = time.perf_counter() - t0
        return neurons * w.ticks * calls / dt, calls, dt

    one, calls1, dt1 = leg(1)
    many, calls_n, dt_n = leg(threads)
    return {'value': many, 'unit': 'neuron-updates/s', 'cores': threads, 'kind': 'port',
            'sample': f'{n_b} brains ({neurons} neurons, {neurons / w.total_neurons:.1%} of the pool): '
                      f'every {step}th brain in size order, {int(w.neuron_counts[pick].min())} to '
                      f'{int(w.neuron_counts[pick].max())} neurons; {calls_n} calls x {w.ticks} ticks '
                      f'in {dt_n:.1f} s, oracle/liboracle.so threaded over brains',
            'single_thread': {'value': one, 'cores': 1, 'calls': calls1, 'seconds': dt1},
            'host': host,
            'threads_note': f'{threads} threads: the CPUs this process may use, capped at the '
                            f'{HOST_THREADS_CAP}-CPU share of one GPU'}


def serving_cpu_baseline(reqs, run_one, work_of, budget=1.5e9):
    """The same request stream on the host cores: the first requests whose
    work (neuron-updates) fits `budget`, all submitted at once to a thread pool
    that runs each through the CPU oracle brain and the task restatement.
    run_one returns the neuron-updates the request took."""
    import concurrent.futures as cf
    sample, est = [], 0.0
    for r in reqs:
        w = work_of(r)
        if sample and est + w > budget:
            break
        sample.append(r)
        est += w
    threads = max(1, min(host_info()['usable_cpus'], HOST_THREADS_CAP))
    t0 = time.perf_counter()
    done, lat = 0.0, []
    with cf.ThreadPoolExecutor(threads) as ex:
        futs = [ex.submit(run_one, r) for r in sample]
        for f in cf.as_completed(futs):
            done += f.result()
            lat.append(time.perf_counter() - t0)
    wall = time.perf_counter() - t0
    lat = np.array(lat) * 1e3
    return {'value': done / wall, 'unit': 'neuron-updates/s', 'cores': threads, 'kind': 'port',
            'sample': f'first {len(sample)} requests of the same stream ({done:.3g} neuron-updates), '
                      f'all submitted at once to a {threads}-thread pool, each run whole by the '
                      f'oracle library (oracle/liboracle.so: brain and task); {wall:.1f} s wall',
            'latency_ms': {'p50': float(np.percentile(lat, 50)), 'p99': float(np.percentile(lat, 99))}}


def load_traffic(kernel, workload, tile):
    """Measured HBM bytes per launch of `kernel` from its committed rocprofv3
    summary, if that summary was taken on this workload and tile."""
    f = ROOT / 'profiles' / f'{kernel}_{workload}.md'
    if not f.exists():
        return None
    text = f.read_text().splitlines()
    want = ','.join(str(x) for x in tile)
    if not text or not text[0].rstrip().endswith(f'tile {want}'):
        return None
    for line in text:
        if line.startswith('bytes_per_launch:'):
            return float(line.split(':')[1])
    return None


def serving_genomes(count, seed, lo=16, hi=4096):
    """C4 episodes: sizes log-uniform on [lo, hi], 8 sensors, 2 motors."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        size = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        nr, sr, mo = W.episode_genome(rng, size)
        out.append((nr, sr, mo, 1000 + i, 5000 + i))
    return out


def serve_run(eps, rate, episode_steps):
    """One continuous-batching run: `eps` submitted at Poisson times of rate
    `rate` per second while earlier ones run; per-episode latency (submit to
    finish), admission wait and whole-run throughput."""
    import ctypes as ct
    L = ns_amd.lib()
    f = ct.POINTER(ct.c_float)
    total_neurons = sum(len(e[0]) for e in eps)
    cfg = ns_amd.ServeConfig(neuron_capacity=1 << 22, synapse_capacity=1 << 24,
                             episode_capacity=8192, steps_per_episode=episode_steps,
                             ticks_per_step=16, steps_per_launch=25, noise_sigma=0.01,
                             input_gain=230.0)
    srv = L.ns_serve_create(ct.byref(cfg))
    if not srv:
        raise SystemExit(ns_amd.error())
    gaps = np.random.default_rng(int(rate)).exponential(1.0 / rate, len(eps))
    t0 = time.perf_counter()
    due = t0
    for k, (nr, sr, mo, seed, tseed) in enumerate(eps):
        due += gaps[k]
        delay = due - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        if L.ns_serve_submit(srv, len(nr), len(sr), nr.ctypes.data_as(f), sr.ctypes.data_as(f),
                             mo.ctypes.data_as(ct.POINTER(ct.c_int)), seed, tseed) < 0:
            raise SystemExit(ns_amd.error())
    if L.ns_serve_drain(srv):
        raise SystemExit(ns_amd.error())
    wall = time.perf_counter() - t0
    res = (ns_amd.ServeResult * len(eps))()
    n = L.ns_serve_poll(srv, res, len(eps))
    stats = ns_amd.ServeStats()
    L.ns_serve_stats_get(srv, ct.byref(stats))
    L.ns_serve_destroy(srv)
    lat = np.array([r.finish_s - r.submit_s for r in res[:n]]) * 1e3
    wait = np.array([r.admit_s - r.submit_s for r in res[:n]]) * 1e3
    nu = total_neurons * episode_steps * 16
    return {'rate_per_s': rate, 'episodes': n, 'wall_s': wall,
            'episodes_per_s': n / wall, 'neuron_updates_per_s': nu / wall,
            'latency_ms': {'p50': float(np.percentile(lat, 50)),
                           'p99': float(np.percentile(lat, 99)),
                           'max': float(lat.max())},
            'admission_wait_ms': {'p50': float(np.percentile(wait, 50)),
                                  'p99': float(np.percentile(wait, 99))},
            'launch_rounds': stats.launches, 'resident_cohorts_peak': stats.resident_peak,
            'device_busy_s': stats.busy_s}


def world_run(steps, warmup):
    """The full C2 world step: 2048 lognormal-size brains under the world rules
    on the 256 x 256 torus with food A/B/C at 0.05 and 1e-3 regrowth, each
    step the senses, 16 neural ticks, motor forces, feeding, bites, deaths,
    world homeostasis and regrowth, all on the device (ns_ecology_step)."""
    w = W.c2()
    batch = ns_amd.GpuBatch(workload=w, world_rules=True)
    eco = None
    try:
        batch.reset(w.seeds())
        eco = ns_amd.Ecology(batch, 256, 256, True, 0.05, 1e-3, seed=1)
        for s in range(1, warmup + 1):
            eco.step(s)
        batch.set_timing(True)
        i0 = batch.info()
        step_ms, neurons = [], 0
        for s in range(warmup + 1, warmup + steps + 1):
            alive = eco.export()[0]['alive'] == 1      # the animals this step runs
            neurons += int(w.neuron_counts[alive].sum())
            t = time.perf_counter()
            eco.step(s)                                # returns once the step is done on the device
            step_ms.append((time.perf_counter() - t) * 1e3)
        i1 = batch.info()
        an = eco.export()[0]
        wall = sum(step_ms) / 1e3
        neural_ms = (i1['resident_ms'] - i0['resident_ms']) / steps
        return {'desc': 'C2 world: 2048 animals, lognormal brains (median 96, 16..1024 neurons, '
                        '20 synapses per neuron, delays 1-16 ms), 256x256 torus, food A/B/C 0.05, '
                        'regrowth 1e-3; senses, 16 ticks, body, feeding, bites, homeostasis, regrowth',
                'steps': steps, 'warmup': warmup,
                'ms_per_world_step': wall * 1e3 / steps,
                'step_ms': {'p50': float(np.percentile(step_ms, 50)),
                            'p99': float(np.percentile(step_ms, 99)), 'max': float(max(step_ms))},
                'neural_launch_ms_per_step': neural_ms,
                'neuron_updates_per_s': neurons * w.ticks / wall,
                'alive_at_end': int((an['alive'] == 1).sum())}
    finally:
        if eco is not None:
            eco.close()
        batch.close()


def bench_serving(args):
    """Continuous batching under Poisson arrivals: per-episode latency (submit
    to finish) and whole-run throughput, one run per arrival rate."""
    rates = [float(x) for x in args.rates.split(',')]
    eps = serving_genomes(args.episodes, seed=17)
    total_neurons = sum(len(e[0]) for e in eps)
    runs = []
    for rate in rates:
        log(f'serving: {len(eps)} episodes at {rate:g}/s')
        runs.append(serve_run(eps, rate, args.episode_steps))
        log(f'  p50 {runs[-1]["latency_ms"]["p50"]:.1f} ms, p99 {runs[-1]["latency_ms"]["p99"]:.1f} ms, '
            f'{runs[-1]["neuron_updates_per_s"]:.3e} neuron-updates/s')
    top = runs[-1]
    cpu = None
    if not args.no_cpu:
        log('timing the CPU oracle on a bounded sample of the same episodes')
        import tasks
        steps = args.episode_steps

        def run_one(e):
            _, ran = tasks.linear_episode_native(e[0], e[1], e[2], e[3], e[4], steps, 0.01)
            return len(e[0]) * ran * 16

        cpu = serving_cpu_baseline(eps, run_one, lambda e: len(e[0]) * steps * 16)
    line = {
        'metric': 'neuron-updates/s + synaptic events/s on 1 GPU; p50/p99 batched-eval latency',
        'value': top['neuron_updates_per_s'], 'unit': 'neuron-updates/s', 'n_gpus': 1,
        'steps': args.episodes, 'warmup': 0, 'ms_per_step': top['wall_s'] * 1e3 / max(top['episodes'], 1),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32',
        'data': 'synthetic (seeded episode genomes and tasks)',
        'config': {'workload': 'c4', 'desc': 'serving: Poisson arrivals of genome-evaluation '
                   'episodes, sizes log-uniform 16..4096 neurons, 4-d linear task, 8 sensors, '
                   f'2 motors, {args.episode_steps} steps x 16 ticks', 'episodes': len(eps),
                   'neurons_total': total_neurons, 'rates_per_s': rates},
        'serving': runs, 'roofline': None, 'cpu_baseline': cpu,
    }
    print(json.dumps(line), flush=True)


def bench_mazes(args):
    """C5: many concurrent maze requests (sizes log-uniform 64..2048 neurons),
    each `--mazes` depth-first mazes of at most `--maze-horizon` steps with the
    brain's state kept; all submitted at once.  Throughput counts the steps
    each request took until its last maze ended."""
    import ctypes as ct
    L = ns_amd.lib()
    rng = np.random.default_rng(23)
    reqs = []
    for i in range(args.maze_requests):
        size = int(np.exp(rng.uniform(np.log(64), np.log(2048))))
        nr, sr, mo = W.maze_genome(rng, size)
        reqs.append((nr, sr, mo, 3000 + i, 9000 + i))
    classes = []
    c = 16.0
    while True:                            # the server's slot classes
        classes.append(int(np.ceil(c)))
        if classes[-1] >= 2048:
            break
        c *= np.sqrt(2.0)
    per = int(np.ceil(args.maze_requests / 10)) + 64   # sizes fall in about 10 classes
    cfg = ns_amd.ServeConfig(neuron_capacity=per * sum(classes), synapse_capacity=4 * per * sum(classes),
                             episode_capacity=per * len(classes), steps_per_episode=1,
                             ticks_per_step=16, steps_per_launch=25, noise_sigma=0.0,
                             input_gain=230.0, max_episode_neurons=2048, task=ns_amd.TASK_MAZE,
                             maze_episodes=args.mazes, maze_horizon=args.maze_horizon)
    t_build = time.perf_counter()
    srv = L.ns_serve_create(ct.byref(cfg))
    if not srv:
        raise SystemExit(ns_amd.error())
    log(f'mazes: slot pool for {per} requests per class built in {time.perf_counter() - t_build:.1f} s; '
        f'submitting {len(reqs)} requests x {args.mazes} mazes')
    f = ct.POINTER(ct.c_float)
    t0 = time.perf_counter()
    for nr, sr, mo, seed, tseed in reqs:
        if L.ns_serve_submit(srv, len(nr), len(sr), nr.ctypes.data_as(f), sr.ctypes.data_as(f),
                             mo.ctypes.data_as(ct.POINTER(ct.c_int)), seed, tseed) < 0:
            raise SystemExit(ns_amd.error())
    if L.ns_serve_drain(srv):
        raise SystemExit(ns_amd.error())
    wall = time.perf_counter() - t0
    res = (ns_amd.ServeResult * len(reqs))()
    n = L.ns_serve_poll(srv, res, len(reqs))
    stats = ns_amd.ServeStats()
    L.ns_serve_stats_get(srv, ct.byref(stats))
    L.ns_serve_destroy(srv)
    order = [r for r in res[:n]]
    steps = np.array([r.final_state[1] for r in order])
    nu = float(sum(len(reqs[r.id][0]) * r.final_state[1] * 16 for r in order))
    lat = np.array([r.finish_s - r.submit_s for r in order]) * 1e3
    solved = np.array([r.final_state[0] for r in order])
    log(f'  {n} requests in {wall:.2f} s: {nu / wall:.3e} neuron-updates/s, '
        f'latency p50 {np.percentile(lat, 50):.0f} ms p99 {np.percentile(lat, 99):.0f} ms')
    cpu = None
    if not args.no_cpu:
        log('timing the CPU oracle on a bounded sample of the same requests')
        import tasks
        mz, hz = args.mazes, args.maze_horizon

        def run_one(r):
            _, ran = tasks.maze_request_native(r[0], r[1], r[2], r[3], r[4], mz, hz)
            return len(r[0]) * ran * 16

        cpu = serving_cpu_baseline(reqs, run_one, lambda r: len(r[0]) * mz * hz * 16)
    line = {
        'metric': 'neuron-updates/s + synaptic events/s on 1 GPU; p50/p99 batched-eval latency',
        'value': nu / wall, 'unit': 'neuron-updates/s', 'n_gpus': 1,
        'steps': int(steps.max()) if n else 0, 'warmup': 0,
        'ms_per_step': wall * 1e3 / max(float(steps.max()) if n else 1.0, 1.0),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32',
        'data': 'synthetic (seeded maze genomes and mazes)',
        'config': {'workload': 'c5', 'desc': f'{len(reqs)} concurrent maze requests, sizes log-uniform '
                   f'64..2048 neurons, {args.mazes} depth-first 31x31 mazes each (state kept), '
                   f'horizon {args.maze_horizon} steps x 16 ticks, reward as a dopamine event',
                   'requests': len(reqs), 'neurons_total': int(sum(len(r[0]) for r in reqs))},
        'serving': {'requests': n, 'wall_s': wall, 'requests_per_s': n / wall,
                    'latency_ms': {'p50': float(np.percentile(lat, 50)),
                                   'p99': float(np.percentile(lat, 99)), 'max': float(lat.max())},
                    'env_steps': {'mean': float(steps.mean()), 'max': float(steps.max())},
                    'mazes_solved': {'total': int(solved.sum()), 'requests_solving_any': int((solved > 0).sum())},
                    'launch_rounds': stats.launches, 'resident_peak': stats.resident_peak,
                    'device_busy_s': stats.busy_s},
        'roofline': None, 'cpu_baseline': cpu,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=50)
    # 40 warmup steps: the first ~20 steps of sustained load after the GPU
    # idles (autotuning, pool build) hold a one-time slowdown of one step
    ap.add_argument('--warmup', type=int, default=40)
    ap.add_argument('--workload', default='c3')
    ap.add_argument('--plain-steps', type=int, default=5)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-autotune', action='store_true')
    ap.add_argument('--eager', action='store_true',
                    help='time eager steps (default: each step one hipGraph replay)')
    ap.add_argument('--tiles', default='')
    ap.add_argument('--churn', type=float, default=None,
                    help='fraction of brains replaced by newborns every step (c3 default 0.01)')
    ap.add_argument('--slack', type=float, default=0.05,
                    help='churn: slot capacity rungs 1 + slack apart (ns_gpu_reserve)')
    ap.add_argument('--spare', type=float, default=0.03,
                    help='churn: empty spare slots per brain (ns_gpu_reserve)')
    ap.add_argument('--rates', default='100,1000,5000', help='c4: Poisson arrival rates (1/s)')
    ap.add_argument('--episodes', type=int, default=300, help='c4: episodes per rate')
    ap.add_argument('--episode-steps', type=int, default=500, help='c4: env steps per episode')
    ap.add_argument('--maze-requests', type=int, default=8192, help='c5: concurrent maze requests')
    ap.add_argument('--mazes', type=int, default=10, help='c5: mazes per request')
    ap.add_argument('--maze-horizon', type=int, default=100, help='c5: steps per maze at most')
    ap.add_argument('--no-serving', action='store_true', help='leave out the C4 serving leg')
    ap.add_argument('--no-world', action='store_true', help='leave out the C2 world-step leg')
    ap.add_argument('--world-steps', type=int, default=100, help='world leg: timed world steps')
    ap.add_argument('--serving-rate', type=float, default=1000.0, help='serving leg: arrivals per s')
    ap.add_argument('--serving-episodes', type=int, default=200, help='serving leg: episodes')
    ap.add_argument('--copy-calibration', type=int, default=0, metavar='MIB',
                    help='only copy MIB MiB device to device (HBM counter calibration)')
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * 2**20 // 4
        src = torch.ones(n, dtype=torch.float32, device='cuda')
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({'copy_bytes_each_way': n * 4, 'copies': 3}))
        return
    if args.gpus != 1:
        print(f'bench.py measures one GPU only (asked for {args.gpus}); brains are independent, '
              'so the update shards by brain with no exchange', file=sys.stderr)
        sys.exit(2)
    if args.workload == 'c4':
        bench_serving(args)
        return
    if args.workload == 'c5':
        bench_mazes(args)
        return
    import torch
    torch.cuda.init()

    log(f'building workload {args.workload}')
    t0 = time.perf_counter()
    w, plastic = make_workload(args.workload)
    log(f'{w.count} brains, {w.total_neurons} neurons, {w.total_synapses} synapses '
        f'({time.perf_counter() - t0:.1f} s)')
    churn = args.churn if args.churn is not None else (0.01 if args.workload == 'c3' else 0.0)
    batch = ns_amd.GpuBatch(workload=w)
    batch.reset(w.seeds())
    tile_key = None          # the tile as set_tiles names it
    if args.tiles:
        tile_key = tuple(int(x) for x in args.tiles.split(','))
        batch.set_tiles(*tile_key)
    elif not args.no_autotune:
        log('autotuning tile shape')
        tile_key = tuple(batch.autotune(w.ticks))
        log(f'tile {tile_key}')
    births = None
    if churn > 0:
        import ctypes as ct
        t1 = time.perf_counter()
        batch.reserve(args.slack, args.spare)
        newborn = W.c3_newborns(count=2048)
        births = ns_amd.GenomeHandles([newborn.brain(i) for i in range(newborn.count)], w.channels,
                                      max_delay=w.max_delay)
        per_step = max(1, int(round(churn * w.count)))
        rng = np.random.default_rng(77)
        plans = []
        for k in range(args.warmup + args.steps + 1):
            idx = np.ascontiguousarray(rng.choice(w.count, per_step, replace=False), np.int32)
            which = rng.integers(0, newborn.count, per_step)
            arr = (ct.c_void_p * per_step)(*[births.ptrs[j] for j in which])
            seeds = np.ascontiguousarray(rng.integers(1, 2**32 - 1, per_step), np.uint32)
            plans.append((idx, arr, seeds, which))
        log(f'churn: {per_step} brains replaced per step; slot reserve built in '
            f'{time.perf_counter() - t1:.1f} s')
    # The live pool, brain by brain, as births and deaths change it: neurons,
    # synapses and plastic excitatory pairs of each brain's current genome.
    live = {'n': w.neuron_counts.astype(np.int64), 's': w.edge_counts.astype(np.int64),
            'p': W.plastic_pair_counts(w)}
    born = None
    if births is not None:
        born = {'n': newborn.neuron_counts.astype(np.int64), 's': newborn.edge_counts.astype(np.int64),
                'p': W.plastic_pair_counts(newborn)}
    totals = {k: int(v.sum()) for k, v in live.items()}
    launched = []            # live totals of every launched step, in order
    info = batch.info()
    log(f'packs {info["resident_packs"]} fallback brains {info["fallback_brains"]} '
        f'({info["fallback_neurons"]} neurons) device {info["device_bytes"] / 2**30:.2f} GiB')

    patterns = 8
    inputs = torch.from_numpy(np.stack([w.inputs(c) for c in range(patterns)])).cuda()
    out = torch.zeros((w.count, w.outputs), dtype=torch.float32, device='cuda')
    # a stream of our own (default priority): torch's default stream is the
    # null stream, on which events would not bracket the library's launches
    stream = torch.cuda.Stream()

    def births_for(k):
        if births is not None:
            import ctypes as ct
            idx, arr, seeds, which = plans[k % len(plans)]
            rc = batch.lib.ns_gpu_replace(batch.ptr, len(idx), idx.ctypes.data_as(ct.POINTER(ct.c_int)),
                                          arr, seeds.ctypes.data_as(ct.POINTER(ct.c_uint32)),
                                          stream.cuda_stream)
            if rc:
                raise RuntimeError(ns_amd.error())
            for key in live:
                totals[key] += int(born[key][which].sum() - live[key][idx].sum())
                live[key][idx] = born[key][which]

    # A world step launches its ticks, then places the births and deaths the
    # next step starts from (host work overlapping this launch, device work
    # queued behind it), as a world loop runs them.
    def step(k):
        batch.step_device(inputs[k % patterns].data_ptr(), w.ticks, w.gain, out.data_ptr(),
                          stream=stream.cuda_stream)
        launched.append(dict(totals))
        births_for(k + 1)

    def live_sum(first, key):
        return sum(t[key] for t in launched[first:])

    # Timed steps: each step one replay of a captured hipGraph (the output
    # clears and the whole 16-tick launch), replacements eager between them.
    batch.set_graph(not args.eager)
    births_for(0)
    for k in range(args.warmup):
        step(k)
    torch.cuda.synchronize()
    batch.set_timing(True)
    i0 = batch.info()
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    torch.cuda.synchronize()
    timed0 = len(launched)
    t0 = time.perf_counter()
    for k in range(args.steps):
        evs[k].record(stream)
        step(args.warmup + k)
    evs[-1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    i1 = batch.info()
    step_ms = np.array([evs[k].elapsed_time(evs[k + 1]) for k in range(args.steps)])
    log(f'{args.steps} steps in {wall:.3f} s; step p50 {np.percentile(step_ms, 50):.3f} ms')

    # neuron- and synapse-updates of the live pool at each timed launch
    nu = live_sum(timed0, 'n') * w.ticks
    su = live_sum(timed0, 's') * w.ticks
    su_pairs = live_sum(timed0, 'p') * w.ticks
    ev = i1['events'] - i0['events']
    value = nu / wall
    res_launches = i1['resident_launches'] - i0['resident_launches']
    res_ms = (i1['resident_ms'] - i0['resident_ms']) / max(res_launches, 1)
    # the resident launch's share per launch (brains on the per-tick path, LP
    # and Prinz cells, excluded; plastic pairs are counted only where no brain
    # takes that path, which holds for the c1/c2/c3 genomes)
    assert i1['fallback_neurons'] == 0 or not plastic
    res_neurons = nu / w.ticks / args.steps - i1['fallback_neurons']
    res_synapses = su / w.ticks / args.steps - i1['fallback_synapses']
    res_pairs = su_pairs / w.ticks / args.steps
    res_events = (i1['resident_events'] - i0['resident_events']) / max(res_launches, 1)
    alg_bytes = (NEURON_UPDATE_BYTES * res_neurons * w.ticks
                 + SYNAPSE_UPDATE_BYTES_PLASTIC * res_pairs * w.ticks
                 + SYNAPSE_UPDATE_BYTES_STATIC * (res_synapses - res_pairs) * w.ticks
                 + EVENT_BYTES * res_events)
    achieved = alg_bytes / (res_ms * 1e-3) / 1e9 if res_ms > 0 else 0.0
    per_tick_share = (i1['per_tick_ms'] - i0['per_tick_ms']) / max(args.steps, 1)

    # plain per-tick form on the same pools (the 5x denominator)
    batch.set_timing(False)
    batch.set_graph(False)
    batch.set_form(ns_amd.FORM_PER_TICK)
    for k in range(2):
        step(k)
    torch.cuda.synchronize()
    plain0 = len(launched)
    t1 = time.perf_counter()
    for k in range(args.plain_steps):
        step(k)
    torch.cuda.synchronize()
    plain_wall = time.perf_counter() - t1
    plain_value = live_sum(plain0, 'n') * w.ticks / plain_wall
    per = max(args.plain_steps, 1)
    plain_alg = (NEURON_UPDATE_BYTES * live_sum(plain0, 'n') / per * w.ticks
                 + SYNAPSE_UPDATE_BYTES_PLASTIC * live_sum(plain0, 'p') / per * w.ticks
                 + SYNAPSE_UPDATE_BYTES_STATIC * (live_sum(plain0, 's') - live_sum(plain0, 'p')) / per
                 * w.ticks + EVENT_BYTES * ev / args.steps)

    def leg(steps, graph):
        # graph: each step one replay of a captured hipGraph (ns_gpu_set_graph);
        # the replacements stay eager work between replays
        batch.set_graph(graph)
        for k in range(2):
            step(k)
        torch.cuda.synchronize()
        first = len(launched)
        t = time.perf_counter()
        for k in range(steps):
            step(k)
        torch.cuda.synchronize()
        t = time.perf_counter() - t
        batch.set_graph(False)
        return t, live_sum(first, 'n') * w.ticks / t

    plain_graph_wall, plain_graph_value = leg(args.plain_steps, True)
    batch.set_form(ns_amd.FORM_RESIDENT)
    # the resident form in the other launch mode
    other_wall, other_value = leg(args.steps, args.eager)
    log(f'resident {value:.3e} ({"eager" if args.eager else "graph"}; '
        f'{"graph" if args.eager else "eager"} {other_value:.3e}) vs plain per-tick '
        f'{plain_value:.3e} (graph replays {plain_graph_value:.3e}) neuron-updates/s')

    serving = None
    if not args.no_serving:
        # p50/p99 batched-evaluation latency: a short C4 run through the
        # continuous-batching loop (its own device pool; the world-step pool
        # is released first)
        batch.close()
        log(f'serving leg: {args.serving_episodes} C4 episodes at {args.serving_rate:g}/s')
        eps = serving_genomes(args.serving_episodes, seed=17)
        serving = serve_run(eps, args.serving_rate, args.episode_steps)
        serving['desc'] = (f'C4 episodes (log-uniform 16..4096 neurons, 4-d linear task, '
                           f'{args.episode_steps} steps x 16 ticks), Poisson arrivals')
        log(f'  p50 {serving["latency_ms"]["p50"]:.1f} ms, p99 {serving["latency_ms"]["p99"]:.1f} ms')

    world = None
    if not args.no_world:
        if args.no_serving:
            batch.close()
        log(f'world leg: {args.world_steps} C2 world steps on the device')
        world = world_run(args.world_steps, 10)
        log(f'  {world["ms_per_world_step"]:.3f} ms per world step '
            f'(neural launch {world["neural_launch_ms_per_step"]:.3f} ms)')

    cpu = None
    if not args.no_cpu:
        log('timing the CPU oracle on a bounded sample (one thread, then the host share)')
        cpu = cpu_baseline(w)

    tile = (i1['threads'], i1['neurons_per_thread'], i1['synapses_per_thread'], i1['min_waves'],
            i1['lds_synapses'])
    if tile[2] == 0:
        kernel = 'k_resident_lds'
        kernel_full = f'k_resident_lds<{tile[0]}, {tile[1]}, {tile[3]}, {str(plastic).lower()}>'
    else:
        kernel = 'k_resident'
        kernel_full = f'k_resident<{tile[0]}, {tile[1]}, {tile[2]}, {tile[3]}>'
    traffic = load_traffic(kernel, args.workload, tile_key) if tile_key else None
    measured = traffic / (res_ms * 1e-3) / 1e9 if traffic and res_ms else None
    # what bounds the launch, read from the measurement: the streaming
    # roofline is HBM's, but a launch that moves far less than the HBM peak
    # allows in its time is held by latency (barriers, dependent LDS and HBM
    # round trips), not bandwidth
    limiter = None
    if measured:
        limiter = ('hbm bandwidth' if measured >= 0.6 * HBM_PEAK_GBPS else
                   f'latency: measured HBM traffic {measured / HBM_PEAK_GBPS:.0%} of peak')
    line = {
        'metric': 'neuron-updates/s + synaptic events/s on 1 GPU; p50/p99 batched-eval latency',
        'value': value,
        'unit': 'neuron-updates/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': wall * 1e3 / args.steps,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32',
        'data': 'synthetic (seeded genomes and Poisson sensory input from ns_workloads.py)',
        'config': describe(args.workload, w, churn),
        'synaptic_events_per_s': ev / wall,
        'synapse_updates_per_s': su / wall,
        'step_latency_ms': {'p50': float(np.percentile(step_ms, 50)),
                            'p99': float(np.percentile(step_ms, 99)),
                            'max': float(step_ms.max()),
                            # timed steps over 1.5x the median: (index, ms)
                            'slow_steps': [[int(k), round(float(step_ms[k]), 3)] for k in
                                           np.flatnonzero(step_ms > 1.5 * np.percentile(step_ms, 50))]},
        'plain_per_tick': {'value': plain_value, 'speedup': value / plain_value,
                           'steps': args.plain_steps,
                           'ms_per_step': plain_wall * 1e3 / max(args.plain_steps, 1),
                           'effective_GBps': plain_alg / (plain_wall / max(args.plain_steps, 1)) / 1e9,
                           'graph': {'value': plain_graph_value, 'speedup': value / plain_graph_value,
                                     'ms_per_step': plain_graph_wall * 1e3 / max(args.plain_steps, 1)}},
        'step_launch': 'eager' if args.eager else 'hipGraph replay',
        'other_launch': {'mode': 'hipGraph replay' if args.eager else 'eager', 'value': other_value,
                         'steps': args.steps, 'ms_per_step': other_wall * 1e3 / args.steps},
        'tile': {'key': list(tile_key) if tile_key else None,
                 'threads': tile[0], 'neurons_per_thread': tile[1], 'synapses_per_thread': tile[2],
                 'min_waves_per_simd': tile[3], 'lds_synapses': tile[4],
                 'packs': i1['resident_packs'],
                 'big_brains': i1['big_brains'], 'big_neurons': i1['big_neurons'],
                 'fallback_brains': i1['fallback_brains'],
                 'fallback_neurons': i1['fallback_neurons'],
                 'per_tick_ms_per_step': per_tick_share},
        'roofline': {'bound': 'hbm', 'kernel': kernel_full, 'achieved': achieved,
                     'peak': HBM_PEAK_GBPS, 'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBPS,
                     'traffic': traffic,
                     'measured_GBps': measured,
                     'measured_frac': measured / HBM_PEAK_GBPS if measured else None,
                     'limiter': limiter,
                     'launch_ms': res_ms, 'algorithmic_bytes_per_launch': alg_bytes,
                     'bytes_model': {'neuron_update': NEURON_UPDATE_BYTES,
                                     'plastic_pair_synapse_update': SYNAPSE_UPDATE_BYTES_PLASTIC,
                                     'other_synapse_update': SYNAPSE_UPDATE_BYTES_STATIC,
                                     'event': EVENT_BYTES,
                                     'per_launch': {'neuron_updates': res_neurons * w.ticks,
                                                    'plastic_pair_synapse_updates': res_pairs * w.ticks,
                                                    'other_synapse_updates':
                                                        (res_synapses - res_pairs) * w.ticks,
                                                    'events': res_events}}},
        'live_pool': {'neurons_per_step': nu / w.ticks / args.steps,
                      'synapses_per_step': su / w.ticks / args.steps,
                      'plastic_pairs_per_step': su_pairs / w.ticks / args.steps},
        'serving_c4': serving,
        'world_c2': world,
        'cpu_baseline': cpu,
    }
    if births is not None:
        line['churn'] = {'fraction_per_step': churn, 'brains_per_step': len(plans[0][0]),
                         'slack': args.slack, 'spare': args.spare,
                         'placements': i1['placements'], 'relayouts': i1['relayouts'],
                         'appended_slots': i1['appended_slots'],
                         'staging_grows': i1['staging_grows'],
                         'slots': i1['slots'], 'free_slots': i1['free_slots'],
                         'slot_neurons': i1['slot_neurons'], 'slot_synapses': i1['slot_synapses']}
    print(json.dumps(line), flush=True)
    if args.no_serving and args.no_world:
        batch.close()
    if births is not None:
        births.close()


if __name__ == '__main__':
    main()
