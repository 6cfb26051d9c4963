"""Graph mode (ns_gpu_set_graph): every step replays one captured hipGraph.
Each replay's outputs, and the state at the end, equal a batch stepped
eagerly on the same inputs, with eager work between replays (replacements
placed into free slots, which keep the graph, and one that rebuilds the
pool, which captures again), in the resident and the per-tick forms, with
LP/Prinz brains on their per-tick kernels inside the same graph."""
import numpy as np
import pytest

import ns_amd
import ns_workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('form', [ns_amd.FORM_RESIDENT, ns_amd.FORM_PER_TICK], ids=['resident', 'per-tick'])
def test_graph_replays_equal_eager_steps(form):
    w = W.small_mix(seed=61, brains=40, calls=12, types=(0, 1, 2, 4, 5, 6))
    donors = W.small_mix(seed=62, brains=6, types=(0, 1, 4, 5, 6))
    slow = W.small_mix(seed=63, brains=1, types=(0, 2, 3))          # LP/Prinz: a rebuild
    print(f'[graph] form={form} brains={w.count} calls={w.calls}', flush=True)
    A = ns_amd.GpuBatch(workload=w)
    B = ns_amd.GpuBatch(workload=w)
    born = ns_amd.GenomeHandles([donors.brain(i) for i in range(donors.count)], w.channels,
                                max_delay=w.max_delay)
    lp = ns_amd.GenomeHandles([slow.brain(0)], w.channels, max_delay=w.max_delay)
    try:
        for b in (A, B):
            b.set_form(form)
            b.reserve(0.19, 0.25)
            b.reset(w.seeds())
        A.set_graph(True)
        rng = np.random.default_rng(5)
        for call in range(w.calls):
            x = w.inputs(call)
            oa, sa = A.step(x, w.ticks, w.gain)
            ob, sb = B.step(x, w.ticks, w.gain)
            assert np.array_equal(oa, ob), f'outputs at call {call}'
            assert np.array_equal(sa, sb)
            if call in (3, 7):                                       # placements between replays
                idx = rng.choice(w.count, 3, replace=False).astype(np.int32)
                which = rng.integers(0, donors.count, 3)
                seeds = rng.integers(1, 2**31, 3).astype(np.uint32)
                for b in (A, B):
                    b.replace(idx, born, which, seeds)
            if call == 9:                                            # a rebuild: capture again
                for b in (A, B):
                    b.replace(np.array([11], np.int32), lp, [0], np.array([77], np.uint32))
        ea, eb = A.export(), B.export()
        for x, y in zip(ea, eb):
            assert np.asarray(x).tobytes() == np.asarray(y).tobytes()
        info = A.info()
        assert info['graph_launches'] == w.calls
        assert info['relayouts'] >= 1 and info['graph_captures'] >= 2
        assert B.info()['graph_launches'] == 0
    finally:
        born.close()
        lp.close()
        A.close()
        B.close()
