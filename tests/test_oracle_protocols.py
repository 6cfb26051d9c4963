"""Known answers the reference publishes for its own physiology protocols
(EXPERIMENTS.md sections 1-10, produced by tests/physiology.cpp and
tests/neuromodulation.cpp), reproduced
on the oracle's function probes.  Each protocol is restated here from its
description; the expected numbers are the reference's printed results, to the
precision it prints them.  A dropped term, a wrong sign or a transposed index
in the oracle's neuron, dendrite, release, STDP, modulated-plasticity or
homeostasis code moves at least one of these.  Protocols that need a receptor
blocker switch (NMDA, L-type, PKA) are left out: the tick path, like the
reference world, always runs with every receptor available."""
import ctypes as ct

import numpy as np
import pytest

import oracle as O

LIB = None
_f4 = ct.POINTER(ct.c_float)


def lib():
    global LIB
    if LIB is None:
        LIB = O.oracle_lib()
    return LIB


def profile(t):
    g = np.zeros(1, O.NEURON_GENE)
    lib().ora_fn_profile(t, g.ctypes.data)
    return g


def recorded(t):
    """recorded_profile(t) (physiology.cuh:82-92): type 0 is the Naud RS fit
    without the L5 afterpotential, type 1 the cNA fit."""
    g = profile(t)
    if t == 0:
        g['f'][0][8] = -53.0      # reset_mv
        g['f'][0][9] = 0.0        # afterdepol_spike_pa
        g['f'][0][10] = 0.0       # afterdepol_tau_ms
    return g


def fresh(g, seed=1):
    s = np.zeros(1, O.NEURON_STATE)
    lib().ora_fn_initial_neuron(g.ctypes.data, seed, s.ctypes.data)
    return s


def advance(g, s, current, exc=None, inh=0.0):
    e = np.zeros(4, np.float32) if exc is None else np.asarray(exc, np.float32)
    lib().ora_fn_neuron_advance(g.ctypes.data, s.ctypes.data, current, e.ctypes.data_as(_f4), inh)


def voltage(s):
    return float(s['f'][0][0])


def synapse_gene(weight, U, rec, fac, delay=1, plasticity=0, sites=14):
    g = np.zeros(1, O.SYNAPSE_GENE)
    g['initial_weight_ns'], g['utilization'], g['recovery_ms'] = weight, U, rec
    g['facilitation_ms'], g['pre'], g['post'] = fac, 0, 1
    g['delay'], g['active'], g['plasticity'], g['release_sites'] = delay, 1, plasticity, sites
    return g


def fresh_syn(g, seed=1):
    s = np.zeros(1, O.SYNAPSE_STATE)
    lib().ora_fn_initial_synapse(g.ctypes.data, seed, s.ctypes.data)
    return s


def release(g, s, tick):
    return lib().ora_fn_synapse_release(g.ctypes.data, s.ctypes.data, tick)


def weight(s):
    return float(s['f'][0][0])


# ---- 1. step currents (EXPERIMENTS.md section 1) ---------------------------
def step_spikes(g, current, ms=2000):
    s = fresh(g)
    times = []
    for t in range(ms):
        advance(g, s, current)
        if s['spike'][0]:
            times.append(t)
    return times


@pytest.mark.parametrize('t,current,count,isi', [(0, 130.0, 32, (17, 64)), (1, 150.0, 73, (23, 27)),
                                                 (0, 98.0, 22, None), (1, 184.0, 100, None),
                                                 (0, 0.0, 0, None)])
def test_step_current_spike_counts(t, current, count, isi):
    times = step_spikes(recorded(t), current)
    assert len(times) == count
    if isi:
        assert (times[1] - times[0], times[-1] - times[-2]) == isi


# ---- 2/5. release statistics (sections 2, 3 and 7) ----------------------------
def test_paired_pulse_depression_and_recovery():
    g = synapse_gene(1.0, 0.55, 450.0, 0.0)
    first = second = recovered = 0.0
    trials = 10000
    for trial in range(1, trials + 1):
        s = fresh_syn(g, trial)
        first += release(g, s, 0)
        second += release(g, s, 50)
        recovered += release(g, s, 2050)
    first, second, recovered = first / trials, second / trials, recovered / trials
    assert round(first, 3) == 0.548
    assert round(second, 3) == 0.279
    assert round(second / first, 3) == 0.509
    assert round(recovered, 3) == 0.546


def test_thirty_hz_facilitation_depression_and_failures():
    fac = synapse_gene(1.0, 0.1, 30.0, 1700.0)
    dep = synapse_gene(1.0, 0.55, 450.0, 0.0)
    trials = 10000
    f1 = f6 = d1 = d6 = 0.0
    ffail = dfail = 0
    for trial in range(1, trials + 1):
        fs, ds = fresh_syn(fac, trial), fresh_syn(dep, trial + trials)
        for pulse in range(6):
            f = release(fac, fs, pulse * 33)
            d = release(dep, ds, pulse * 33)
            if pulse == 0:
                f1 += f; d1 += d
                ffail += f == 0; dfail += d == 0
            if pulse == 5:
                f6 += f; d6 += d
    assert round(f6 / f1, 3) == 3.741
    assert round(d6 / d1, 3) == 0.134
    assert round(100.0 * ffail / trials, 2) == 22.28
    assert dfail == 0


# ---- 4. dendritic summation (section 4) -------------------------------------
def dendritic_peak(each, same, two=True):
    g = recorded(0)
    g['f'][0][3] = 100.0          # threshold out of reach: subthreshold readout
    s = fresh(g)
    rest = float(g['f'][0][2])
    peak = 0.0
    for t in range(30):
        exc = np.zeros(4, np.float32)
        if t == 0:
            exc[0] = each
            if two:
                exc[0 if same else 1] += each
        advance(g, s, 0.0, exc)
        peak = max(peak, voltage(s) - rest)
    return peak


def test_dendritic_peaks():
    got = [dendritic_peak(0.5, True), dendritic_peak(0.5, False), dendritic_peak(2.5, True),
           dendritic_peak(2.5, False), dendritic_peak(8.0, True), dendritic_peak(8.0, False)]
    assert [round(x, 3) for x in got] == [1.446, 1.445, 8.695, 6.943, 22.324, 25.792]


# ---- 3/6. pair STDP (sections 3 and 6) --------------------------------------
def paired(interval, excitatory_target, profile_id=1, depol=0.0, pairs=60):
    g = synapse_gene(1.0, 0.55, 450.0, 0.0, delay=5, plasticity=profile_id)
    s = fresh_syn(g)
    f = lib().ora_fn_plasticity
    for t in range(pairs * 1000):
        phase = t % 1000
        f(g.ctypes.data, s.ctypes.data, int(phase == 100), int(phase == 100 + interval),
          int(excitatory_target), depol, 0)
    return weight(s)


def test_hippocampal_pair_stdp():
    assert round(paired(10, True), 3) == 1.415
    assert round(paired(-10, True), 3) == 0.836
    assert paired(50, True) == 1.0 and paired(-50, True) == 1.0
    assert paired(10, False) == 1.0


def test_cortical_l5_pair_stdp():
    assert paired(10, True, 2, 0.0, 50) == 1.0
    assert round(paired(10, True, 2, 3.0, 50), 3) == 1.360
    assert round(paired(-10, True, 2, 0.0, 50), 3) == 0.861


# ---- 6. afterpotential and the 40 Hz train (section 6) ----------------------
def test_isolated_cortical_afterpotential():
    g = profile(0)
    s = fresh(g)
    rest = float(g['f'][0][2])
    first, spikes, at = -1, 0, {}
    for t in range(200):
        advance(g, s, 3500.0 if t == 100 else 0.0)
        if s['spike'][0]:
            spikes += 1
            if first < 0:
                first = t
        for k in (2, 10, 20, 80):
            if first >= 0 and t == first + k:
                at[k] = voltage(s) - rest
    assert spikes == 1
    assert [round(at[k], 2) for k in (2, 10, 20, 80)] == [-1.78, 1.61, 3.51, 0.46]


def cortical_train(bursts, afterdepol):
    g = profile(0)
    if not afterdepol:
        g['f'][0][9] = 0.0
    cell = fresh(g)
    rest = float(g['f'][0][2])
    sg = synapse_gene(1.0, 0.55, 450.0, 0.0, delay=1, plasticity=2)
    syn = fresh_syn(sg)
    evoked, residual, nres = 0, 0.0, 0
    duration = 150000 if bursts else 500000
    plast = lib().ora_fn_plasticity
    exc_on = np.array([0.2, 0, 0, 0], np.float32)
    exc_off = np.zeros(4, np.float32)
    for t in range(duration):
        m = t % 10000
        if bursts:
            pre = 100 <= m <= 200 and (m - 100) % 25 == 0
            post = 110 <= m <= 210 and (m - 110) % 25 == 0
        else:
            pre, post = m == 100, m == 110
        if bursts and post and m > 110:
            residual += voltage(cell) - rest
            nres += 1
        advance(g, cell, (3500.0 if afterdepol else 5000.0) if post else 0.0,
                exc_on if pre else exc_off)
        spike = int(cell['spike'][0])
        evoked += spike
        plast(sg.ctypes.data, syn.ctypes.data, int(pre), spike, 1,
              float(cell['f'][0][4]), 0)
    return weight(syn), evoked, residual / max(nres, 1)


@pytest.mark.slow
def test_cortical_40hz_train():
    w, n, res = cortical_train(True, True)
    assert n == 75 and round(res, 2) == 7.24 and round(w, 3) == 1.297
    w, n, _ = cortical_train(False, True)
    assert n == 50 and w == 1.0
    w, n, res = cortical_train(True, False)
    assert n == 75 and round(res, 2) == -15.85 and round(w, 3) == 0.792


# ---- 8. homeostatic scaling (section 8) -------------------------------------
def test_homeostatic_scaling_48h():
    pyr, inter = recorded(0), recorded(1)
    states = {k: fresh(pyr) for k in ('quiet', 'control', 'active')}
    states['inter'] = fresh(inter)
    rates = {'quiet': 0.0, 'control': 0.4, 'active': 1.06, 'inter': 0.0}
    for _ in range(48):
        for k, s in states.items():
            lib().ora_fn_homeostasis((inter if k == 'inter' else pyr).ctypes.data, s.ctypes.data,
                                     3600000.0, rates[k])
    scale = {k: float(np.exp(np.float32(s['f'][0][5]))) for k, s in states.items()}
    assert round(scale['quiet'], 2) == 1.92
    assert round(scale['active'], 2) == 0.70
    assert abs(scale['control'] - 1) < 1e-5 and abs(scale['inter'] - 1) < 1e-5


# ---- 7, 9, 10. modulated plasticity profiles (EXPERIMENTS.md 7, 9, 10) -----
MOD_NE, MOD_5HT, MOD_DA = 1, 2, 4


def plast_step(g, s, arrived, post, mod):
    lib().ora_fn_plasticity(g.ctypes.data, s.ctypes.data, int(arrived), int(post), 1, 0.0, mod)


def cortical_eligibility(modulator, delay_ms, pair=True, receptors=MOD_NE | MOD_5HT):
    """Two pathways, 200 pairings at 10 Hz, one arrival and four spikes at 100 Hz."""
    g = synapse_gene(1.0, 0.55, 450.0, 0.0, delay=1, plasticity=3)
    g['mod_receptors'] = receptors
    fwd, rev = fresh_syn(g), fresh_syn(g)
    mod_tick = -1 if delay_ms < 0 else 20000 + delay_ms
    for t in range(20001 if mod_tick < 0 else mod_tick + 1):
        phase = t % 100
        training = pair and t < 20000
        event = modulator if t == mod_tick else 0
        plast_step(g, fwd, training and phase == 0, training and phase in (10, 20, 30, 40), event)
        plast_step(g, rev, training and phase == 40, training and phase in (0, 10, 20, 30), event)
    return weight(fwd), weight(rev)


def test_eligibility_profile_ne_5ht():
    assert cortical_eligibility(0, -1) == (1.0, 1.0)
    f, r = cortical_eligibility(MOD_NE, 0)
    assert round(f, 3) == 1.322 and r == 1.0
    f, r = cortical_eligibility(MOD_5HT, 0)
    assert f == 1.0 and round(r, 3) == 0.732
    assert round(cortical_eligibility(MOD_NE, 10000)[0], 3) == 1.080
    assert round(cortical_eligibility(MOD_5HT, 5000)[1], 3) == 0.933
    assert cortical_eligibility(MOD_NE, 0, pair=False)[0] == 1.0
    assert cortical_eligibility(MOD_NE, 0, receptors=MOD_5HT)[0] == 1.0


def striatal(offset_ms):
    """60 pairings at 0.1 Hz: one arrival 10 ms before a three-spike burst;
    dopamine once per cycle at `offset_ms` (None: no dopamine)."""
    g = synapse_gene(1.0, 0.55, 450.0, 0.0, delay=1, plasticity=4)
    g['mod_receptors'] = MOD_DA
    s = fresh_syn(g)
    mod_phase = None if offset_ms is None else (10000 + offset_ms if offset_ms < 0 else offset_ms)
    for t in range(600000):
        phase = t % 10000
        plast_step(g, s, phase == 0, phase in (10, 20, 30),
                   MOD_DA if mod_phase is not None and phase == mod_phase else 0)
    return weight(s)


@pytest.mark.slow
def test_d1_profile_without_l_type_block():
    assert round(striatal(None), 3) == 0.677
    assert round(striatal(2000), 3) == 0.949


def spine(offset_ms, pair=True):
    """Ten pairings at 10 Hz from 1 s, dopamine: ten 30 Hz pulses at offset."""
    g = synapse_gene(1.0, 0.55, 450.0, 0.0, delay=1, plasticity=5)
    g['mod_receptors'] = MOD_DA
    stim, nb = fresh_syn(g), fresh_syn(g)
    for t in range(6501):
        training = pair and 1000 <= t < 2000
        phase = t - 1000
        pre = training and phase % 100 == 0
        post = training and phase % 100 == 10
        since = t - 1000 - (offset_ms if offset_ms is not None else 0)
        da = offset_ms is not None and 0 <= since <= 300 and since % 33 == 0
        plast_step(g, stim, pre, post, MOD_DA if da else 0)
        plast_step(g, nb, False, post, MOD_DA if da else 0)
    return float(stim['f'][0][7]), float(nb['f'][0][7])


def test_spine_profile_timing():
    assert spine(None)[0] == 1.0
    assert spine(-1000)[0] == 1.0
    assert round(spine(600)[0], 3) == 1.216
    v, neighbor = spine(1000)
    assert round(v, 3) == 1.241 and neighbor == 1.0
    assert round(spine(5000)[0], 3) == 1.0      # printed as 1.000
