"""GPU: the live population (include/mq_live.h; mqviz Sim::block, engine.rs:448-538) through the C ABI.

* Transport: ions spawned into the live pool are the batched path's ions. With dt chosen so that the RF advance
  per step is dyadic (0.0625 cycles), ions spawned at step 0 and a second group spawned a whole number of RF
  periods later (and no excitation, whose phase the live clock sets) end bit-identical to the same job's ions
  in mq_run_batch_ions, their random streams numbered by spawn order.
* Host stream: without recirculation the block's draws (Poisson count, species by weight) depend on the seed
  only; a Python restatement of the reference's xorshift64, poisson and pick_species (engine.rs:318-328,
  366-370) predicts the number spawned per species exactly.
* Bookkeeping: alive + finished = spawned; ion-steps equal the sum of the ions' steps; with recirculation and a
  zero mean delay every returned ion re-enters in the same block; tags split the tallies of an AC sweep.
* Sweep (start_sweep, stop_sweep and the block's sweep bookkeeping, engine.rs:420-446, 512-535): on a clock where
  every tone's phase is 0 at block boundaries, each point's ions equal, bit for bit, the same-numbered ions of a
  population whose live tone is that point's frequency, and its counts are theirs; the live tone's ions are moved
  to each new frequency with a continuous phase; the state machine (101-point cap, a single point when the upper
  limit is below the lower, passes, drain, stop before and after the end, bad configurations) as the reference's.
"""
import math
import sys

import numpy as np
import pytest

import mqhip
import synth

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    e = mqhip.Engine(synth.c1_grid())
    yield e
    e.close()


def physics(dt=1e-7, max_steps=3000):
    return mqhip.Physics.correct(0.133322).copy(dt=dt, max_steps=max_steps, injection=mqhip.INJ_GAUSSIAN,
                                                inj_sigma_mm=0.1, legacy_force=1)


SP = [mqhip.Species(499.0, 1.0, 230.0), mqhip.Species(505.0, 1.0, 235.0), mqhip.Species(2 * 863.473, 2.0, 425.0)]


def test_spawned_ions_are_the_batched_paths_ions(eng):
    ph = physics(dt=1.25e-7)       # f_rf dt = 0.0625: one RF period is 16 steps
    rf = mqhip.RfAc(100.0, 500e3, 0.0, 100e3)
    dc = [0, 5, 0, 0]
    seed, n, B = 777, 1500, 64
    live = mqhip.Live(eng, SP[:1], None, rf, dc, ph, block_steps=B, capacity=2 * n, seed=seed)
    try:
        live.inject(0, n)
        st = live.run(3)                # 192 steps: 12 RF periods
        assert st["step"] == 3 * B and st["injected"] == n
        live.inject(0, n)
        while live.run(16)["alive"]:
            pass
        st = live.run(0)
        s = live.slots()
    finally:
        live.close()
    say(f"[gpu] live pool: {st}")
    assert st["finished"] == 2 * n and st["alive"] == 0
    assert np.array_equal(s["sid"], np.arange(2 * n)) and np.all(s["species"] == 0)
    assert np.all(s["birth"][:n] == 0) and np.all(s["birth"][n:] == 3 * B)
    want = eng.run_batch_ions([mqhip.job(SP[0].mass, SP[0].z, SP[0].ccs_n2, rf.v_rf, rf.f_rf, 0.0, rf.f_ac, 2 * n,
                                         seed)], dc, ph)
    for k in ("x", "y", "z", "vx", "vy", "vz", "steps", "ncoll", "status"):
        assert np.array_equal(s[k], want[k]), k
    assert st["ion_steps"] == int(want["steps"].sum())


def xs_u01(r):
    r ^= (r << 13) & 0xFFFFFFFFFFFFFFFF
    r ^= r >> 7
    r ^= (r << 17) & 0xFFFFFFFFFFFFFFFF
    return r, (r >> 11) / float(1 << 53)


def poisson(r, lam):
    """engine.rs:318-328."""
    if lam < 30.0:
        l, k, p = math.exp(-lam), 0, 1.0
        while True:
            r, u = xs_u01(r)
            p *= u
            if p <= l:
                return r, k
            k += 1
    r, u1 = xs_u01(r)
    r, u2 = xs_u01(r)
    g = math.sqrt(-2.0 * math.log(max(u1, 1e-300))) * math.cos(2 * math.pi * u2)
    return r, int(max(0.0, round(lam + math.sqrt(lam) * g)))


@pytest.mark.parametrize("rate", [3e5, 4e6])
def test_host_stream_spawns_as_the_reference_draws(eng, rate):
    ph = physics()
    rf = mqhip.RfAc(100.0, 500e3, 0.5, 100e3)
    w = [0.5, 0.3, 0.2]
    seed, B, nb = 4242, 128, 40
    live = mqhip.Live(eng, SP, w, rf, [0, 5, 0, 0], ph, block_steps=B, capacity=200000, inj_rate=rate, seed=seed)
    try:
        st = live.run(nb)
        tal = live.tally(0)
        s = live.slots()
    finally:
        live.close()
    # replay of the block's draws: Poisson(rate B dt) spawns, species by the cumulative weights
    r, cdf = seed | 1, np.cumsum(np.array(w) / sum(w))
    per = [0, 0, 0]
    for _ in range(nb):
        r, k = poisson(r, rate * B * ph.dt)
        for _ in range(k):
            r, u = xs_u01(r)
            j = next((i for i, c in enumerate(cdf) if u <= c), 2)
            per[j] += 1
    say(f"[gpu] live spawns {st['injected']} (replay {sum(per)}), per species {per}; state {st}")
    assert st["injected"] == sum(per) and st["dropped"] == 0
    alive_sp = [int(np.sum((s["status"] == 0) & (s["species"] == k))) for k in range(3)]
    assert [tal[k].n + alive_sp[k] for k in range(3)] == per
    assert st["alive"] + st["finished"] == st["injected"]
    assert st["alive"] == int(np.sum(s["status"] == 0))
    assert sum(o.n for o in tal) == st["finished"]
    # the ion-steps of the run: every finished ion's steps plus every live ion's age
    live_age = int(np.sum(s["steps"][s["status"] == 0]))
    assert st["ion_steps"] == sum(int(round(o.mean_time / ph.dt * o.n)) for o in tal) + live_age


def test_recirculation_and_tags(eng):
    ph = physics()
    rf = mqhip.RfAc(100.0, 500e3, 0.5, 100e3)
    B = 128
    live = mqhip.Live(eng, SP[:2], None, rf, [0, 5, 0, 0], ph, block_steps=B, capacity=100000, inj_rate=2e6,
                      recirculate=True, residence_s=0.0, max_tags=3, seed=99)
    try:
        st = live.run(20)
        t0 = live.tally(0)
        # a returned ion with a zero mean delay is due at its own return step: it re-enters in that block
        assert st["reentered"] == sum(o.returned for o in t0) > 0
        live.set_ac(95e3, 1)
        st1 = live.run(20)
        live.set_ac(105e3, 2)
        st2 = live.run(20)
        tags = [live.tally(k) for k in range(3)]
    finally:
        live.close()
    say(f"[gpu] live with recirculation: {st2}")
    assert st2["alive"] + st2["finished"] == st2["injected"] + st2["reentered"]
    assert st2["reentered"] == sum(o.returned for t in tags for o in t)
    assert sum(o.n for t in tags for o in t) == st2["finished"]
    assert all(sum(o.n for o in t) > 0 for t in tags)
    assert st1["step"] == 40 * B and st2["step"] == 60 * B


def test_live_rejects_bad_options(eng):
    rf = mqhip.RfAc(100.0, 500e3, 0.5, 100e3)
    with pytest.raises(mqhip.MqError, match="multiple of 4"):
        mqhip.Live(eng, SP[:1], None, rf, [0, 5, 0, 0], physics(), block_steps=30)
    with pytest.raises(mqhip.MqError, match="Philox"):
        mqhip.Live(eng, SP[:1], None, rf, [0, 5, 0, 0], physics().copy(rng=mqhip.RNG_COMPAT))
    live = mqhip.Live(eng, SP[:1], None, rf, [0, 5, 0, 0], physics(), capacity=10)
    try:
        live.inject(0, 25)
        st = live.run(0)
        assert st["injected"] == 10 and st["dropped"] == 15 and st["alive"] == 10
        with pytest.raises(mqhip.MqError, match="species"):
            live.inject(3, 1)
    finally:
        live.close()


def _sweep_pop(eng, f_ac, sweep, n0=300, n1=200, ipp=400, rate=1.5e6, seed=31):
    """A population on the RF-aligned clock (f_rf dt = 1/16, 64-step blocks: four RF periods) with an
    injected group at step 0, two blocks, then optionally a sweep of 125 and 250 kHz, an injected group and
    blocks until the sweep has ended (or as many blocks as `sweep` says when it is an int)."""
    ph = physics(dt=1.25e-7)
    rf = mqhip.RfAc(100.0, 500e3, 0.5, f_ac)
    cap = 60000
    live = mqhip.Live(eng, SP[:2], [0.6, 0.4], rf, [0, 5, 0, 0], ph, block_steps=64, capacity=cap, inj_rate=rate,
                      seed=seed)
    live.inject(0, n0)
    pre = live.run(2)
    if sweep is True:
        live.sweep_start(125.0, 250.0, 125.0, ipp, 1)
    live.inject(1, n1)
    nb = 2
    while True:
        st = live.run(1)
        nb += 1
        sw = live.sweep()
        if (sweep is True and not sw["running"]) or (sweep is not True and nb >= sweep):
            break
        assert nb < 400
    assert st["injected"] + st["reentered"] < cap and st["dropped"] == 0   # every ion keeps its own slot
    return live, pre, st, nb


def test_sweep_points_run_on_their_own_tones(eng):
    """Sweep ions are the ions of a population whose live tone is the point's frequency.

    Tones of 125 and 250 kHz advance 1 and 2 cycles per 64-step block, so every tone's phase at a block
    boundary is 0, and a population without a sweep at that frequency gives its ions the same birth phases.
    Populations with the same seed draw the same spawns in the same order until the sweep drains, and with
    capacity to spare ion k sits in slot k: every ion of sweep point p equals the ion of the same birth order
    in the population at f_p, bit for bit. The live tone starts at 93.75 kHz (0.75 cycles a block), so the
    untagged ions of the first two blocks carry AC phase 0.5 after the sweep's start moves them to 125 kHz.
    """
    live, pre, st, nb = _sweep_pop(eng, 93.75e3, True)
    try:
        sw = live.sweep()
        s = live.slots()
    finally:
        live.close()
    say(f"[gpu] live sweep: {nb} blocks, {st}, sweep {dict((k, v) for k, v in sw.items() if k != 'acc')}")
    assert sw["npoints"] == 2 and list(sw["f_khz"]) == [125.0, 250.0]
    assert not sw["running"] and not sw["draining"] and sw["point"] == 1 and sw["sweeps_done"] == 1
    n0, n1, ipp = 300, 200, 400
    n_pre = pre["injected"]                                   # spawned before the sweep: the live tone
    nsp = n_pre + n1 + 2 * ipp                                # then n1 injected and 2 ipp Poisson ions on sweep tones
    tone = np.zeros(st["injected"], np.int64)
    tone[n_pre:n_pre + n1 + ipp] = 1
    tone[n_pre + n1 + ipp:nsp] = 2
    assert np.array_equal(s["sid"][:st["injected"]], np.arange(st["injected"]))
    refs = {}
    for f, t in ((125e3, 1), (250e3, 2)):
        ref, _, rst, _ = _sweep_pop(eng, f, nb)
        try:
            refs[t] = ref.slots()
        finally:
            ref.close()
        sel = np.flatnonzero(tone == t)
        for k in ("x", "y", "z", "vx", "vy", "vz", "steps", "ncoll", "status", "phi_ac0", "species"):
            assert np.array_equal(s[k][sel], refs[t][k][sel]), (f, k)
        assert np.all(s["status"][sel] != 0)
        # the point's counts: transmitted, crashed, returned, other, per species
        for sp in range(2):
            st_sp = s["status"][sel][s["species"][sel] == sp]
            want = [np.sum(st_sp == mqhip.TRANSMITTED), np.sum(st_sp == mqhip.CRASHED),
                    np.sum(st_sp == mqhip.RETURNED), 0]
            want[3] = len(st_sp) - sum(want)
            assert list(sw["acc"][t - 1, sp]) == want, (t, sp)
        say(f"[gpu]   point {t - 1} ({f / 1e3:.0f} kHz): {len(sel)} ions, acc {sw['acc'][t - 1].tolist()}")
    # fresh ions per point: exactly ions_per_point (the injected group of point 0 does not count)
    assert sw["acc"][0].sum() == n1 + ipp and sw["acc"][1].sum() == ipp
    # the untagged ions: born on the 93.75 kHz tone (phase 0, 0.75, 0.5 at steps 0, 64, 128); those alive when the
    # sweep started at step 128 moved to 125 kHz at its phase 0.5, and kept it (1 and 2 cycles a block) through the
    # move to 250 kHz; every one of them ended before the sweep did
    b = s["birth"][:n_pre].astype(np.int64)
    alive = s["steps"][:n_pre].astype(np.int64) > 128 - b
    born = np.where(b == 0, 0.0, np.where(b == 64, 0.75, 0.5))
    assert set(b.tolist()) <= {0, 64, 128} and np.sum(b == 0) == n0
    assert np.array_equal(s["phi_ac0"][:n_pre], np.where(alive, 0.5, born))
    assert np.all(s["phi_ac0"][n_pre:nsp] == 0.0)


def test_sweep_state_machine(eng):
    """start_sweep / stop_sweep / the block's bookkeeping (engine.rs:420-446, 512-535) through the C ABI."""
    ph = physics(dt=1.25e-7)
    rf = mqhip.RfAc(100.0, 500e3, 0.5, 100e3)
    live = mqhip.Live(eng, SP[:1], None, rf, [0, 5, 0, 0], ph, block_steps=64, capacity=100000, inj_rate=2e6,
                      seed=0x5DEECE66D)
    try:
        assert live.sweep()["npoints"] == 0
        live.sweep_stop()                                    # no sweep yet: nothing happens
        assert live.sweep()["npoints"] == 0
        # 301 steps requested: 101 points at most (start_sweep, engine.rs:423-424)
        live.sweep_start(70.0, 100.0, 0.1, 1000, 1)
        sw = live.sweep()
        assert sw["npoints"] == 101 and sw["f_khz"][100] == 70.0 + 100 * 0.1 and sw["running"] == 1
        with pytest.raises(mqhip.MqError, match="sweep is running"):
            live.sweep_start(80.0, 90.0, 1.0, 10, 1)
        live.sweep_stop()
        sw = live.sweep()
        assert sw["draining"] == 1 and sw["running"] == 1
        inj = live.run(0)["injected"]
        st = live.run(1)
        # draining: no fresh ions; no ion was on a sweep tone, so the sweep ends in that block
        assert st["injected"] == inj
        sw = live.sweep()
        assert sw["running"] == 0 and sw["draining"] == 0 and sw["point"] == 100
        assert live.run(3)["injected"] > inj                  # the live beam resumes
        # an upper limit below the lower one: a single point
        live.sweep_start(90.0, 80.0, 1.0, 50, 2)
        sw = live.sweep()
        assert sw["npoints"] == 1 and list(sw["f_khz"]) == [90.0]
        # two passes over one point of 50 fresh ions, then the drain
        prog = []
        while live.sweep()["running"]:
            live.run(1)
            prog.append(live.sweep()["progress"])
        sw = live.sweep()
        assert sw["sweeps_done"] == 2 and sw["acc"].sum() == 100 and prog[-1] == 1.0
        assert all(0.0 <= a <= b <= 1.0 for a, b in zip(prog, prog[1:]))
        # stop after the end: one block without fresh ions, then the beam again (stop_sweep, engine.rs:431-432)
        live.sweep_stop()
        assert live.sweep()["running"] == 1 and live.sweep()["draining"] == 1
        inj = live.run(0)["injected"]
        assert live.run(1)["injected"] == inj and live.sweep()["running"] == 0
        for bad in ((80.0, 90.0, 0.0, 10, 1), (80.0, 90.0, 1.0, 0, 1), (80.0, 90.0, 1.0, 10, 0),
                    (float("nan"), 90.0, 1.0, 10, 1), (-5.0, 90.0, 1.0, 10, 1)):
            with pytest.raises(mqhip.MqError):
                live.sweep_start(*bad)
    finally:
        live.close()
