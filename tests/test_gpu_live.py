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
