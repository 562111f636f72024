"""GPU paths with the reference's random streams (MQ_RNG_COMPAT).

* FP64 plain path vs the oracle: every ion bit-identical (injection on the
  host from the job's xorshift64 stream, collisions from each ion's
  xoshiro128** stream in channel.metal's draw order).
* FP32 batched path: class counts within 3 sigma of the oracle; per-ion
  results independent of the batch composition, the launch tuning and the
  collision batching (the stream is the ion's own).
* The two stream families describe the same physics: PHILOX and COMPAT
  outcome counts agree within 3 sigma.
"""
import sys

import numpy as np
import pytest

import mqhip
import synth
from gridutil import binom_pull

pytestmark = pytest.mark.gpu
KEYS = ("x", "y", "z", "vx", "vy", "vz")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def c1():
    g = synth.c1_grid()
    say("[gpu] compat streams: creating engine on the 64x64x256 grid")
    e = mqhip.Engine(g)
    yield g, e
    e.close()


def physics(case):
    C = mqhip.RNG_COMPAT
    base = mqhip.Physics.correct(0.133322).copy(dt=1e-7, max_steps=20000, injection=mqhip.INJ_GAUSSIAN,
                                                inj_sigma_mm=0.1, rng=C)
    return {
        "full_thermal": base,
        "half_thermal": base.copy(legacy_force=1),
        "thermal_disc": mqhip.Physics.correct(0.133322).copy(max_steps=20000, legacy_force=1, rng=C),
        "legacy": mqhip.Physics.legacy(1e-3 * 133.322).copy(max_steps=20000, rng=C),
        "legacy_trilinear": mqhip.Physics.legacy(1e-3 * 133.322).copy(max_steps=20000, legacy_fields=0, rng=C),
    }[case]


def jobs_for(case, n):
    if case.startswith("legacy"):
        return [mqhip.job(5 * 499.0, 5.0, 0.0, 100.0, 500e3, 0.5, 100e3, n, 1499),
                mqhip.job(5 * 503.0, 5.0, 0.0, 100.0, 500e3, 0.5, 100e3, n // 2, 1503)]
    return [mqhip.job(499.0, 1.0, 230.0, 100.0, 500e3, 0.5, 100e3, n, 1499),
            mqhip.job(505.0, 1.0, 235.0, 100.0, 480e3, 0.8, 95e3, n // 2, 77, 0.5)]


@pytest.mark.parametrize("case", ["full_thermal", "half_thermal", "thermal_disc", "legacy", "legacy_trilinear"])
def test_plain_compat_matches_oracle(c1, oracle_mod, case):
    g, e = c1
    ph, jobs = physics(case), jobs_for(case, 200)
    say(f"[gpu] compat plain FP64 vs oracle: {case}")
    got = e.run_batch_ions(jobs, [0, 5, 0, 0], ph, mqhip.PATH_PLAIN)
    want = oracle_mod.run_batch(g, jobs, [0, 5, 0, 0], ph, 8)
    for k in ("job", "status", "steps", "ncoll") + KEYS:
        assert np.array_equal(got[k], want[k]), k
    assert np.array_equal(got["cd"], want["cd"])


@pytest.mark.parametrize("case", ["half_thermal", "legacy"])
def test_batched_compat_within_3sigma(c1, oracle_mod, case):
    g, e = c1
    ph, jobs = physics(case), jobs_for(case, 3000)
    say(f"[gpu] compat batched FP32 vs oracle: {case}")
    got = e.run_batch_ions(jobs, [0, 5, 0, 0], ph, mqhip.PATH_BATCHED)
    want = oracle_mod.run_batch(g, jobs, [0, 5, 0, 0], ph, 8)
    for j in range(len(jobs)):
        sg, sw = got["status"][got["job"] == j], want["status"][want["job"] == j]
        for k in range(1, 6):
            pull = binom_pull(int((sg == k).sum()), len(sg), int((sw == k).sum()), len(sw))
            assert abs(pull) <= 3.0, (j, k, pull)


def test_batched_compat_independent_of_batch_tuning_and_collision_batching(c1):
    g, e = c1
    ph, jobs = physics("half_thermal"), jobs_for("half_thermal", 400)
    say("[gpu] compat batched: independence")
    full = e.run_batch_ions(jobs, [0, 5, 0, 0], ph)
    alone = e.run_batch_ions(jobs[1:], [0, 5, 0, 0], ph)
    sel = full["job"] == 1
    for k in KEYS + ("steps", "ncoll", "status"):
        assert np.array_equal(full[k][sel], alone[k]), k
    t0 = e.get_tune()
    try:
        for v in (dict(block=64, waves_per_cu=4, refill_period=4, chunk=96, chunk_cap=96, resort=0, coll_batch=1),
                  dict(block=256, waves_per_cu=12, refill_period=16, chunk=512, chunk_cap=512, resort=1,
                       coll_batch=16),
                  # staged tiles are Philox-only: COMPAT runs keep the unstaged worker
                  dict(block=256, waves_per_cu=8, refill_period=8, chunk=64, chunk_cap=64, resort=2,
                       coll_batch=4)):
            e.set_tune(**v)
            r = e.run_batch_ions(jobs, [0, 5, 0, 0], ph)
            for k in KEYS + ("steps", "ncoll", "status"):
                assert np.array_equal(r[k], full[k]), (v, k)
    finally:
        e.set_tune(**t0)


def test_compat_and_philox_describe_the_same_physics(c1):
    g, e = c1
    for case in ("half_thermal", "legacy"):
        ph = physics(case)
        jobs = jobs_for(case, 4000)[:1]
        say(f"[gpu] compat vs philox streams: {case}")
        a = e.run_batch(jobs, [0, 5, 0, 0], ph, mqhip.PATH_BATCHED)[0]
        b = e.run_batch(jobs, [0, 5, 0, 0], ph.copy(rng=mqhip.RNG_PHILOX), mqhip.PATH_BATCHED)[0]
        for k in ("crashed", "returned", "transmitted", "left", "timeout"):
            assert abs(binom_pull(getattr(a, k), a.n, getattr(b, k), b.n)) <= 3.0, (case, k)


def test_compat_rejected_by_adjoint_and_tandem(c1):
    g, e = c1
    ph = mqhip.Physics.correct(0.0).copy(dt=1e-7, rng=mqhip.RNG_COMPAT)
    j = mqhip.job(499.0, 1.0, 0.0, 100.0, 500e3, 0.5, 100e3, 64, 1)
    with pytest.raises(mqhip.MqError, match="Philox"):
        e.forward(j, [0, 0.5, 0, 0], ph, mqhip.Objective(100, 0, 18.6e-3, 2e-5, 1e-3, 0.3))
