"""GPU: mq_run_batch_ions_noalloc, the entry point that allocates no device memory, against
mq_run_batch_ions through the C ABI. Per-ion results are bit-identical on both paths, on two calls in a
row (the second smaller, so the scratch array holds the first call's data when it starts), and sizes
beyond the maximum are rejected with an error."""
import sys

import numpy as np
import pytest

import mqhip
import synth

pytestmark = pytest.mark.gpu

KEYS = ("x", "y", "z", "vx", "vy", "vz", "steps", "ncoll", "status", "job", "phi_ac0", "cd")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    say("[gpu] no-alloc: engine on the 64x64x256 grid")
    e = mqhip.Engine(synth.c1_grid())
    yield e
    e.close()


def _ph(case):
    base = mqhip.Physics.correct(0.133322).copy(dt=1e-7, max_steps=20000, injection=mqhip.INJ_GAUSSIAN,
                                                inj_sigma_mm=0.1, legacy_force=1)
    return base if case == "thermal" else mqhip.Physics.legacy(1e-3 * 133.322).copy(max_steps=20000)


def _jobs(case, n):
    if case == "thermal":
        return [mqhip.job(499.0, 1.0, 230.0, 100.0, 500e3, 0.5, 100e3, n, 1499),
                mqhip.job(505.0, 1.0, 235.0, 100.0, 480e3, 0.8, 95e3, n // 2, 77, 0.5)]
    return [mqhip.job(5 * 499.0, 5.0, 0.0, 100.0, 500e3, 0.5, 100e3, n, 1499),
            mqhip.job(5 * 503.0, 5.0, 0.0, 100.0, 500e3, 0.5, 100e3, n // 3, 1503)]


def _same(a, b, what):
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), (what, k)


@pytest.mark.parametrize("path", [mqhip.PATH_BATCHED, mqhip.PATH_PLAIN])
@pytest.mark.parametrize("case", ["thermal", "legacy"])
def test_noalloc_matches_run_batch_ions(eng, case, path):
    ph = _ph(case)
    dc = [0, 5, 0, 0]
    for n in (20000, 3001):   # two calls in a row, the second smaller
        jobs = _jobs(case, n)
        say(f"[gpu] no-alloc vs run_batch_ions: {case}, path {path}, {sum(j.n for j in jobs)} ions")
        want = eng.run_batch_ions(jobs, dc, ph, path)
        got = eng.run_batch_ions(jobs, dc, ph, path, noalloc=True)
        _same(want, got, (case, path, n))
    # a new DC configuration is recomputed in place; then the same call again
    jobs = _jobs(case, 500)
    got = eng.run_batch_ions(jobs, [0, 3, 0, 0], ph, path, noalloc=True)
    want = eng.run_batch_ions(jobs, [0, 3, 0, 0], ph, path)
    _same(want, got, (case, path, "dc"))


def test_noalloc_across_tunings(eng):
    """The no-alloc path runs every tuning (graph mode runs uncaptured) to the same results."""
    ph = _ph("thermal")
    jobs = _jobs("thermal", 4000)
    want = eng.run_batch_ions(jobs, [0, 5, 0, 0], ph)
    t0 = eng.get_tune()
    try:
        for v in (dict(block=64, waves_per_cu=3, refill_period=4, chunk=100, chunk_cap=400, resort=0, sync_every=1,
                       graph=0),
                  dict(block=128, waves_per_cu=8, refill_period=16, chunk=64, chunk_cap=64, resort=1, sync_every=4,
                       graph=1),
                  dict(block=256, waves_per_cu=8, refill_period=8, chunk=64, chunk_cap=64, resort=2, sync_every=4,
                       graph=0)):
            eng.set_tune(**v)
            say(f"[gpu] no-alloc under tune {v}")
            _same(want, eng.run_batch_ions(jobs, [0, 5, 0, 0], ph, noalloc=True), v)
    finally:
        eng.set_tune(**t0)


def test_noalloc_rejects_what_it_cannot_run(eng):
    ph = _ph("thermal")
    say("[gpu] no-alloc limits")
    big = [mqhip.job(499.0, 1.0, 230.0, 100.0, 500e3, 0.5, 100e3, mqhip.NOALLOC_MAX_IONS + 1, 1)]
    with pytest.raises(mqhip.MqError, match="maximum"):
        eng.run_batch_ions(big, [0, 5, 0, 0], ph, noalloc=True)
    many = [mqhip.job(499.0, 1.0, 230.0, 100.0, 500e3, 0.5, 100e3, 1, i) for i in range(mqhip.NOALLOC_MAX_JOBS + 1)]
    with pytest.raises(mqhip.MqError, match="maximum"):
        eng.run_batch_ions(many, [0, 5, 0, 0], ph, noalloc=True)
    with pytest.raises(mqhip.MqError, match="Philox"):
        eng.run_batch_ions(_jobs("thermal", 10), [0, 5, 0, 0], ph.copy(rng=mqhip.RNG_COMPAT), noalloc=True)
    # exactly the maximum runs
    top = [mqhip.job(499.0, 1.0, 230.0, 100.0, 500e3, 0.5, 100e3, mqhip.NOALLOC_MAX_IONS, 5)]
    r = eng.run_batch_ions(top, [0, 5, 0, 0], ph.copy(max_steps=300), noalloc=True)
    assert np.all(r["status"] != 0) and len(r["status"]) == mqhip.NOALLOC_MAX_IONS


def test_noalloc_needs_built_tables():
    """On a fresh engine the field tables do not exist yet: the call is refused, and succeeds after one
    ordinary call on the same path."""
    e = mqhip.Engine(synth.c1_grid())
    try:
        ph = _ph("thermal")
        jobs = _jobs("thermal", 200)
        say("[gpu] no-alloc on a fresh engine")
        with pytest.raises(mqhip.MqError, match="field tables"):
            e.run_batch_ions(jobs, [0, 5, 0, 0], ph, noalloc=True)
        want = e.run_batch_ions(jobs, [0, 5, 0, 0], ph)
        _same(want, e.run_batch_ions(jobs, [0, 5, 0, 0], ph, noalloc=True), "fresh")
    finally:
        e.close()
