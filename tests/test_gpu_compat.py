"""GPU runs exported as the reference's 64-byte ChIon records and hashed as
check.rs does (check.rs:6-21 compares records, :100-104 hashes them).

* FP64 plain path: the FNV hash of the exported records equals the hash of
  the oracle's states exported the same way (the states are bit-identical).
* Both paths: a job run alone and the same job inside a larger batch give
  byte-identical records (the check-tiled comparison, check.rs:24-62).
* summarize over records equals run_batch's outcomes.
"""
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
def c1():
    g = synth.c1_grid()
    say("[gpu] compat: creating engine on the 64x64x256 grid")
    e = mqhip.Engine(g)
    yield g, e
    e.close()


def _ph(legacy):
    if legacy:
        return mqhip.Physics.legacy(1e-3 * 133.322).copy(max_steps=20000)
    return mqhip.Physics.correct(0.133322).copy(dt=1e-7, max_steps=20000, injection=mqhip.INJ_GAUSSIAN,
                                                inj_sigma_mm=0.1, legacy_force=1)


def _jobs(legacy):
    if legacy:
        return [mqhip.job(5 * 499.0, 5.0, 0.0, 100.0, 500e3, 0.5, 100e3, 150, 1499),
                mqhip.job(5 * 507.0, 5.0, 0.0, 100.0, 500e3, 0.5, 100e3, 90, 1507),
                mqhip.job(5 * 523.0, 5.0, 0.0, 100.0, 500e3, 0.5, 100e3, 60, 1523)]
    return [mqhip.job(499.0, 1.0, 230.0, 100.0, 500e3, 0.5, 100e3, 150, 1499),
            mqhip.job(505.0, 1.0, 235.0, 100.0, 480e3, 0.8, 95e3, 90, 77, 0.5),
            mqhip.job(477.0, 1.0, 226.0, 100.0, 520e3, 0.5, 101e3, 60, 1477)]


@pytest.mark.parametrize("legacy", [0, 1])
def test_plain_hash_equals_oracle_hash(c1, oracle_mod, legacy):
    g, e = c1
    ph, jobs = _ph(legacy), _jobs(legacy)
    say(f"[gpu] compat hash, plain vs oracle, legacy={legacy}")
    got = e.run_batch_ions(jobs, [0, 5, 0, 0], ph, mqhip.PATH_PLAIN)
    want = oracle_mod.run_batch(g, jobs, [0, 5, 0, 0], ph, 8)
    want = dict(want, phi_ac0=want["phi_ac"])
    rg = mqhip.export_chion(got, jobs, ph)
    rw = mqhip.export_chion(want, jobs, ph)
    assert mqhip.fnv1a64(rg) == mqhip.fnv1a64(rw)
    assert rg.tobytes() == rw.tobytes()


@pytest.mark.parametrize("path", [mqhip.PATH_PLAIN, mqhip.PATH_BATCHED])
@pytest.mark.parametrize("legacy", [0, 1])
def test_job_alone_equals_job_in_batch(c1, path, legacy):
    g, e = c1
    ph, jobs = _ph(legacy), _jobs(legacy)
    say(f"[gpu] compat check-tiled: path={path} legacy={legacy}")
    batch = e.run_batch_ions(jobs, [0, 5, 0, 0], ph, path)
    rb = mqhip.export_chion(batch, jobs, ph)
    off = 0
    for k, j in enumerate(jobs):
        alone = e.run_batch_ions([j], [0, 5, 0, 0], ph, path)
        ra = mqhip.export_chion(alone, [j], ph)
        assert ra.tobytes() == rb[off:off + j.n].tobytes(), f"job {k}"
        off += j.n


@pytest.mark.parametrize("path", [mqhip.PATH_PLAIN, mqhip.PATH_BATCHED])
def test_summarize_records_equals_run_batch(c1, path):
    g, e = c1
    ph, jobs = _ph(0), _jobs(0)
    say(f"[gpu] compat summarize over records: path={path}")
    st = e.run_batch_ions(jobs, [0, 5, 0, 0], ph, path)
    out = e.run_batch(jobs, [0, 5, 0, 0], ph, path)
    rec = mqhip.summarize_chion(mqhip.export_chion(st, jobs, ph), st["job"], len(jobs))
    for a, b in zip(rec, out):
        assert [a.n, a.crashed, a.returned, a.transmitted, a.left, a.timeout] == \
               [b.n, b.crashed, b.returned, b.transmitted, b.left, b.timeout]
        assert a.mean_coll == pytest.approx(b.mean_coll, rel=1e-12)
        assert a.mean_time == pytest.approx(b.mean_time, rel=1e-6)
