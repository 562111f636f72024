"""Record compatibility: the 64-byte ChIon export (channel.rs:292-294),
summarize over records (channel.rs:496-506) and the FNV-1a hash of check.rs:100-104.
Host-only entry points: these run without a GPU."""
import numpy as np
import pytest

import mqhip


def test_fnv1a64_published_vectors():
    # FNV-1a 64-bit test vectors of the FNV reference suite (Noll).
    assert mqhip.fnv1a64(b"") == 0xCBF29CE484222325
    assert mqhip.fnv1a64(b"a") == 0xAF63DC4C8601EC8C
    assert mqhip.fnv1a64(b"foobar") == 0x85944171F73967E8


def test_fnv1a64_matches_bytewise_definition():
    data = np.random.default_rng(5).integers(0, 256, 1000).astype(np.uint8).tobytes()
    h = 14695981039346656037
    for b in data:
        h = ((h ^ b) * 1099511628211) % (1 << 64)
    assert mqhip.fnv1a64(data) == h


def _states(rng, jobs):
    n = sum(j.n for j in jobs)
    o = {
        "x": rng.uniform(-5e-3, 5e-3, n), "y": rng.uniform(-5e-3, 5e-3, n), "z": rng.uniform(0, 2.5e-2, n),
        "vx": rng.normal(0, 300, n), "vy": rng.normal(0, 300, n), "vz": rng.normal(0, 300, n),
        "steps": rng.integers(0, 20001, n).astype(np.uint32), "ncoll": rng.integers(0, 50, n).astype(np.uint32),
        "status": rng.integers(1, 6, n).astype(np.uint8), "phi_ac0": rng.uniform(0, 1, n),
        "cd": rng.integers(0, 30, n).astype(np.uint32),
        "job": np.repeat(np.arange(len(jobs), dtype=np.uint32), [j.n for j in jobs]),
    }
    return o


def _jobs():
    return [mqhip.job(499.0, 1.0, 0.0, 100.0, 500e3, 0.5, 100e3, 7, 1499),
            mqhip.job(2495.0, 5.0, 0.0, 90.0, 480e3, 0.3, 91e3, 5, (1 << 40) + 3)]


def test_export_fields_follow_the_record_layout():
    jobs = _jobs()
    ph = mqhip.Physics.correct(0.1)
    ph.dt = 1e-7
    o = _states(np.random.default_rng(1), jobs)
    r = mqhip.export_chion(dict(o), jobs, ph)
    assert r.dtype.itemsize == 64 and len(r) == 12
    # fixed point: nearest 62.5 pm quantum
    for k, c in enumerate("xyz"):
        assert np.array_equal(r["p"][:, k], np.rint(o[c] * mqhip.POS_SCALE).astype(np.int32))
        assert np.all(np.abs(r["p"][:, k] / mqhip.POS_SCALE - o[c]) <= 0.5 / mqhip.POS_SCALE * (1 + 1e-9))
    assert np.array_equal(r["p"][:, 3], o["status"].astype(np.int32))
    for k, c in enumerate(("vx", "vy", "vz")):
        assert np.array_equal(r["v"][:, k], o[c].astype(np.float32))
    assert np.array_equal(r["v"][:, 3], (o["steps"] * ph.dt).astype(np.float32))
    f_ac = np.repeat([j.drive.f_ac for j in jobs], [j.n for j in jobs])
    f_rf = np.repeat([j.drive.f_rf for j in jobs], [j.n for j in jobs])
    ac = (o["phi_ac0"] + f_ac * ph.dt * o["steps"]) % 1.0
    rf = (f_rf * ph.dt * o["steps"]) % 1.0
    assert np.all(np.abs(r["aux"][:, 0] - ac) < 1e-6)
    assert np.all(np.abs(r["aux"][:, 3] - rf) < 1e-6)
    assert np.all((r["aux"][:, 0] >= 0) & (r["aux"][:, 0] <= 1)) and np.all(r["aux"][:, 3] >= 0)
    assert np.array_equal(r["aux"][:, 1], o["ncoll"].astype(np.float32))
    assert np.array_equal(r["aux"][:, 2], o["cd"].astype(np.float32))
    seeds = np.repeat([j.seed for j in jobs], [j.n for j in jobs]).astype(np.uint64)
    assert np.array_equal(r["rng"][:, 0], (seeds & 0xFFFFFFFF).astype(np.uint32))
    assert np.array_equal(r["rng"][:, 1], (seeds >> 32).astype(np.uint32))
    assert list(r["rng"][:, 2]) == list(range(7)) + list(range(5))
    assert np.array_equal(r["rng"][:, 3], o["steps"])


def test_export_without_optional_arrays_writes_zero():
    jobs = _jobs()
    ph = mqhip.Physics.correct(0.1)
    o = _states(np.random.default_rng(2), jobs)
    del o["phi_ac0"], o["cd"]
    r = mqhip.export_chion(o, jobs, ph)
    assert np.all(r["aux"][:, 0] == 0) and np.all(r["aux"][:, 2] == 0)


def test_export_rejects_ions_out_of_job_order():
    jobs = _jobs()
    o = _states(np.random.default_rng(3), jobs)
    o["job"] = o["job"][::-1].copy()
    with pytest.raises(mqhip.MqError, match="job order"):
        mqhip.export_chion(o, jobs, mqhip.Physics.correct(0.1))


def test_export_rejects_positions_beyond_the_fixed_point_range():
    jobs = _jobs()
    o = _states(np.random.default_rng(4), jobs)
    o["x"][3] = 0.2  # 0.2 m * 1.6e10 > 2^31
    with pytest.raises(mqhip.MqError, match="fixed-point"):
        mqhip.export_chion(o, jobs, mqhip.Physics.correct(0.1))


def test_summarize_over_records_equals_summarize_over_states():
    jobs = _jobs()
    ph = mqhip.Physics.correct(0.1)
    ph.dt = 1e-7
    o = _states(np.random.default_rng(6), jobs)
    r = mqhip.export_chion(dict(o), jobs, ph)
    a = mqhip.summarize_chion(r, o["job"], len(jobs))
    b = mqhip.summarize(o["status"], o["ncoll"], o["steps"], o["job"], len(jobs), ph.dt)
    for x, y in zip(a, b):
        assert [x.n, x.crashed, x.returned, x.transmitted, x.left, x.timeout] == \
               [y.n, y.crashed, y.returned, y.transmitted, y.left, y.timeout]
        assert x.mean_coll == pytest.approx(y.mean_coll, rel=1e-12)
        # v[3] is the f32 time, as in the reference record
        assert x.mean_time == pytest.approx(y.mean_time, rel=1e-6)
