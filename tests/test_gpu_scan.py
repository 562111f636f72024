"""GPU: scan::cached_batch and scan::adaptive_scan through the C ABI (scan.rs:31-78)."""
import numpy as np
import pytest

import mqhip
import synth

pytestmark = pytest.mark.gpu

GEOM = "Geom { synthetic ideal channel 64x64x256 @ 0.1 mm }"


@pytest.fixture(scope="module")
def eng():
    e = mqhip.Engine(synth.c1_grid())
    yield e
    e.close()


def _ph():
    return mqhip.Physics.correct(0.133322).copy(dt=1e-7, max_steps=3000, injection=mqhip.INJ_GAUSSIAN,
                                                inj_sigma_mm=0.1, legacy_force=1)


def test_cached_batch_roundtrip(eng, tmp_path):
    print("[gpu] cached_batch round trip", flush=True)
    cache = str(tmp_path / "results.tsv")
    ph = _ph()
    jobs = [mqhip.job(499.0 + 2 * i, 1.0, 230.0, 100.0, 500e3, 0.5, 100e3, 200, 1000 + i) for i in range(5)]
    want = eng.run_batch(jobs, [0, 5, 0, 0], ph)
    got, computed = mqhip.cached_batch(eng, GEOM, jobs, [0, 5, 0, 0], ph, cache)
    assert computed == 5 and [o.as_dict() for o in got] == [o.as_dict() for o in want]
    lines = open(cache).read().splitlines()
    assert len(lines) == 5
    for line, j, o in zip(lines, jobs, want):
        key, val = line.split("\t")
        assert key == mqhip.scan_key(GEOM, j, [0, 5, 0, 0], ph)
        f = val.split(",")
        assert [int(x) for x in f[:6]] == [o.n, o.crashed, o.returned, o.transmitted, o.left, o.timeout]
        assert float(f[6]) == o.mean_coll and float(f[7]) == o.mean_time
    # all hits now, including a duplicated job; a new job runs alone
    again, computed = mqhip.cached_batch(eng, GEOM, jobs + [jobs[0]], [0, 5, 0, 0], ph, cache)
    assert computed == 0 and [o.as_dict() for o in again] == [o.as_dict() for o in want + [want[0]]]
    extra = mqhip.job(600.0, 1.0, 230.0, 100.0, 500e3, 0.5, 100e3, 100, 7)
    out, computed = mqhip.cached_batch(eng, GEOM, jobs[:2] + [extra], [0, 5, 0, 0], ph, cache)
    assert computed == 1 and out[2].n == 100 and len(open(cache).read().splitlines()) == 6
    # an entry written by someone else (reference line format) is used as is
    k = mqhip.scan_key(GEOM, mqhip.job(700.0, 1.0, 230.0, 100.0, 500e3, 0.5, 100e3, 10, 1), [0, 5, 0, 0], ph)
    with open(cache, "a") as fh:
        fh.write(f"{k}\t10,1,2,3,4,0,7.5,0.0012\n")
    out, computed = mqhip.cached_batch(eng, GEOM, [mqhip.job(700.0, 1.0, 230.0, 100.0, 500e3, 0.5, 100e3, 10, 1)],
                                       [0, 5, 0, 0], ph, cache)
    assert computed == 0 and (out[0].n, out[0].crashed, out[0].returned, out[0].transmitted) == (10, 1, 2, 3)
    assert out[0].mean_coll == 7.5 and out[0].mean_time == 0.0012


def test_adaptive_scan_structure(eng, tmp_path):
    print("[gpu] adaptive_scan", flush=True)
    cache = str(tmp_path / "results.tsv")
    ph = _ph()
    sp, rf = mqhip.Species(499.0, 1.0, 230.0), mqhip.RfAc(100.0, 500e3, 0.5, 100e3)
    # base = -1: every point with T >= 0.35 T_max is refined
    pts = mqhip.adaptive_scan(eng, GEOM, sp, rf, [0, 5, 0, 0], ph, 80.0, 90.0, 2.0, 50, 80, -1.0, 100, cache)
    fs = [80.0 + 2 * i for i in range(6)]
    coarse = eng.run_batch([mqhip.job(499.0, 1.0, 230.0, 100.0, 500e3, 0.5, f * 1e3, 50, 100 + i)
                            for i, f in enumerate(fs)], [0, 5, 0, 0], ph)
    t1 = [o.transmitted / o.n for o in coarse]
    fine = mqhip.adaptive_fine_points(fs, t1, 50, -1.0, 2.0)
    assert [p[0] for p in pts] == sorted(set(fine) | {f for f in fs if all(abs(f - g) >= 1e-6 for g in fine)})
    assert all(p[3] == (80 if p[0] in fine else 50) for p in pts)
    lines = open(cache).read().splitlines()
    assert len(lines) == len(fs) + len(fine)


def test_cached_batch_keeps_streams_and_paths_apart(eng, tmp_path):
    """A cached Philox / FP32-batched line is not returned to a COMPAT or FP64-plain request for the same
    job, and each request then reads back its own line."""
    print("[gpu] cached_batch: streams and paths keyed apart", flush=True)
    cache = str(tmp_path / "results.tsv")
    ph = _ph()
    jobs = [mqhip.job(499.0, 1.0, 230.0, 100.0, 500e3, 0.5, 100e3, 150, 11)]
    outs = {}
    for rng in (mqhip.RNG_PHILOX, mqhip.RNG_COMPAT):
        for path in (mqhip.PATH_BATCHED, mqhip.PATH_PLAIN):
            p2 = ph.copy(rng=rng)
            got, computed = mqhip.cached_batch(eng, GEOM, jobs, [0, 5, 0, 0], p2, cache, path)
            assert computed == 1, (rng, path)
            want = eng.run_batch(jobs, [0, 5, 0, 0], p2, path)
            assert got[0].as_dict() == want[0].as_dict()
            outs[(rng, path)] = got[0].as_dict()
    assert len(open(cache).read().splitlines()) == 4
    for (rng, path), o in outs.items():
        got, computed = mqhip.cached_batch(eng, GEOM, jobs, [0, 5, 0, 0], ph.copy(rng=rng), cache, path)
        assert computed == 0 and got[0].as_dict() == o
