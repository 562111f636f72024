"""GPU tandem batch (include/mq_tandem.h) against the CPU oracle.

* FP64 plain path: every ion's final state, region and last port identical
  to the oracle's, and the per-port counters, the retry matrix and the totals
  equal exactly (integer bookkeeping of identical trajectories), on a run in
  which ions are forwarded through the ports and detected past the combiner.
* FP32 batched path: the same flow balance holds; per-port entries, returns,
  forwards and losses and the end classes (detections included) agree with
  the FP64 path within 4 sigma (Poisson); per-ion results do not depend on
  the batch composition.
"""
import sys

import numpy as np
import pytest

import mqhip
import synth
from test_oracle_tandem import closed_form_case, flow_checks

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def small():
    cube, port, geom, d, jobs, ph = synth.tandem_setup(mqhip, "small", n_per_charge=96, max_steps=15000)
    say("[gpu] tandem: creating the small configuration")
    t = mqhip.Tandem(cube, port, geom)
    yield t, (cube, port, geom, d, jobs, ph)
    t.close()


def test_plain_matches_oracle_exactly(small, oracle_mod):
    t, (cube, port, geom, d, jobs, ph) = small
    say("[gpu] tandem FP64 plain vs oracle, 4 x 96 ions")
    c, tt, s = t.run_ions(jobs, d, ph, mqhip.PATH_PLAIN)
    co, to, so = oracle_mod.tandem_run(cube, port, geom, jobs, d, ph, 8)
    assert tt == to
    assert c == co
    for k in ("status", "region", "last", "steps", "ncoll", "job"):
        assert np.array_equal(s[k], so[k]), k
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.array_equal(s[k], so[k]), k
    flow_checks(c, tt, s, len(jobs))
    assert sum(p["entered"] for p in c) > 50 and sum(p["returned"] for p in c) > 0
    assert sum(p["forwarded"] for p in c) > 0 and tt["detected"] > 0


def test_plain_cell_records_keep_results(small):
    """The FP64 path reads per-cell records of the four field sets (splitter, combiner, driven and idle port;
    mq_tandem_set_plain_recs(1), the default) or gathers the node tables (0): the counters and every ion's final
    state are the same bit for bit."""
    t, (cube, port, geom, d, jobs, ph) = small
    big = [mqhip.job(j.sp.mass, j.sp.z, j.sp.ccs_n2, j.drive.v_rf, j.drive.f_rf, j.drive.v_ac, j.drive.f_ac, 1024,
                     j.seed) for j in jobs]
    out = {}
    try:
        for on in (True, False):
            say(f"[gpu] tandem FP64, 4 x 1024 ions, cell records {on}")
            t.set_plain_recs(on)
            out[on] = t.run_ions(big, d, ph, mqhip.PATH_PLAIN)
            assert t.stats()["plain_recs"] == int(on)
    finally:
        t.set_plain_recs(True)
    (ca, ta, sa), (cb, tb, sb) = out[True], out[False]
    assert ta == tb and ca == cb
    for k in ("status", "region", "last", "steps", "ncoll", "job", "x", "y", "z", "vx", "vy", "vz"):
        assert np.array_equal(sa[k], sb[k]), k
    assert sum(p["entered"] for p in ca) > 0 and ta["detected"] > 0


@pytest.mark.parametrize("case", ["entry", "bounce", "forward_floor", "reverse", "from_other", "board_top",
                                  "board_mouth", "back", "wall", "detect"])
def test_closed_form_cases_match_oracle(oracle_mod, case):
    """The single-ion cases of tests/test_oracle_tandem.py (closed forms there), on the GPU's FP64 path."""
    cube, port, geom, d, jobs, ph = closed_form_case(case, oracle_mod)
    t = mqhip.Tandem(cube, port, geom)
    try:
        say(f"[gpu] tandem closed-form case {case}")
        c, tt, s = t.run_ions(jobs, d, ph, mqhip.PATH_PLAIN)
    finally:
        t.close()
    co, to, so = oracle_mod.tandem_run(cube, port, geom, jobs, d, ph, 1)
    assert c == co and tt == to
    for k in ("x", "y", "z", "vx", "vy", "vz", "steps", "status", "region", "last"):
        assert np.array_equal(s[k], so[k]), k


def test_batched_flow_balance_and_agreement(small):
    t, (cube, port, geom, d, jobs, ph) = small
    big = [mqhip.job(j.sp.mass, j.sp.z, j.sp.ccs_n2, j.drive.v_rf, j.drive.f_rf, j.drive.v_ac, j.drive.f_ac, 2048,
                     j.seed) for j in jobs]
    say("[gpu] tandem FP32 batched vs FP64 plain, 4 x 2048 ions")
    cb, tb, sb = t.run_ions(big, d, ph, mqhip.PATH_BATCHED)
    cp, tp = t.run(big, d, ph, mqhip.PATH_PLAIN)
    flow_checks(cb, tb, sb, len(big))
    say(f"   batched {tb}\n   plain   {tp}")

    def close(a, b):
        return abs(a - b) <= 4.0 * np.sqrt(max(a + b, 1)) + 2

    for k in ("split", "channel", "combine", "lost", "back", "detected"):
        assert close(tb[k], tp[k]), (k, tb[k], tp[k])
    for p in range(9):
        for k in ("entered", "returned", "forwarded", "lost", "revisits"):
            assert close(cb[p][k], cp[p][k]), (p, k, cb[p][k], cp[p][k])
    assert sum(p["forwarded"] for p in cb) > 0 and tb["detected"] > 50 and tp["detected"] > 50


@pytest.mark.parametrize("path", [mqhip.PATH_BATCHED, mqhip.PATH_PLAIN])
def test_job_alone_equals_job_in_batch(small, path):
    t, (cube, port, geom, d, jobs, ph) = small
    say(f"[gpu] tandem batch independence, path {path}")
    c, tt, s = t.run_ions(jobs, d, ph, path)
    c1, t1, s1 = t.run_ions(jobs[2:3], d, ph, path)
    sel = s["job"] == 2
    for k in ("x", "y", "z", "vx", "vy", "vz", "steps", "ncoll", "status", "region", "last"):
        assert np.array_equal(s[k][sel], s1[k]), k


def test_tandem_rejects_bad_inputs(small):
    t, (cube, port, geom, d, jobs, ph) = small
    with pytest.raises(mqhip.MqError, match="legacy"):
        t.run(jobs, d, mqhip.Physics.legacy(0.1))
    with pytest.raises(mqhip.MqError, match="dc_scale"):
        t.run([mqhip.job(8565.0, 7.0, 1000.0, 115.0, 500e3, 0.5, 0.0, 10, 1, 0.5)], d, ph)
    with pytest.raises(mqhip.MqError, match="at most"):
        t.run(jobs * 3, d, ph)
    mouths = [(geom.mouth[k][0], geom.mouth[k][1]) for k in range(9)]
    ent = (geom.entrance[0], geom.entrance[1])
    bad = mqhip.tandem_geom(30.0e-3, mouths, geom.ap_r, ent, geom.l_ext)
    with pytest.raises(mqhip.MqError, match="seam"):
        mqhip.Tandem(cube, port, bad)
    bad = mqhip.tandem_geom(geom.seam_z, mouths, geom.ap_r, (geom.entrance[0], 0.1e-3), geom.l_ext)
    with pytest.raises(mqhip.MqError, match="entrance"):
        mqhip.Tandem(cube, port, bad)
    bad = mqhip.tandem_geom(geom.seam_z, mouths, geom.ap_r, ent, 0.0)
    with pytest.raises(mqhip.MqError, match="l_ext"):
        mqhip.Tandem(cube, port, bad)


@pytest.mark.timeout(900)
def test_full_size_plain_matches_oracle_on_a_subset(oracle_mod):
    """The full-size tandem grids (cube and port 128 x 128 x 512 at 0.05 mm): 4 x 64 ions of the 1M batch's
    jobs on the FP64 path, every ion's final state and every counter equal to the oracle's."""
    cube, port, geom, d, jobs, ph = synth.tandem_setup(mqhip, "c5", n_per_charge=64)
    t = mqhip.Tandem(cube, port, geom)
    try:
        say("[gpu] tandem full-size grids, FP64 plain vs oracle, 4 x 64 ions")
        c, tt, s = t.run_ions(jobs, d, ph, mqhip.PATH_PLAIN)
    finally:
        t.close()
    co, to, so = oracle_mod.tandem_run(cube, port, geom, jobs, d, ph, 16)
    assert tt == to and c == co
    for k in ("status", "region", "last", "steps", "ncoll", "job", "x", "y", "z", "vx", "vy", "vz"):
        assert np.array_equal(s[k], so[k]), k
    flow_checks(c, tt, s, len(jobs))
    say(f"[gpu]   totals {tt}")
