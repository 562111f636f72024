"""GPU parity: the HIP paths through the C ABI against the CPU FP64 oracle.

* FP64 plain path: per-ion outcomes, step counts and collision counts
  identical, final positions and velocities within 1e-10 relative (the code
  follows the oracle's operation order, so they are expected bit-identical).
* FP32 batched path: outcome counts per class within a 3-sigma binomial bound
  of the oracle on the same seeded inputs; per-ion results independent of the
  batch composition and of the launch tuning; deterministic.
"""
import os
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
    say("[gpu] creating engine on the 64x64x256 grid")
    e = mqhip.Engine(g)
    yield g, e
    e.close()


def physics(case):
    base = mqhip.Physics.correct(0.133322).copy(dt=1e-7, max_steps=20000, injection=mqhip.INJ_GAUSSIAN,
                                                inj_sigma_mm=0.1)
    if case == "full_thermal":
        return base
    if case == "half_thermal":
        return base.copy(legacy_force=1)
    if case == "half_thermal_nearest":
        return base.copy(legacy_force=1, legacy_fields=1)
    if case == "thermal_disc":
        return mqhip.Physics.correct(0.133322).copy(max_steps=20000, legacy_force=1)
    if case == "legacy":
        return mqhip.Physics.legacy(1e-3 * 133.322).copy(max_steps=20000)
    if case == "legacy_trilinear":
        return mqhip.Physics.legacy(1e-3 * 133.322).copy(max_steps=20000, legacy_fields=0)
    raise KeyError(case)


def jobs_for(case, n):
    if case.startswith("legacy"):
        return [mqhip.job(5 * 499.0, 5.0, 0.0, 100.0, 500e3, 0.5, 100e3, n, 1499),
                mqhip.job(5 * 503.0, 5.0, 0.0, 100.0, 500e3, 0.5, 100e3, n // 2, 1503)]
    return [mqhip.job(499.0, 1.0, 230.0, 100.0, 500e3, 0.5, 100e3, n, 1499),
            mqhip.job(505.0, 1.0, 235.0, 100.0, 480e3, 0.8, 95e3, n // 2, 77, 0.5)]


CASES = ["full_thermal", "half_thermal", "half_thermal_nearest", "thermal_disc", "legacy", "legacy_trilinear"]


@pytest.fixture(scope="module")
def c1_push():
    """The C1 grid with a uniform axial push added to its DC basis, so that every trajectory depends on the DC
    weight (the gate alone acts only near z = 16.8 mm)."""
    g = synth.c1_grid()
    push = np.zeros((g["nx"] * g["ny"] * g["nz"], 3))
    push[:, 2] = 2.0
    g["dc"] = [None, g["dc"][1] + push.reshape(-1), None, None]
    say("[gpu] creating engine on the 64x64x256 grid with an axial DC push")
    e = mqhip.Engine(g)
    yield g, e
    e.close()


@pytest.mark.parametrize("case", CASES + ["compat"])
def test_plain_cell_records_keep_results(c1_push, case):
    """The plain path reads the field from per-cell FP64 records of the corner node values, kept per lane while
    its ion stays in a cell (mq_tune.plain_recs = 1, the default), or gathers the node tables at every step (0).
    Per-ion results are the same bit for bit, for every collision model and for the reference's own streams, and
    after a change of the DC set (the records are rebuilt for the new combined DC)."""
    g, e = c1_push
    ph = physics("half_thermal").copy(rng=mqhip.RNG_COMPAT) if case == "compat" else physics(case)
    jobs = jobs_for("half_thermal" if case == "compat" else case, 2000)
    t0 = e.get_tune()
    runs = {}
    try:
        for dc in ([0, 5, 0, 0], [0, 3.5, 0, 0]):
            for recs in (1, 0):
                say(f"[gpu] plain path, {case}, dc {dc}: plain_recs={recs}")
                e.set_tune(plain_recs=recs)
                runs[(tuple(dc), recs)] = e.run_batch_ions(jobs, dc, ph, mqhip.PATH_PLAIN)
                assert e.stats()["plain_recs"] == recs
    finally:
        e.set_tune(**t0)
    for dc in ([0, 5, 0, 0], [0, 3.5, 0, 0]):
        a, b = runs[(tuple(dc), 1)], runs[(tuple(dc), 0)]
        for k in KEYS + ("steps", "ncoll", "status"):
            assert np.array_equal(a[k], b[k]), (dc, k)
    assert not np.array_equal(runs[((0, 5, 0, 0), 1)]["z"], runs[((0, 3.5, 0, 0), 1)]["z"])


@pytest.mark.parametrize("case", CASES)
def test_plain_fp64_matches_oracle(c1, oracle_mod, case):
    g, e = c1
    ph = physics(case)
    jobs = jobs_for(case, 200)
    say(f"[gpu] plain FP64 vs oracle: {case}, {sum(j.n for j in jobs)} ions")
    got = e.run_batch_ions(jobs, [0, 5, 0, 0], ph, mqhip.PATH_PLAIN)
    want = oracle_mod.run_batch(g, jobs, [0, 5, 0, 0], ph, 8)
    assert np.array_equal(got["job"], want["job"])
    assert np.array_equal(got["status"], want["status"])
    assert np.array_equal(got["steps"], want["steps"])
    assert np.array_equal(got["ncoll"], want["ncoll"])
    worst = 0.0
    for k in KEYS:
        d = np.abs(got[k] - want[k]) / np.maximum(np.abs(want[k]), 1e-300)
        worst = max(worst, float(d.max()))
    say(f"[gpu]   max relative difference {worst:.3e}; bit-identical: "
        f"{all(np.array_equal(got[k], want[k]) for k in KEYS)}")
    assert worst <= 1e-10
    assert all(np.array_equal(got[k], want[k]) for k in KEYS)


@pytest.mark.parametrize("case", ["full_thermal", "half_thermal", "half_thermal_nearest", "legacy"])
def test_batched_fp32_counts_within_3sigma(c1, oracle_mod, case):
    g, e = c1
    ph = physics(case)
    jobs = jobs_for(case, 3000)
    say(f"[gpu] batched FP32 vs oracle: {case}, {sum(j.n for j in jobs)} ions")
    got = e.run_batch_ions(jobs, [0, 5, 0, 0], ph, mqhip.PATH_BATCHED)
    want = oracle_mod.run_batch(g, jobs, [0, 5, 0, 0], ph, 8)
    same = float(np.mean(got["status"] == want["status"]))
    for j in range(len(jobs)):
        sg, sw = got["status"][got["job"] == j], want["status"][want["job"] == j]
        n = len(sw)
        assert len(sg) == n
        for k in range(1, 6):
            pull = binom_pull(int((sg == k).sum()), n, int((sw == k).sum()), n)
            say(f"[gpu]   job {j} status {k}: gpu {(sg == k).sum()} oracle {(sw == k).sum()} pull {pull:+.2f}")
            assert abs(pull) <= 3.0
    say(f"[gpu]   per-ion outcome identical to the FP64 oracle for {same:.4f} of ions")


def test_batched_independent_of_batch_and_tuning(c1):
    """Each ion's FP32 computation depends only on its own state and counters, so it
    is bit-identical alone or in a batch, and under any launch shape (check.rs:22-60)."""
    g, e = c1
    ph = physics("half_thermal")
    jobs = jobs_for("half_thermal", 500)
    say("[gpu] batched path: batch-composition and tuning independence")
    full = e.run_batch_ions(jobs, [0, 5, 0, 0], ph)
    alone = e.run_batch_ions(jobs[1:], [0, 5, 0, 0], ph)
    sel = full["job"] == 1
    for k in KEYS + ("steps", "ncoll", "status"):
        assert np.array_equal(full[k][sel], alone[k]), k
    t0 = e.get_tune()
    variants = [
        dict(block=64, waves_per_cu=3, refill_period=4, chunk=100, chunk_cap=400, resort=0, sync_every=1, graph=0),
        dict(block=128, waves_per_cu=8, refill_period=8, chunk=64, chunk_cap=64, resort=1, sync_every=3, graph=0),
        dict(block=256, waves_per_cu=4, refill_period=16, chunk=96, chunk_cap=96, resort=1, sync_every=4, graph=1),
        dict(block=256, waves_per_cu=8, refill_period=8, chunk=20000, chunk_cap=200000, resort=0, sync_every=2,
             graph=1),
        dict(block=256, waves_per_cu=8, refill_period=16, chunk=256, chunk_cap=256, resort=1, sync_every=16, graph=0,
             coll_batch=6),
        dict(block=64, waves_per_cu=4, refill_period=4, chunk=128, chunk_cap=128, resort=1, sync_every=8, graph=0,
             coll_batch=64),
        # staged tiles: workgroup chunks of the sorted list with their records in LDS
        dict(block=256, waves_per_cu=8, refill_period=8, chunk=64, chunk_cap=64, resort=2, sync_every=4, graph=0,
             coll_batch=4),
        dict(block=256, waves_per_cu=8, refill_period=8, chunk=40, chunk_cap=40, resort=2, sync_every=4, graph=1,
             coll_batch=1),
    ]
    runs = []
    try:
        for v in variants:
            e.set_tune(**v)
            runs.append(e.run_batch_ions(jobs, [0, 5, 0, 0], ph))
            say(f"[gpu]   tune {v}: {e.stats()['launches']} launches")
        runs.append(e.run_batch_ions(jobs, [0, 5, 0, 0], ph))  # graph mode again, same engine
    finally:
        e.set_tune(**t0)
    for other in runs:
        for k in KEYS + ("steps", "ncoll", "status"):
            assert np.array_equal(full[k], other[k]), k
    again = e.run_batch_ions(jobs, [0, 5, 0, 0], ph)
    for k in KEYS + ("steps", "ncoll", "status"):
        assert np.array_equal(full[k], again[k]), k


def test_batched_collision_batching_is_exact(c1):
    """Deferring collisions to a wavefront-wide batch leaves every ion's result unchanged
    (thermal and legacy collision models)."""
    g, e = c1
    t0 = e.get_tune()
    try:
        for case in ("half_thermal", "legacy"):
            ph = physics(case)
            jobs = jobs_for(case, 600)
            e.set_tune(coll_batch=1, resort=1)
            a = e.run_batch_ions(jobs, [0, 5, 0, 0], ph)
            for cb, rs in ((4, 1), (16, 1), (4, 2), (16, 2)):
                e.set_tune(coll_batch=cb, resort=rs, graph=t0["graph"], tail_lanes=t0["tail_lanes"])
                b = e.run_batch_ions(jobs, [0, 5, 0, 0], ph)
                for k in KEYS + ("steps", "ncoll", "status"):
                    assert np.array_equal(a[k], b[k]), (case, cb, rs, k)
    finally:
        e.set_tune(**t0)


def test_batched_tail_switch_keeps_results(c1):
    """A batch larger than a launch's lanes: re-sorted launches with one queue per XCD until the
    live count fits the lanes, then launches without re-sort and with doubling chunks. Per-ion
    results equal those of launches that never re-sort."""
    g, e = c1
    ph = physics("half_thermal")
    jobs = jobs_for("half_thermal", 80000)   # 120,000 ions
    t0 = e.get_tune()
    say("[gpu] batched path: re-sorted launches, then the tail switch, vs no re-sort")
    try:
        # 256 CUs x 3 waves x 64 lanes = 49,152 lanes: the first launches carry more ions than lanes
        e.set_tune(block=64, waves_per_cu=3, refill_period=8, chunk=64, chunk_cap=64, resort=1, sync_every=1, graph=0)
        sorted_run = e.run_batch_ions(jobs, [0, 5, 0, 0], ph)
        n_sorted = e.stats()["launches"]
        # the tail spread over every wavefront (tail_lanes 0) and at 3 ions per wavefront
        spread = []
        for tl in (0, 3):
            e.set_tune(block=64, waves_per_cu=3, refill_period=8, chunk=64, chunk_cap=64, resort=1, sync_every=1,
                       graph=0, tail_lanes=tl)
            spread.append(e.run_batch_ions(jobs, [0, 5, 0, 0], ph))
        e.set_tune(block=256, waves_per_cu=8, refill_period=8, chunk=20000, chunk_cap=200000, resort=0,
                   sync_every=2, graph=0, tail_lanes=64)
        plain_launches = e.run_batch_ions(jobs, [0, 5, 0, 0], ph)
    finally:
        e.set_tune(**t0)
    say(f"[gpu]   {n_sorted} launches with re-sort and tail")
    for k in KEYS + ("steps", "ncoll", "status"):
        assert np.array_equal(sorted_run[k], plain_launches[k]), k
        for r in spread:
            assert np.array_equal(sorted_run[k], r[k]), k
    assert sorted_run["steps"].max() > 64 * 8   # the run outlasted several re-sort intervals


@pytest.mark.parametrize("case, n", [("half_thermal", 40000), ("legacy", 15000), ("half_thermal_nearest", 40000),
                                     ("full_thermal", 40000)])
def test_batched_quad_tail_keeps_results(c1, case, n):
    """The four-lane tail worker (k_tail4: each ion's step split over a quad of lanes, packed 16 ions per
    wavefront, 1024-step launches) takes over once at most 16 ions per wavefront are left. Per-ion results
    equal, bit for bit, those of the one-lane worker and of launches that never re-sort, for the thermal
    and the legacy collision models."""
    g, e = c1
    ph = physics(case)
    jobs = jobs_for(case, n)   # 1.5 n ions (the legacy ions end within a narrower span of steps)
    t0 = e.get_tune()
    say(f"[gpu] batched path, {case}: four-lane tail vs one lane")
    # 256 CUs x 3 waves x 64 lanes = 49,152 lanes; the four-lane tail holds 16 x 768 = 12,288 ions
    shape = dict(block=64, waves_per_cu=3, refill_period=8, chunk=64, chunk_cap=64, resort=1, sync_every=1,
                 graph=0, tail_lanes=0)
    try:
        e.set_tune(tail_quads=1, **shape)
        quad = e.run_batch_ions(jobs, [0, 5, 0, 0], ph)
        st_quad = e.stats()
        e.set_tune(tail_quads=0, **shape)
        one = e.run_batch_ions(jobs, [0, 5, 0, 0], ph)
        st_one = e.stats()
        e.set_tune(block=256, waves_per_cu=8, refill_period=8, chunk=20000, chunk_cap=200000, resort=0,
                   sync_every=2, graph=0, tail_lanes=64, tail_quads=0)
        plain_launches = e.run_batch_ions(jobs, [0, 5, 0, 0], ph)
    finally:
        e.set_tune(**t0)
    say(f"[gpu]   {st_quad['launches']} launches, {st_quad['quad_launches']} of them four-lane")
    assert st_quad["quad_launches"] > 0 and st_one["quad_launches"] == 0
    assert st_quad["ion_steps"] == st_one["ion_steps"]
    for k in KEYS + ("steps", "ncoll", "status"):
        assert np.array_equal(quad[k], one[k]), k
        assert np.array_equal(quad[k], plain_launches[k]), k
    if case == "legacy":
        assert quad["ncoll"].sum() > 0


def test_quad_tail_not_used_with_reference_streams(c1):
    """The reference's per-ion streams draw sequentially, which the four-lane worker's split of the Philox blocks
    does not apply to: with MQ_RNG_COMPAT the tail stays on one lane per ion, whatever tail_quads says."""
    g, e = c1
    ph = physics("half_thermal").copy(rng=mqhip.RNG_COMPAT)
    jobs = jobs_for("half_thermal", 20000)
    t0 = e.get_tune()
    shape = dict(block=64, waves_per_cu=3, refill_period=8, chunk=64, chunk_cap=64, resort=1, sync_every=1,
                 graph=0, tail_lanes=0)
    try:
        e.set_tune(tail_quads=1, **shape)
        a = e.run_batch_ions(jobs, [0, 5, 0, 0], ph)
        st = e.stats()
        e.set_tune(tail_quads=0, **shape)
        b = e.run_batch_ions(jobs, [0, 5, 0, 0], ph)
    finally:
        e.set_tune(**t0)
    assert st["quad_launches"] == 0 and st["launches"] >= 4   # re-sorted launches, then the tail
    for k in KEYS + ("steps", "ncoll", "status"):
        assert np.array_equal(a[k], b[k]), k


def test_plain_independent_of_batch(c1):
    g, e = c1
    ph = physics("legacy")
    jobs = jobs_for("legacy", 200)
    say("[gpu] plain path: batch-composition independence")
    full = e.run_batch_ions(jobs, [0, 5, 0, 0], ph, mqhip.PATH_PLAIN)
    alone = e.run_batch_ions(jobs[1:], [0, 5, 0, 0], ph, mqhip.PATH_PLAIN)
    sel = full["job"] == 1
    for k in KEYS + ("steps", "ncoll", "status"):
        assert np.array_equal(full[k][sel], alone[k]), k


@pytest.mark.parametrize("path", [mqhip.PATH_BATCHED, mqhip.PATH_PLAIN])
def test_outcomes_and_stats_consistent(c1, path):
    g, e = c1
    ph = physics("half_thermal")
    jobs = jobs_for("half_thermal", 400)
    say(f"[gpu] run_batch vs summarize(run_batch_ions), path {path}")
    outs = e.run_batch(jobs, [0, 5, 0, 0], ph, path)
    st = e.stats()
    ions = e.run_batch_ions(jobs, [0, 5, 0, 0], ph, path)
    ref = mqhip.summarize(ions["status"], ions["ncoll"], ions["steps"], ions["job"], len(jobs), ph.dt)
    for a, b in zip(outs, ref):
        assert a.as_dict() == b.as_dict()
    assert st["ion_steps"] == int(ions["steps"].sum())
    assert st["n_ions"] == len(ions["status"])
    assert st["kernel_ms"] > 0 and st["launches"] >= 1


@pytest.mark.parametrize("path", [mqhip.PATH_BATCHED, mqhip.PATH_PLAIN])
def test_edge_cases(c1, oracle_mod, path):
    g, e = c1
    ph = physics("half_thermal")
    say(f"[gpu] edge cases, path {path}")
    # an empty job between two others, and a batch with no ions at all
    jobs = [mqhip.job(499.0, 1.0, 230.0, 100.0, 500e3, 0.5, 100e3, 3, 1), mqhip.job(499.0, 1.0, 230.0, 100.0,
            500e3, 0.5, 100e3, 0, 2), mqhip.job(499.0, 1.0, 230.0, 100.0, 500e3, 0.5, 100e3, 1, 3)]
    outs = e.run_batch(jobs, [0, 5, 0, 0], ph, path)
    assert [o.n for o in outs] == [3, 0, 1]
    assert outs[1].mean_time == 0.0
    outs = e.run_batch(jobs[1:2], [0, 5, 0, 0], ph, path)
    assert outs[0].n == 0 and e.stats()["ion_steps"] == 0
    # a one-step cap: every ion ends after exactly one step
    r = e.run_batch_ions(jobs, [0, 5, 0, 0], ph.copy(max_steps=1), path)
    assert np.all(r["steps"] == 1) and np.all(r["status"] != 0)
    # ions injected outside the grid leave the box on their first step
    g2 = dict(g)
    g2["axis_x"] = -1e-3
    e2 = mqhip.Engine(g2)
    try:
        r = e2.run_batch_ions(jobs, [0, 5, 0, 0], ph, path)
        assert np.all(r["status"] == mqhip.LEFT) and np.all(r["steps"] == 1)
        w = oracle_mod.run_batch(g2, jobs, [0, 5, 0, 0], ph, 1)
        assert np.array_equal(w["status"], r["status"])
    finally:
        e2.close()


def test_rejects_bad_arguments(c1):
    g, e = c1
    ph = physics("half_thermal")
    with pytest.raises(mqhip.MqError, match="no jobs"):
        e.run_batch([], [0, 0, 0, 0], ph)
    with pytest.raises(mqhip.MqError, match="path"):
        e.run_batch(jobs_for("half_thermal", 2), [0, 0, 0, 0], ph, 7)
    with pytest.raises(mqhip.MqError, match="dt"):
        e.run_batch(jobs_for("half_thermal", 2), [0, 0, 0, 0], ph.copy(dt=0.0))
    with pytest.raises(mqhip.MqError, match="block"):
        e.set_tune(block=100)
    with pytest.raises(mqhip.MqError, match="block"):
        e.set_tune(block=512)
    with pytest.raises(mqhip.MqError, match="refill"):
        e.set_tune(refill_period=2)


def test_autotune_persists_and_keeps_results(c1, tmp_path):
    g, e = c1
    ph = physics("half_thermal")
    jobs = jobs_for("half_thermal", 2000)
    say("[gpu] autotune")
    ref = e.run_batch_ions(jobs, [0, 5, 0, 0], ph)
    t0 = e.get_tune()
    cache = str(tmp_path / "tune.tsv")
    try:
        best, rate = e.autotune(jobs, [0, 5, 0, 0], ph, sample_ions=1500, cache_path=cache)
        assert rate > 0 and e.get_tune() == best
        assert len(open(cache).read().splitlines()) == 1
        again, rate2 = e.autotune(jobs, [0, 5, 0, 0], ph, cache_path=cache)
        assert again == best and rate2 == 0.0
        tuned = e.run_batch_ions(jobs, [0, 5, 0, 0], ph)
        # an entry tuned elsewhere with a smaller block, another refill (compaction) period and graph replay:
        # the last entry for a key wins, and the results stay the same
        line = open(cache).read().splitlines()[0]
        key = line.split("\t")[0]
        with open(cache, "a") as fh:
            fh.write(f"{key}\t128,4,16,128,128,1,8,1,2\t0\n")
        other, rate3 = e.autotune(jobs, [0, 5, 0, 0], ph, cache_path=cache)
        assert [other[k] for k in ("block", "refill_period", "waves_per_cu", "graph", "coll_batch")] == [128, 16, 4, 1, 2]
        assert e.get_tune() == other and rate3 == 0.0
        tuned2 = e.run_batch_ions(jobs, [0, 5, 0, 0], ph)
    finally:
        e.set_tune(**t0)
    say(f"[gpu]   autotuned: {best}")
    for k in KEYS + ("steps", "ncoll", "status"):
        assert np.array_equal(ref[k], tuned[k]), k
        assert np.array_equal(ref[k], tuned2[k]), k
