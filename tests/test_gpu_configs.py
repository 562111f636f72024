"""GPU parity on the mass-scan and operating-setting-sweep workloads, each on its own grid and job structure.

Mass-scan slice: m/z 475..523 step 2 (25) x 4 of the 64 RF frequencies (k = 0, 21, 42, 63: 450, 483.3, 516.7,
550 kHz) on the 128 x 128 x 512 grid at 0.05 mm (FP32 cell table of 2.6 GB); Fig-4 drive (100 V RF, 0.5 V AC at
100 kHz), half kick, thermal N2 at 1 mTorr and 300 K, Gaussian 0.1 mm spot with a 300 K Maxwellian start,
dt 100 ns. With the 5 V gate of the workload no ion crosses the gate (an ion at 300 K has about 0.04 eV), so the
slice also runs a 0.02 V gate, where about a quarter of the ions are transmitted and that class is compared too.

Sweep slice: 32 Latin-hypercube settings of the sweep (V_rf 50-150 V, f_exc 80-120 kHz, V_exc 0.1-1 V, V_dc
0-10 V as the job's dc_scale on the plate basis, N2 0.5-2 mTorr as the job's pressure) at m/z 499 on the
192 x 192 x 768 grid at 1/30 mm (FP32 cell table of 8.9 GB).

Bars:
* FP64 plain path vs the oracle: every ion's status, steps, collisions, position and velocity identical (stricter
  than agreement to 1e-10 relative), on a 2,000-step cap.
* FP32 batched path vs the oracle at the full 20,000-step cap: the totals of each class within 3 sigma (binomial);
  per job and class, at least 99 % of the cells within 3 sigma and every cell within the Bonferroni bound of all
  the cells compared (overall false-alarm probability 0.001; z = 4.75 for 500 cells), since hundreds of cells
  at 3 sigma each would fail by chance.
"""
import os
import sys

import numpy as np
import pytest
from scipy.stats import norm

import mqhip
import synth
from gridutil import binom_pull

pytestmark = pytest.mark.gpu

KEYS = ("x", "y", "z", "vx", "vy", "vz")
CLASSES = {1: "crashed", 2: "returned", 3: "transmitted", 4: "left", 5: "timeout"}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def threads():
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return 8


def physics(max_steps=20000):
    return mqhip.Physics.correct(0.133322).copy(dt=1e-7, max_steps=max_steps, injection=mqhip.INJ_GAUSSIAN,
                                                inj_sigma_mm=0.1, legacy_force=1)


def mass_scan_jobs(n):
    """The mass-scan jobs of synth.c2_jobs at RF-frequency indices 0, 21, 42, 63, n ions each."""
    jobs = synth.c2_jobs(mqhip, n=n)
    return [j for i, j in enumerate(jobs) if i % 64 in (0, 21, 42, 63)]


def sweep_jobs(n):
    """32 settings of the sweep's Latin hypercube (every 128th of its 4096), n ions each."""
    jobs = synth.c3_jobs(mqhip, n_settings=4096, n=n)
    return jobs[::128]


def assert_bit_identical(got, want):
    for k in ("job", "status", "steps", "ncoll"):
        assert np.array_equal(got[k], want[k]), k
    worst = 0.0
    for k in KEYS:
        d = np.abs(got[k] - want[k]) / np.maximum(np.abs(want[k]), 1e-300)
        worst = max(worst, float(d.max()))
    say(f"[gpu]   max relative difference {worst:.3e}")
    assert worst <= 1e-10
    for k in KEYS:
        assert np.array_equal(got[k], want[k]), k


def assert_counts_agree(got, want, njobs, need_transmitted=False):
    """Class totals within 3 sigma; per job and class, >= 99 % of cells within 3 sigma and all within the
    Bonferroni bound (overall p = 0.001)."""
    n = len(want["status"])
    assert len(got["status"]) == n and np.array_equal(got["job"], want["job"])
    for k, name in CLASSES.items():
        a, b = int((got["status"] == k).sum()), int((want["status"] == k).sum())
        p = binom_pull(a, n, b, n) if a or b else 0.0
        say(f"[gpu]   {name}: FP32 {a} oracle {b} pull {p:+.2f}")
        assert abs(p) <= 3.0, name
    if need_transmitted:
        assert (want["status"] == 3).mean() > 0.1
    pulls = []
    for j in range(njobs):
        sg, sw = got["status"][got["job"] == j], want["status"][want["job"] == j]
        m = len(sw)
        for k in CLASSES:
            a, b = int((sg == k).sum()), int((sw == k).sum())
            pulls.append(binom_pull(a, m, b, m) if a or b else 0.0)
    pulls = np.abs(np.array(pulls))
    zb = float(norm.isf(0.001 / (2 * len(pulls))))
    say(f"[gpu]   {len(pulls)} job x class cells: {np.mean(pulls <= 3):.4f} within 3 sigma, max |pull| "
        f"{pulls.max():.2f} (Bonferroni bound {zb:.2f})")
    assert np.mean(pulls <= 3.0) >= 0.99
    assert pulls.max() <= zb


# ------------------------------------------------------------------ mass scan


@pytest.fixture(scope="module")
def c2():
    say("[gpu] mass-scan grid 128x128x512 @ 0.05 mm")
    g = synth.c2_grid()
    e = mqhip.Engine(g)
    yield g, e
    e.close()


@pytest.mark.timeout(900)
@pytest.mark.parametrize("gate", [5.0, 0.02])
def test_mass_scan_plain_fp64_bit_identical(c2, oracle_mod, gate):
    g, e = c2
    jobs = mass_scan_jobs(40)
    ph = physics(2000)
    say(f"[gpu] mass scan, plain FP64 vs oracle, gate {gate} V: {len(jobs)} jobs, {sum(j.n for j in jobs)} ions, "
        "2000-step cap")
    got = e.run_batch_ions(jobs, [0, gate, 0, 0], ph, mqhip.PATH_PLAIN)
    want = oracle_mod.run_batch(g, jobs, [0, gate, 0, 0], ph, threads())
    assert_bit_identical(got, want)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("gate", [5.0, 0.02])
def test_mass_scan_batched_fp32_counts(c2, oracle_mod, gate):
    g, e = c2
    jobs = mass_scan_jobs(400)
    ph = physics()
    say(f"[gpu] mass scan, batched FP32 vs oracle, gate {gate} V: {len(jobs)} jobs, {sum(j.n for j in jobs)} ions")
    got = e.run_batch_ions(jobs, [0, gate, 0, 0], ph, mqhip.PATH_BATCHED)
    want = oracle_mod.run_batch(g, jobs, [0, gate, 0, 0], ph, threads())
    assert_counts_agree(got, want, len(jobs), need_transmitted=gate < 1.0)


# ------------------------------------------------------------------ sweep


@pytest.fixture(scope="module")
def c3():
    say("[gpu] sweep grid 192x192x768 @ 1/30 mm")
    g = synth.c3_grid()
    e = mqhip.Engine(g)
    yield g, e
    e.close()


@pytest.mark.timeout(900)
def test_sweep_plain_fp64_bit_identical(c3, oracle_mod):
    g, e = c3
    jobs = sweep_jobs(64)
    assert len({j.pressure_pa for j in jobs}) == len(jobs) and len({j.dc_scale for j in jobs}) == len(jobs)
    ph = physics(2000)
    say(f"[gpu] sweep, plain FP64 vs oracle: {len(jobs)} settings, {sum(j.n for j in jobs)} ions, 2000-step cap")
    got = e.run_batch_ions(jobs, [0, 1, 0, 0], ph, mqhip.PATH_PLAIN)
    want = oracle_mod.run_batch(g, jobs, [0, 1, 0, 0], ph, threads())
    assert_bit_identical(got, want)


@pytest.mark.timeout(900)
def test_sweep_batched_fp32_counts(c3, oracle_mod):
    g, e = c3
    jobs = sweep_jobs(1000)
    ph = physics()
    say(f"[gpu] sweep, batched FP32 vs oracle: {len(jobs)} settings, {sum(j.n for j in jobs)} ions")
    got = e.run_batch_ions(jobs, [0, 1, 0, 0], ph, mqhip.PATH_BATCHED)
    want = oracle_mod.run_batch(g, jobs, [0, 1, 0, 0], ph, threads())
    assert_counts_agree(got, want, len(jobs))
