"""GPU: the Optimize search on the engine (include/mq_optimize.h).

* Grid search (research.rs:133-163) on the corrected S14 channel the field build solves: four AC amplitudes,
  AC-frequency scans of 70-100 kHz at 0.5 kHz with 500 ions per point, RF 70 V at 500 kHz, DC [18, 24, 15, 20],
  N2 0.1 Pa, t_max 3 ms. The chosen row is the one the reference's rule (eligible, then the largest T - 2 sigma,
  restated in tests/test_optimize.py) picks from the returned rows; the confirmation's centre agrees with the
  chosen row's within 3 combined standard errors; its flag is the rule applied to it. Each row's scan point
  carries its counts, the Wilson error and the leakage of the Na adduct (500 ions per point of its own) with the
  reference's ratio statistics; out-of-range requests fail with research_start's message.
* Gradient search: projected gradient ascent of the differentiable objective over f_exc, from inside the peak,
  ends within one grid step of the best point of a grid of forward evaluations of the same objective (2,000
  collision-free steps, where the adjoint agrees with central differences).
"""
import sys

import numpy as np
import pytest

import mqhip
import synth
from test_optimize import best_row, eligible

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.timeout(900)
def test_grid_search_on_the_s14_channel():
    say("[gpu] optimize grid: building the corrected S14 channel fields")
    grid = mqhip.channel_grid(mqhip.Geom.s14(), tol=1e-7)
    e = mqhip.Engine(grid)
    try:
        sp = mqhip.Species(2 * 863.473, 2.0, 425.0)
        rf = mqhip.RfAc(70.0, 500e3, 0.35, 84.25e3)
        ph = mqhip.Physics.correct(0.1).copy(t_max=3e-3)
        min_r = 5.0
        na = mqhip.Species(2 * 863.473 + 21.982, 2.0, 430.0)   # the Na adduct (Na for H): the leakage species
        res = mqhip.optimize_grid(e, sp, rf, [18, 24, 15, 20], ph, [0.35, 0.2625, 0.175, 0.0875], 70.0, 100.0, 0.5,
                                  500, min_r, 0xA1100101, leak=na, leak_n=500)
        for r in res["rows"]:
            fit = r["fit"]
            say(f"[gpu]   AC {r['v_ac']:.4f} V: f0 {fit['f0_khz']:.2f} kHz FWHM {fit['fwhm_khz']:.2f} R "
                f"{r['r_mz']:.1f} +- {r['r_mz_err']:.1f}, T {r['t_best']:.3f} at {r['f_best_khz']:.2f} kHz, "
                f"issue {r['issue']}, refined {r['refined']}")
        assert any(r["fit"]["converged"] for r in res["rows"])
        assert res["best"] == best_row(res["rows"], min_r)
        assert res["best"] >= 0
        b, c = res["rows"][res["best"]], res["confirmation"]
        say(f"[gpu]   best row {res['best']}; confirmation f0 {c['fit']['f0_khz']:.2f} kHz, T {c['t_best']:.3f}, "
            f"confirmed {res['confirmed']}")
        # the confirmation rescans the chosen row's frequencies with three times the base ions (research.rs:155)
        assert c["ions_per_point"] == 3 * 500 and c["v_ac"] == b["v_ac"]
        assert (c["lo_khz"], c["step_khz"]) == (b["lo_khz"], b["step_khz"])
        assert res["confirmed"] == eligible(c, min_r)
        if c["fit"]["converged"]:
            se = np.hypot(b["fit"]["f0_err"], c["fit"]["f0_err"])
            assert abs(b["fit"]["f0_khz"] - c["fit"]["f0_khz"]) <= 3 * se + 1e-9
        for r in res["rows"] + [c]:
            assert 70.0 <= r["lo_khz"] < r["hi_khz"] <= 100.0 + 1e-9
            if not r["fit"]["converged"]:
                continue
            # the scan point's counts and statistics (stats.rs:17-22, 51-69; annotate_leakage, research.rs:88-95)
            assert r["t_n"] == r["ions_per_point"] and r["t_best"] == r["t_k"] / r["t_n"]
            assert r["t_best_err"] == mqhip.frac_sigma(r["t_k"], r["t_n"])
            assert r["leak_n"] == r["ions_per_point"] and 0 <= r["leak_k"] <= r["leak_n"]
            want = mqhip.ratio(r["leak_k"], r["leak_n"], r["t_k"], r["t_n"])
            got = (r["leakage"], r["leakage_err"], r["leakage_ul95"])
            assert all((np.isnan(a) and np.isnan(b)) or a == b for a, b in zip(got, want))
            say(f"[gpu]   AC {r['v_ac']:.4f} V leakage {r['leak_k']}/{r['leak_n']}: {r['leakage']:.4g} "
                f"+- {r['leakage_err']:.2g}, 95 % UL {r['leakage_ul95']:.3g}")
        # research_start's check (research.rs:167)
        for bad in (dict(n=499), dict(n=5001), dict(min_r=4.9), dict(min_r=151.0)):
            kw = dict(n=500, min_r=5.0) | bad
            with pytest.raises(mqhip.MqError, match="Choose resolving power 5-150 and 500-5000"):
                mqhip.optimize_grid(e, sp, rf, [18, 24, 15, 20], ph, [0.35], 70.0, 100.0, 0.5, kw["n"], kw["min_r"], 1)
        with pytest.raises(mqhip.MqError, match="leakage species"):
            mqhip.optimize_grid(e, sp, rf, [18, 24, 15, 20], ph, [0.35], 70.0, 100.0, 0.5, 500, 5.0, 1, leak_n=10)
    finally:
        e.close()


@pytest.mark.timeout(600)
def test_gradient_search_ends_at_the_grid_optimum():
    g = synth.c1_grid()
    e = mqhip.Engine(g)
    try:
        ph = mqhip.Physics.correct(0.0).copy(dt=1e-7, injection=mqhip.INJ_GAUSSIAN, inj_sigma_mm=0.1, legacy_force=1)
        # reward radial excitation: the resonance of the AC drive with the secular motion (w_leak < 0)
        obj = mqhip.Objective(2000, 0, 18.6e-3, 2e-4, 0.5e-3, -1.0)
        dc = [0.0, 1.0, 0.0, 0.0]

        def job(f):
            return mqhip.job(499.0, 1.0, 0.0, 100.0, 500e3, 0.5, f, 16384, 4499, dc_scale=0.5)

        fs = np.arange(90e3, 104e3 + 1, 1e3)
        Js = np.array([e.forward(job(f), dc, ph, obj) for f in fs])
        say("[gpu]   J(f_exc): " + ", ".join(f"{f / 1e3:.0f}:{J:.4f}" for f, J in zip(fs, Js)))
        fbest = fs[int(np.argmax(Js))]
        start = fbest - 2e3
        r = mqhip.optimize_grad(e, job(start), dc, ph, obj, mask=1 << 3, lo=[0, 0, 0, 90e3, 0],
                                hi=[1e3, 1e7, 10, 104e3, 10], step0=[1, 1, 1, 1e3, 1], max_iter=40, tol=1e-3)
        say(f"[gpu]   gradient search from {start / 1e3:.1f} kHz: f_exc {r['theta'][3] / 1e3:.3f} kHz, J {r['J']:.4f} "
            f"after {r['n_eval']} evaluations; grid optimum {fbest / 1e3:.0f} kHz, J {Js.max():.4f}")
        assert abs(r["theta"][3] - fbest) <= 1e3
        assert r["J"] >= e.forward(job(start), dc, ph, obj)
        # only the searched parameter moved
        assert r["theta"][:3] == [100.0, 500e3, 0.5] and r["theta"][4] == 0.5
        h = r["history"]
        assert len(h) == r["n_eval"] and np.all(h[:, 3] >= 90e3) and np.all(h[:, 3] <= 104e3)
        assert abs(h[:, 5].max() - r["J"]) <= 1e-15 * abs(r["J"]) + 1e-300
    finally:
        e.close()
