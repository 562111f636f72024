"""CPU: the peak fit of the Optimize driver (mq_fit_peak) and its selection rule, no GPU.

* An exact Gaussian on a baseline is recovered to rounding.
* On binomial-noise curves the fit equals scipy's curve_fit with the same model and weights (parameters to 1e-6
  relative, standard errors to 1e-4 after the same max(1, chi2_red) scaling).
* Curves with no peak, or too few points, do not converge.
"""
import numpy as np
import pytest
from scipy.optimize import curve_fit

import mqhip

FW = 2.0 * np.sqrt(2.0 * np.log(2.0))


def gauss(f, amp, f0, s, base):
    return base + amp * np.exp(-0.5 * ((f - f0) / s) ** 2)


def test_exact_gaussian():
    f = np.arange(70.0, 100.01, 0.5)
    t = gauss(f, 0.3, 84.25, 1.7, 0.04)
    r = mqhip.fit_peak(f, t, 2000)
    assert r["converged"] == 1 and r["npts"] == len(f)
    assert abs(r["f0_khz"] - 84.25) < 1e-9 and abs(r["fwhm_khz"] - FW * 1.7) < 1e-9
    assert abs(r["amp"] - 0.3) < 1e-12 and abs(r["base"] - 0.04) < 1e-12
    assert abs(r["r_f"] - 84.25 / (FW * 1.7)) < 1e-9


@pytest.mark.parametrize("seed", range(8))
def test_matches_scipy_on_noisy_curves(seed):
    rng = np.random.default_rng(seed)
    f = np.arange(70.0, 100.01, 0.5)
    f0, s, amp, base = rng.uniform(78, 92), rng.uniform(0.8, 3.0), rng.uniform(0.1, 0.5), rng.uniform(0.0, 0.1)
    n = 2000
    t = rng.binomial(n, gauss(f, amp, f0, s, base)) / n
    r = mqhip.fit_peak(f, t, n)
    assert r["converged"] == 1
    tt = np.clip(t, 0.5 / n, 1 - 0.5 / n)
    sig = np.sqrt(tt * (1 - tt) / n)
    p, cov = curve_fit(gauss, f, t, p0=[amp, f0, s, base], sigma=sig, absolute_sigma=True, maxfev=20000,
                       xtol=1e-14, ftol=1e-14, gtol=1e-14)
    chi2 = float(np.sum(((t - gauss(f, *p)) / sig) ** 2)) / (len(f) - 4)
    e = np.sqrt(np.diag(cov) * max(1.0, chi2))
    assert abs(r["f0_khz"] - p[1]) <= 1e-6 * abs(p[1])
    assert abs(r["fwhm_khz"] - FW * abs(p[2])) <= 1e-6 * FW * abs(p[2])
    assert abs(r["amp"] - p[0]) <= 1e-6 * abs(p[0]) and abs(r["base"] - p[3]) <= 1e-5
    assert abs(r["chi2_red"] - chi2) <= 1e-6 * chi2
    assert abs(r["f0_err"] - e[1]) <= 1e-4 * e[1]
    assert abs(r["fwhm_err"] - FW * e[2]) <= 1e-4 * FW * e[2]
    assert abs(r["f0_khz"] - f0) < 5 * r["f0_err"] + 0.05


def test_no_peak_no_fit():
    f = np.arange(70.0, 100.01, 0.5)
    assert mqhip.fit_peak(f, np.full_like(f, 0.2), 1000)["converged"] == 0
    assert mqhip.fit_peak(f[:4], gauss(f[:4], 0.3, 71, 1, 0), 1000)["converged"] == 0


def eligible(row, min_r):
    """research.rs:96-99: no issue, resolving power minus two standard errors at least min_r."""
    return row["issue"] == 0 and row["r_mz"] - 2 * row["r_mz_err"] >= min_r


def best_row(rows, min_r):
    """research.rs:101-107: the eligible row with the largest transmission minus two standard errors."""
    cand = [(r["t_best"] - 2 * r["t_best_err"], i) for i, r in enumerate(rows) if eligible(r, min_r)]
    return max(cand)[1] if cand else -1


# ---- the statistics of the rows (mq_frac_sigma, mq_ratio; mqviz/src/stats.rs), pinned as the reference's own
# tests pin them (stats.rs:170-240): interval coverage by simulation, the Poisson limit's level, the boundaries.

def test_wilson_error_boundaries_and_large_n():
    import math
    assert mqhip.frac_sigma(0, 50) > 0.0 and mqhip.frac_sigma(50, 50) > 0.0
    assert abs(mqhip.frac_sigma(215000, 1000000) - math.sqrt(0.215 * 0.785 / 1e6)) < 1e-6
    assert math.isnan(mqhip.frac_sigma(0, 0))


@pytest.mark.parametrize("p", [0.02, 0.2, 0.5, 0.9])
def test_wilson_error_coverage(p):
    rng = np.random.default_rng(int(p * 1000) + 3)
    n, reps = 400, 4000
    k = rng.binomial(n, p, reps)
    sig = np.array([mqhip.frac_sigma(int(x), n) for x in k])
    c1 = np.mean(np.abs(k / n - p) <= sig)
    c2 = np.mean(np.abs(k / n - p) <= 2 * sig)
    assert abs(c1 - 0.6827) < 0.04 and abs(c2 - 0.9545) < 0.03, (c1, c2)


def test_poisson_upper_limit_is_95_percent():
    from scipy.stats import poisson
    for k in range(0, 30):
        v, e, ul = mqhip.ratio(k, 1000, 1000, 1000)   # denominator exactly 1: ul = poisson_ul95(k) / 1000
        lim = ul * 1000
        assert abs(poisson.cdf(k, lim) - 0.05) < 0.012, (k, lim)
    assert abs(mqhip.ratio(0, 1, 1, 1)[2] - 2.97) < 0.01


def test_ratio_coverage_and_branches():
    import math
    rng = np.random.default_rng(21)
    na, nb, pa, pb, reps = 20000, 8000, 0.02, 0.2, 6000
    ka, kb = rng.binomial(na, pa, reps), rng.binomial(nb, pb, reps)
    hit = np.mean([abs(mqhip.ratio(int(a), na, int(b), nb)[0] - pa / pb) <= mqhip.ratio(int(a), na, int(b), nb)[1]
                   for a, b in zip(ka, kb)])
    assert abs(hit - 0.6827) < 0.05, hit
    # no numerator count: the value 0 with a one-count error; no denominator count: undefined
    v, e, ul = mqhip.ratio(0, 500, 100, 1000)
    assert v == 0.0 and e > 0.0 and ul > e
    assert all(math.isnan(x) for x in mqhip.ratio(3, 500, 0, 1000))
    # the upper limit covers the true ratio at least 95 % of the time, one-sided (the numerator's Poisson limit
    # and the denominator's term add linearly, so it errs on the safe side)
    for pa, na in ((0.002, 20000), (0.0002, 20000)):
        ka, kb = rng.binomial(na, pa, reps), rng.binomial(nb, pb, reps)
        cov = np.mean([pa / pb <= mqhip.ratio(int(a), na, int(b), nb)[2] for a, b in zip(ka, kb)])
        assert 0.94 <= cov <= 0.999, (pa, cov)
