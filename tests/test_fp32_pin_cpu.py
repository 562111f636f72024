"""CPU: the FP32 per-ion bounds of tests/test_gpu_fp32_pin.py are tight enough to catch a wrong coefficient.

A float32 restatement of the batched path's cell record, polynomial and collision-free step (tests/fp32_pin.py)
stays within both bounds against the FP64 oracle; the same restatement with one coefficient of the record
computed wrongly (the tx ty tz term c7, or the tx ty term c4) breaks the field bound at most points and the
trajectory bound for most ions. The mistake is planted in this CPU restatement only.
"""
import numpy as np
import pytest

import mqhip
from fp32_pin import (drive_response, emulated_field, emulated_run, field_bound, smooth_grid, smooth_jobs,
                      state_errors)

AMU, QE = 1.66053906660e-27, 1.602176634e-19


@pytest.fixture(scope="module")
def grid():
    return smooth_grid(48, 48, 80, seed=7)


def oracle_field(oracle_mod, g, dc, x, y, z, amps):
    from fp32_pin import combined_dc
    dcc = combined_dc(g, dc)
    out = np.zeros((len(x), 3))
    for F, a in ((g["rf"], amps[0]), (dcc, amps[1]), (g["ac"], amps[2])):
        for i in range(len(x)):
            out[i] += a * oracle_mod.sample(F, g["nx"], g["ny"], g["nz"], (x[i] / g["h"], y[i] / g["h"],
                                                                            z[i] / g["h"]))
    return out


@pytest.mark.parametrize("wrong", [None, 7, 4])
def test_field_bound_catches_a_wrong_coefficient(grid, oracle_mod, wrong):
    g = grid
    rng = np.random.default_rng(3)
    L = np.array([g["nx"] - 1, g["ny"] - 1, g["nz"] - 1]) * g["h"]
    p = rng.uniform(0, 1, (1500, 3)) * L
    dc, amps = [0.0, 1.0, 0.0, 0.0], (37.5, -0.8, 0.6)
    got = emulated_field(g, dc, p[:, 0], p[:, 1], p[:, 2], *amps, wrong=wrong)
    want = oracle_field(oracle_mod, g, dc, p[:, 0], p[:, 1], p[:, 2], amps)
    bound = field_bound(g, dc, p[:, 0], p[:, 1], p[:, 2], *amps)
    inside = np.max(np.abs(got - want), axis=1) <= bound
    if wrong is None:
        assert inside.all()
    else:
        assert inside.mean() < 0.2


def _start(g, jobs, ph):
    import oracle
    n = sum(j.n for j in jobs)
    st = {k: np.empty(n) for k in ("x", "y", "z", "vx", "vy", "vz", "phi_ac0")}
    jp = {k: np.empty(n) for k in ("kq", "v_rf", "w_rf", "v_ac", "w_ac", "s_dc")}
    i = 0
    for j in jobs:
        for k in range(j.n):
            s = oracle.inject(g, j, ph, k)
            st["x"][i], st["y"][i], st["z"][i] = s["x"]
            st["vx"][i], st["vy"][i], st["vz"][i] = s["v"]
            st["phi_ac0"][i] = s["phi_ac"]
            m = j.sp.mass * AMU
            jp["kq"][i] = (0.5 if ph.legacy_force else 1.0) * j.sp.z * QE / m * ph.dt
            jp["v_rf"][i], jp["w_rf"][i] = j.drive.v_rf, j.drive.f_rf * ph.dt
            jp["v_ac"][i], jp["w_ac"][i] = j.drive.v_ac, j.drive.f_ac * ph.dt
            jp["s_dc"][i] = j.dc_scale
            i += 1
    jp["dt"] = ph.dt
    return st, jp


@pytest.mark.parametrize("half", [0, 1])
@pytest.mark.parametrize("wrong", [None, 7])
def test_trajectory_bound_catches_a_wrong_coefficient(oracle_mod, half, wrong):
    g = smooth_grid()
    jobs, ph = smooth_jobs(g, 150, half)
    dc = [0.0, 0.2, 0.0, 0.0]
    want = oracle_mod.run_batch(g, jobs, dc, ph, 8)
    assert np.all(want["status"] == 5) and np.all(want["steps"] == 256)
    st, jp = _start(g, jobs, ph)
    got = emulated_run(g, dc, st, jp, 256, wrong=wrong)
    rx, rv = drive_response(oracle_mod, g, jobs, dc, ph, want, 8)
    ex, ev = state_errors(got, want)
    qx, qv = ex / (4 * rx + 1e-15), ev / (4 * rv + 1e-9)
    if wrong is None:
        assert np.all(qx <= 1.0) and np.all(qv <= 1.0)
        assert np.median(qx) <= 0.25 and np.median(qv) <= 0.25
    else:
        assert np.mean(qx > 1.0) > 0.5
