"""Pins of the oracle's dual-number gradient of the differentiable form.

Closed forms: in a uniform DC field, z_N = z0 + N v0 dt + kq s_dc V E0 dt N(N+1)/2;
in a uniform oscillating RF field along z, z_N = z0 + N v0 dt + kq dt sum_n (N - n) a_rf(n) E1,
whose derivatives in V_rf and f_rf are sums computed here directly. Then central
finite differences of the oracle's own objective on the synthetic channel.
"""
import math

import numpy as np
import pytest

import mqhip
import synth
from gridutil import make_grid

AMU, QE = mqhip.AMU, mqhip.QE


def sig(u):
    return 1.0 / (1.0 + math.exp(-u))


def _ph(**kw):
    return mqhip.Physics.correct(0.0).copy(dt=1e-7, injection=mqhip.INJ_GAUSSIAN, inj_sigma_mm=0.0, temp_k=1.0, **kw)


def test_uniform_dc_closed_form(oracle_mod):
    E0, V, s_dc, N = 2.0, 3.0, 0.7, 800
    g = make_grid(8, 8, 64, 1e-3, dc1=lambda X, Y, Z: (0 * X, 0 * X, 0 * X + E0), axis=(3.5e-3, 3.5e-3),
                  z_inject=20e-3)
    ph = _ph()
    job = mqhip.job(499.0, 1.0, 0.0, 0.0, 500e3, 0.0, 100e3, 4, 3, dc_scale=s_dc)
    obj = mqhip.Objective(N, 0, 21e-3, 2e-5, 0.0, 0.0)
    J, grad, gs = oracle_mod.grad_reference(g, job, [0, V, 0, 0], ph, obj, 2)
    kq = QE / (499.0 * AMU) * ph.dt
    dzds = kq * V * E0 * ph.dt * N * (N + 1) / 2
    want_J, want_g, want_z0, want_vz0 = 0.0, 0.0, [], []
    for i in range(4):
        s0 = oracle_mod.inject(g, job, ph, i)
        zN = s0["x"][2] + N * s0["v"][2] * ph.dt + s_dc * dzds
        u = (zN - 21e-3) / 2e-5
        d = sig(u) * (1 - sig(u)) / 2e-5
        want_J += sig(u) / 4
        want_g += d * dzds / 4
        want_z0.append(d / 4)
        want_vz0.append(d * N * ph.dt / 4)
    assert J == pytest.approx(want_J, rel=1e-9)
    assert grad[4] == pytest.approx(want_g, rel=1e-7)
    assert np.allclose(gs[2], want_z0, rtol=1e-7) and np.allclose(gs[5], want_vz0, rtol=1e-7)
    assert np.all(grad[:4] == 0.0)


def test_uniform_rf_frequency_derivative(oracle_mod):
    E1, V, f, N = 1.5, 20.0, 300e3, 500
    g = make_grid(8, 8, 64, 1e-3, rf=lambda X, Y, Z: (0 * X, 0 * X, 0 * X + E1), axis=(3.5e-3, 3.5e-3), z_inject=20e-3)
    ph = _ph()
    job = mqhip.job(499.0, 1.0, 0.0, V, f, 0.0, 100e3, 1, 5)
    obj = mqhip.Objective(N, 0, 20.0005e-3, 1e-5, 0.0, 0.0)
    J, grad, gs = oracle_mod.grad_reference(g, job, [0, 0, 0, 0], ph, obj, 1)
    kq = QE / (499.0 * AMU) * ph.dt
    n = np.arange(N)
    ph_c = (f * ph.dt * n) % 1.0
    dz_dV = kq * ph.dt * E1 * np.sum((N - n) * np.sin(2 * np.pi * ph_c))
    dz_df = kq * ph.dt * E1 * np.sum((N - n) * V * np.cos(2 * np.pi * ph_c) * 2 * np.pi * ph.dt * n)
    s0 = oracle_mod.inject(g, job, ph, 0)
    zN = s0["x"][2] + N * s0["v"][2] * ph.dt + V * dz_dV
    u = (zN - obj.z_exit) / obj.scale
    d = sig(u) * (1 - sig(u)) / obj.scale
    assert J == pytest.approx(sig(u), rel=1e-9)
    assert grad[0] == pytest.approx(d * dz_dV, rel=1e-7)
    assert grad[1] == pytest.approx(d * dz_df, rel=1e-7)


def _c1_setting():
    grid = synth.c1_grid()
    ph = mqhip.Physics.correct(0.0).copy(dt=1e-7, injection=mqhip.INJ_GAUSSIAN, inj_sigma_mm=0.1, legacy_force=1)
    job = mqhip.job(499.0, 1.0, 0.0, 100.0, 500e3, 0.5, 100e3, 48, 1499, dc_scale=0.1)  # 0.5 V gate
    obj = mqhip.Objective(1500, 0, 2.0e-3, 2e-5, 1.0e-3, 0.3)
    return grid, ph, job, obj


def test_dual_numbers_match_central_differences(oracle_mod):
    grid, ph, job, obj = _c1_setting()
    dc = [0, 5, 0, 0]
    J, grad, _ = oracle_mod.grad_reference(grid, job, dc, ph, obj, 8)
    names = [("drive", "v_rf"), ("drive", "f_rf"), ("drive", "v_ac"), ("drive", "f_ac"), (None, "dc_scale")]
    def central(grp, name, h):
        base = getattr(job.drive if grp else job, name)
        vals = []
        for sgn in (+1, -1):
            j2 = mqhip.Job()
            j2.sp, j2.drive, j2.n, j2.seed, j2.dc_scale, j2.pressure_pa = job.sp, job.drive, job.n, job.seed, job.dc_scale, 0.0
            setattr(j2.drive if grp else j2, name, base + sgn * h)
            vals.append(oracle_mod.grad_reference(grid, j2, dc, ph, obj, 8)[0])
        return (vals[0] - vals[1]) / (2 * h)

    for k, (grp, name) in enumerate(names):
        h = 1e-6 * abs(getattr(job.drive if grp else job, name))
        rich = (4 * central(grp, name, h / 2) - central(grp, name, h)) / 3
        assert rich == pytest.approx(grad[k], rel=1e-4, abs=1e-12 * max(1.0, abs(grad).max())), (name, rich, grad[k])


def test_coupled_linear_fields_central_differences(oracle_mod):
    """Ideal quadrupole (E_x depends on y and E_y on x) with a radial leakage term: every
    tangent component must be propagated from the state before the step."""
    r0, ax = 2.765e-3, 3.2e-3
    g = make_grid(32, 32, 64, 0.2e-3, rf=lambda X, Y, Z: (2 * (Y - ax) / r0**2, 2 * (X - ax) / r0**2, 0 * X),
                  ac=lambda X, Y, Z: (0 * X - 1 / r0, 0 * X, 0 * X), dc1=lambda X, Y, Z: (0 * X, 0 * X, -(Z - 6e-3) * 100),
                  axis=(ax, ax), z_inject=5e-3)
    ph = mqhip.Physics.correct(0.0).copy(dt=1e-7, injection=mqhip.INJ_GAUSSIAN, inj_sigma_mm=0.1, legacy_force=1)
    job = mqhip.job(499.0, 1.0, 0.0, 100.0, 500e3, 0.5, 100e3, 16, 1499)
    obj = mqhip.Objective(1500, 0, 6.0e-3, 2e-4, 1.0e-3, 0.3)
    J, grad, _ = oracle_mod.grad_reference(g, job, [0, 1, 0, 0], ph, obj, 8)
    for k, (grp, name) in enumerate([("drive", "v_rf"), ("drive", "f_rf"), ("drive", "v_ac"), ("drive", "f_ac"),
                                     (None, "dc_scale")]):
        base = getattr(job.drive if grp else job, name)
        h = 1e-6 * abs(base)
        vals = []
        for sgn in (+1, -1):
            j2 = mqhip.Job()
            j2.sp, j2.drive, j2.n, j2.seed, j2.dc_scale = job.sp, job.drive, job.n, job.seed, job.dc_scale
            setattr(j2.drive if grp else j2, name, base + sgn * h)
            vals.append(oracle_mod.grad_reference(g, j2, [0, 1, 0, 0], ph, obj, 8)[0])
        assert (vals[0] - vals[1]) / (2 * h) == pytest.approx(grad[k], rel=1e-5), name


def test_per_ion_subset_matches_batch(oracle_mod):
    """grad_ions on a shuffled subset returns, slot by slot, the per-ion terms whose mean is
    grad_reference: the same ions in a different order and with duplicates give the same
    values, and the full list sums to n times the batch gradient."""
    r0, ax = 2.765e-3, 3.2e-3
    g = make_grid(32, 32, 64, 0.2e-3, rf=lambda X, Y, Z: (2 * (Y - ax) / r0**2, 2 * (X - ax) / r0**2, 0 * X),
                  ac=lambda X, Y, Z: (0 * X - 1 / r0, 0 * X, 0 * X), dc1=lambda X, Y, Z: (0 * X, 0 * X, -(Z - 6e-3) * 100),
                  axis=(ax, ax), z_inject=5e-3)
    ph = mqhip.Physics.correct(0.0).copy(dt=1e-7, injection=mqhip.INJ_GAUSSIAN, inj_sigma_mm=0.1, legacy_force=1)
    n = 24
    job = mqhip.job(499.0, 1.0, 0.0, 100.0, 500e3, 0.5, 100e3, n, 1499)
    obj = mqhip.Objective(400, 0, 6.0e-3, 2e-4, 1.0e-3, 0.3)
    J, grad, _ = oracle_mod.grad_reference(g, job, [0, 1, 0, 0], ph, obj, 4)
    Ji, gi = oracle_mod.grad_ions(g, job, [0, 1, 0, 0], ph, obj, np.arange(n), 3)
    assert Ji.mean() == pytest.approx(J, rel=1e-13)
    assert np.allclose(gi.sum(axis=1) / n, grad, rtol=1e-12, atol=1e-15 * np.abs(grad).max())
    sub = np.array([17, 3, 3, 23, 0])
    Js, gs = oracle_mod.grad_ions(g, job, [0, 1, 0, 0], ph, obj, sub, 2)
    assert np.array_equal(Js, Ji[sub]) and np.array_equal(gs, gi[:, sub])
    with pytest.raises(RuntimeError):
        oracle_mod.grad_ions(g, job, [0, 1, 0, 0], ph, obj, np.array([n]), 1)


def _quad_grid():
    r0, ax = 2.765e-3, 3.2e-3
    return make_grid(32, 32, 64, 0.2e-3, rf=lambda X, Y, Z: (2 * (Y - ax) / r0**2, 2 * (X - ax) / r0**2, 0 * X),
                     ac=lambda X, Y, Z: (0 * X - 1 / r0, 0 * X, 0 * X),
                     dc1=lambda X, Y, Z: (0 * X, 0 * X, -(Z - 6e-3) * 100), axis=(ax, ax), z_inject=5e-3)


def test_frozen_collisions_match_central_differences(oracle_mod):
    """Thermal N2 at 1 mTorr (about five collisions per ion in 1,500 steps): with the Philox draws frozen, the
    dual numbers carry tangents through each collision's velocity map, and central differences of the oracle's
    own objective (same seeds) measure the same pathwise derivative, ion by ion (Richardson over h = 1e-5 and h/2
    relative). An ion whose collision decisions flip between theta - h and theta + h shows as a difference
    quotient that changes with h; such ions are set aside, at most two per parameter and one on average."""
    g = _quad_grid()
    ph = mqhip.Physics.correct(0.133322).copy(dt=1e-7, injection=mqhip.INJ_GAUSSIAN, inj_sigma_mm=0.1,
                                              legacy_force=1)
    n = 24
    job = mqhip.job(499.0, 1.0, 230.0, 100.0, 500e3, 0.5, 100e3, n, 1499)
    obj = mqhip.Objective(1500, 0, 6.0e-3, 2e-4, 1.0e-3, 0.3)
    dc = [0, 1, 0, 0]
    ions = np.arange(n)
    Ji, gi = oracle_mod.grad_ions(g, job, dc, ph, obj, ions, 8)
    # the same ions without gas: the collisions change the result
    Jf, gf = oracle_mod.grad_ions(g, job, dc, ph.copy(pressure_pa=0.0), obj, ions, 8)
    assert not np.allclose(gi, gf, rtol=1e-3)
    kept = 0
    for k, (grp, name) in enumerate([("drive", "v_rf"), ("drive", "f_rf"), ("drive", "v_ac"), ("drive", "f_ac"),
                                     (None, "dc_scale")]):
        base = getattr(job.drive if grp else job, name)

        def central(h):
            vals = []
            for sgn in (+1, -1):
                j2 = mqhip.Job()
                j2.sp, j2.drive, j2.n, j2.seed, j2.dc_scale = job.sp, job.drive, job.n, job.seed, job.dc_scale
                setattr(j2.drive if grp else j2, name, base + sgn * h)
                vals.append(oracle_mod.grad_ions(g, j2, dc, ph, obj, ions, 8)[0])
            return (vals[0] - vals[1]) / (2 * h)

        h = 1e-5 * abs(base)
        c1, c2 = central(h), central(h / 2)
        rich = (4 * c2 - c1) / 3
        scale = np.abs(gi[k]).max()
        smooth = np.abs(c1 - c2) <= 1e-4 * scale
        assert smooth.sum() >= n - 2, (name, int(smooth.sum()))
        kept += int(smooth.sum())
        err = np.abs(rich - gi[k])[smooth] / scale
        assert err.max() <= 1e-5 and np.median(err) <= 1e-6, (name, err.max(), np.median(err))
    assert kept >= 5 * (n - 1)


def test_gradient_rejects_legacy_collisions_and_compat_streams_with_gas(oracle_mod):
    g = _quad_grid()
    job = mqhip.job(499.0, 1.0, 230.0, 100.0, 500e3, 0.5, 100e3, 4, 1499)
    obj = mqhip.Objective(100, 0, 6.0e-3, 2e-4, 1.0e-3, 0.3)
    ph = mqhip.Physics.correct(0.133322).copy(dt=1e-7, injection=mqhip.INJ_GAUSSIAN, inj_sigma_mm=0.1)
    for bad in (ph.copy(legacy_collisions=1), ph.copy(rng=mqhip.RNG_COMPAT)):
        with pytest.raises(RuntimeError):
            oracle_mod.grad_reference(g, job, [0, 1, 0, 0], bad, obj, 2)
    # without gas both are accepted (no collision is drawn)
    oracle_mod.grad_reference(g, job, [0, 1, 0, 0], ph.copy(legacy_collisions=1, pressure_pa=0.0), obj, 2)
