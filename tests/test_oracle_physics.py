"""Pins of the oracle's integrator against what the mathematics fixes.

Each test is chosen so that a plausible mistake fails it: a wrong kick/drift
order or kick factor (closed-form kinematics, the symplectic-Euler map of a
harmonic well, the Mathieu stability of the full and half kick), a wrong
interpolation corner or weight (affine exactness, brute-force triple sum), a
wrong collision acceptance or scattering law (thermalisation to the bath
temperature, the hard-sphere collision rate), and a wrong classification
order (hand-placed ions on two boundaries at once).
"""
import math

import numpy as np
import pytest

import mqhip
from gridutil import make_grid, mathieu_beta

AMU, QE, KB = mqhip.AMU, mqhip.QE, mqhip.KB


def base_physics(**kw):
    ph = mqhip.Physics.correct(0.0).copy(dt=1e-7, injection=mqhip.INJ_GAUSSIAN, inj_sigma_mm=0.0, temp_k=300.0)
    return ph.copy(**kw)


# ---------------------------------------------------------------- sampling

def test_trilinear_exact_for_affine_field(oracle_mod):
    rng = np.random.default_rng(0)
    nx, ny, nz = 7, 6, 5
    A = rng.normal(size=(3, 3))
    b = rng.normal(size=3)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    G = np.stack([i, j, k], -1).astype(float)
    F = (G @ A.T + b).reshape(-1)
    for _ in range(200):
        g = rng.random(3) * [nx - 1, ny - 1, nz - 1]
        got = oracle_mod.sample(F, nx, ny, nz, g)
        assert np.allclose(got, A @ g + b, rtol=0, atol=1e-12)


def test_trilinear_matches_triple_sum(oracle_mod):
    rng = np.random.default_rng(1)
    nx, ny, nz = 5, 4, 6
    F = rng.normal(size=(nz, ny, nx, 3))
    for _ in range(200):
        g = rng.random(3) * [nx - 1, ny - 1, nz - 1]
        i, j, k = min(int(g[0]), nx - 2), min(int(g[1]), ny - 2), min(int(g[2]), nz - 2)
        t = g - [i, j, k]
        want = np.zeros(3)
        for dk in (0, 1):
            for dj in (0, 1):
                for di in (0, 1):
                    w = (t[0] if di else 1 - t[0]) * (t[1] if dj else 1 - t[1]) * (t[2] if dk else 1 - t[2])
                    want += w * F[k + dk, j + dj, i + di]
        got = oracle_mod.sample(F.reshape(-1), nx, ny, nz, g)
        assert np.allclose(got, want, rtol=0, atol=1e-13)
        # nearest mode: lower-corner node (channel.metal:60)
        assert np.array_equal(oracle_mod.sample(F.reshape(-1), nx, ny, nz, g, 1), F[k, j, i])
    # clamping outside the grid
    g = np.array([-3.0, 10.0, 2.5])
    assert np.allclose(oracle_mod.sample(F.reshape(-1), nx, ny, nz, g),
                       0.5 * (F[2, ny - 1, 0] + F[3, ny - 1, 0]), atol=1e-13)


# ---------------------------------------------------------------- pushes

def _uniform_grid(E0):
    h = 1e-3
    return make_grid(8, 8, 64, h, dc1=lambda X, Y, Z: (0 * X, 0 * X, 0 * X + E0), axis=(3.5e-3, 3.5e-3),
                     z_inject=20e-3)


@pytest.mark.parametrize("force", [0, 1])
def test_uniform_field_closed_form(oracle_mod, force):
    """Kick-drift in a uniform field: x_n = x0 + n v0 dt + dv dt n(n+1)/2."""
    E0, V, n = 3.0, 2.0, 1500
    g = _uniform_grid(E0)
    ph = base_physics(legacy_force=force, max_steps=n, temp_k=1.0)
    job = mqhip.job(499.0, 1.0, 230.0, 0.0, 500e3, 0.0, 100e3, 64, 11)
    r = oracle_mod.run_batch(g, [job], [0, V, 0, 0], ph, 4)
    assert np.all(r["status"] == mqhip.TIMEOUT) and np.all(r["steps"] == n)
    qm = QE / (499.0 * AMU)
    dv = (0.5 if force else 1.0) * qm * ph.dt * V * E0
    for i in range(8):
        x0 = oracle_mod.inject(g, job, ph, i)
        v0z = x0["v"][2]
        want = x0["x"][2] + n * v0z * ph.dt + dv * ph.dt * n * (n + 1) / 2
        assert abs(r["z"][i] - want) <= 1e-12 * abs(want)
        assert abs(r["vz"][i] - (v0z + n * dv)) <= 1e-10 * abs(v0z + n * dv)


def test_harmonic_well_symplectic_map(oracle_mod):
    """E_z = -(z - zc) per volt: the step is v' = v - w^2 dt u, u' = u + v' dt (kick then drift)."""
    zc, V, n = 32e-3, 0.05, 4000
    g = make_grid(8, 8, 64, 1e-3, dc1=lambda X, Y, Z: (0 * X, 0 * X, -(Z - zc)), axis=(3.5e-3, 3.5e-3),
                  z_inject=32.4e-3)
    ph = base_physics(max_steps=n, temp_k=1.0)
    job = mqhip.job(499.0, 1.0, 230.0, 0.0, 500e3, 0.0, 100e3, 16, 5)
    r = oracle_mod.run_batch(g, [job], [0, V, 0, 0], ph, 4)
    w2 = QE / (499.0 * AMU) * V
    dt = ph.dt
    M = np.array([[1 - w2 * dt * dt, dt], [-w2 * dt, 1.0]])
    Mn = np.linalg.matrix_power(M, n)
    assert abs(np.trace(M) / 2 - (1 - w2 * dt * dt / 2)) < 1e-15
    for i in range(16):
        s = oracle_mod.inject(g, job, ph, i)
        u, v = Mn @ np.array([s["x"][2] - zc, s["v"][2]])
        assert abs(r["z"][i] - zc - u) <= 1e-9 * 4e-4
        assert abs(r["vz"][i] - v) <= 1e-9 * max(abs(v), 1.0)


def _rf_quad_grid(r0=2.765e-3):
    h = 0.3e-3
    nx = 24
    ax = (nx // 2) * h
    return make_grid(nx, nx, 8, h, rf=lambda X, Y, Z: (2 * (Y - ax) / r0**2, 2 * (X - ax) / r0**2, 0 * X),
                     mask=lambda X, Y, Z: np.abs((X - ax) * (Y - ax)) >= r0**2 / 2, z_inject=1.05e-3,
                     z_return=0.05e-3, z_exit=2.0e-3)


def test_secular_frequency_matches_mathieu(oracle_mod):
    """Half kick (q ~ 0.51): the transverse motion oscillates at beta f_rf / 2."""
    g = _rf_quad_grid()
    r0, v_rf, f_rf, mz = 2.765e-3, 100.0, 500e3, 499.0
    ph = base_physics(legacy_force=1, temp_k=1e-6, dt=2e-8)
    job = mqhip.job(mz, 1.0, 230.0, v_rf, f_rf, 0.0, 100e3, 1, 3)
    # displace the ion: inject on a shifted axis
    g["axis_x"] += 0.3e-3
    xs = []
    n = 3000
    for k in range(1, n + 1, 3):
        r = oracle_mod.run_batch(g, [job], [0, 0, 0, 0], ph.copy(max_steps=k), 1)
        assert r["status"][0] == mqhip.TIMEOUT
        xs.append(r["x"][0] - (g["axis_x"] - 0.3e-3))
    xs = np.array(xs)
    spec = np.abs(np.fft.rfft(xs * np.hanning(len(xs)), 16 * len(xs)))
    freqs = np.fft.rfftfreq(16 * len(xs), d=3 * ph.dt)
    lo = freqs > 20e3
    f_meas = freqs[lo][np.argmax(spec[lo] * (freqs[lo] < 200e3))]
    q = 0.5 * 4 * QE * v_rf / (mz * AMU * r0**2 * (2 * math.pi * f_rf) ** 2)
    assert 0.45 < q < 0.55
    f_sec = mathieu_beta(0.0, q) * f_rf / 2
    assert abs(f_meas - f_sec) / f_sec < 0.02, (f_meas, f_sec)


@pytest.mark.parametrize("force,stable", [(0, False), (1, True)])
def test_mathieu_stability_full_vs_half_kick(oracle_mod, force, stable):
    """q_full ~ 1.02 > 0.908 (unstable: every ion is lost radially); q_half ~ 0.51 (stable)."""
    g = _rf_quad_grid()
    ph = base_physics(legacy_force=force, temp_k=10.0, inj_sigma_mm=0.2, max_steps=4000, dt=1e-7)
    job = mqhip.job(499.0, 1.0, 230.0, 100.0, 500e3, 0.0, 100e3, 200, 9)
    r = oracle_mod.run_batch(g, [job], [0, 0, 0, 0], ph, 8)
    lost = np.mean((r["status"] == mqhip.CRASHED) | (r["status"] == mqhip.LEFT))
    if stable:
        assert lost < 0.02
    else:
        assert lost > 0.97


# ---------------------------------------------------------------- collisions

def _bath_grid():
    return make_grid(6, 6, 6, 80e-3, axis=(200e-3, 200e-3), z_inject=200e-3)


def test_thermalisation_to_bath(oracle_mod):
    """Ions injected hot (legacy cone, ~0.1 eV) relax to a Maxwellian at the bath temperature.

    The candidate probability n sigma (|v| + 6 s_gas) dt must stay well below 1 (here ~0.03):
    the one-candidate-per-step null-collision scheme of channel.metal:94-95 saturates above it."""
    g = _bath_grid()
    ph = base_physics(pressure_pa=1.0, dt=2e-8, injection=mqhip.INJ_REFERENCE, legacy_injection=1, max_steps=20000)
    job = mqhip.job(499.0, 1.0, 230.0, 0.0, 500e3, 0.0, 100e3, 3000, 21)
    r = oracle_mod.run_batch(g, [job], [0, 0, 0, 0], ph, 8)
    assert np.all(r["status"] == mqhip.TIMEOUT)
    m = 499.0 * AMU
    v = np.concatenate([r["vx"], r["vy"], r["vz"]])
    kt = m * np.mean(v**2) / KB
    se = 300.0 * math.sqrt(2.0 / len(v))
    assert abs(kt - 300.0) < 4 * se, kt
    # component distribution is Gaussian: kurtosis 3
    k4 = np.mean(v**4) / np.mean(v**2) ** 2
    assert abs(k4 - 3.0) < 0.15, k4


def test_hard_sphere_collision_rate(oracle_mod):
    """At equilibrium the rate per step is n sigma <g> dt with <g> = sqrt(8 kT / (pi mu))."""
    g = _bath_grid()
    p, ccs, steps, n = 1.0, 230.0, 4000, 2000
    ph = base_physics(pressure_pa=p, max_steps=steps, inj_sigma_mm=0.0)
    job = mqhip.job(499.0, 1.0, ccs, 0.0, 500e3, 0.0, 100e3, n, 22)
    r = oracle_mod.run_batch(g, [job], [0, 0, 0, 0], ph, 8)
    m, M = 499.0 * AMU, 28.0 * AMU
    mu = m * M / (m + M)
    gbar = math.sqrt(8 * KB * 300.0 / (math.pi * mu))
    rate = p / (KB * 300.0) * ccs * 1e-20 * gbar * ph.dt
    expect = rate * steps
    mean = r["ncoll"].mean()
    se = math.sqrt(expect / n)
    assert abs(mean - expect) < 4 * se, (mean, expect)


# ---------------------------------------------------------------- classification

def _placed(oracle_mod, z_inject, axis, **gk):
    g = make_grid(10, 10, 10, 1e-3, axis=axis, z_inject=z_inject, **gk)
    ph = base_physics(temp_k=0.0, max_steps=1)
    job = mqhip.job(499.0, 1.0, 230.0, 0.0, 500e3, 0.0, 100e3, 1, 1)
    return int(oracle_mod.run_batch(g, [job], [0, 0, 0, 0], ph, 1)["status"][0]), g, ph


def test_classification_order(oracle_mod):
    """Return before exit before box before crash before timeout (channel.metal:114-120)."""
    inside = (4.5e-3, 4.5e-3)
    assert _placed(oracle_mod, 4e-3, inside)[0] == mqhip.TIMEOUT
    assert _placed(oracle_mod, 1e-3, inside, z_return=2e-3)[0] == mqhip.RETURNED
    assert _placed(oracle_mod, 1e-3, (-1e-3, 4.5e-3), z_return=2e-3)[0] == mqhip.RETURNED
    assert _placed(oracle_mod, 8e-3, (-1e-3, 4.5e-3), z_exit=7e-3)[0] == mqhip.TRANSMITTED
    assert _placed(oracle_mod, 4e-3, (-1e-3, 4.5e-3), mask=lambda X, Y, Z: X < 1)[0] == mqhip.LEFT
    assert _placed(oracle_mod, 4e-3, inside, mask=lambda X, Y, Z: X < 1)[0] == mqhip.CRASHED
    # crash uses the nearest node: a conductor plane at x = 5 mm catches x = 4.6 mm but not 4.4 mm
    plane = dict(mask=lambda X, Y, Z: np.abs(X - 5e-3) < 1e-9)
    assert _placed(oracle_mod, 4e-3, (4.6e-3, 4.5e-3), **plane)[0] == mqhip.CRASHED
    assert _placed(oracle_mod, 4e-3, (4.4e-3, 4.5e-3), **plane)[0] == mqhip.TIMEOUT


def test_legacy_criteria_window(oracle_mod):
    """Legacy criteria test the conductor only inside the rod window (channel.rs:406-407)."""
    g = make_grid(10, 10, 10, 1e-3, axis=(4.5e-3, 4.5e-3), z_inject=4e-3, mask=lambda X, Y, Z: X < 1,
                  crash=(5e-3, 8e-3))
    job = mqhip.job(499.0, 1.0, 230.0, 0.0, 500e3, 0.0, 100e3, 1, 1)
    ph = base_physics(temp_k=0.0, max_steps=1, legacy_criteria=1)
    assert oracle_mod.run_batch(g, [job], [0, 0, 0, 0], ph, 1)["status"][0] == mqhip.TIMEOUT
    g["z_inject"] = 6e-3
    assert oracle_mod.run_batch(g, [job], [0, 0, 0, 0], ph, 1)["status"][0] == mqhip.CRASHED


# ---------------------------------------------------------------- injection

def test_injection_distributions(oracle_mod):
    g = make_grid(10, 10, 10, 1e-3, axis=(4.5e-3, 4.5e-3), z_inject=1e-3)
    n = 4000
    m = 499.0 * AMU
    job = mqhip.job(499.0, 1.0, 230.0, 0.0, 500e3, 0.0, 100e3, n, 77)
    # Gaussian spot + Maxwell-Boltzmann
    ph = base_physics(inj_sigma_mm=0.1, temp_k=300.0)
    s = [oracle_mod.inject(g, job, ph, i) for i in range(n)]
    x = np.array([q["x"][0] for q in s]) - 4.5e-3
    v = np.array([q["v"] for q in s])
    assert abs(x.std() - 1e-4) < 4 * 1e-4 / math.sqrt(2 * n)
    assert abs(np.mean(v**2) * m / KB - 300.0) < 4 * 300.0 * math.sqrt(2 / (3 * n))
    assert all(q["x"][2] == 1e-3 for q in s)
    phis = np.array([q["phi_ac"] for q in s])
    assert 0 < phis.min() and phis.max() < 1 and abs(phis.mean() - 0.5) < 4 * math.sqrt(1 / 12 / n)
    # legacy cone (channel.rs:350-355): KE in [0.075 z, 0.15 z] eV, |angle| <= 10 degrees, on axis
    ph = base_physics(injection=mqhip.INJ_REFERENCE, legacy_injection=1)
    job5 = mqhip.job(2495.0, 5.0, 0.0, 0.0, 500e3, 0.0, 100e3, n, 1499)
    s = [oracle_mod.inject(g, job5, ph, i) for i in range(500)]
    ke = np.array([0.5 * 2495.0 * AMU * np.dot(q["v"], q["v"]) / QE for q in s])
    assert ke.min() >= 0.075 * 5 * 0.99 and ke.max() <= 0.15 * 5 * 1.0001
    ang = np.array([math.degrees(math.atan2(math.hypot(q["v"][0], q["v"][1]), q["v"][2])) for q in s])
    assert ang.max() <= 10.0 * 0.7071 * 1.4143 + 1e-9
    assert all(q["x"][:2] == [4.5e-3, 4.5e-3] for q in s)
    # thermal disc (channel.rs:357-361): r <= r_inj, flux-weighted vz > 0
    ph = base_physics(injection=mqhip.INJ_REFERENCE, legacy_injection=0, r_inject_mm=2.0)
    s = [oracle_mod.inject(g, job, ph, i) for i in range(n)]
    r = np.array([math.hypot(q["x"][0] - 4.5e-3, q["x"][1] - 4.5e-3) for q in s])
    assert r.max() <= 2e-3 and abs(np.mean(r**2) - 2e-6) < 4 * 2e-6 / math.sqrt(3 * n)
    vz = np.array([q["v"][2] for q in s])
    assert vz.min() > 0
    # Rayleigh: <vz^2> = 2 kT/m
    assert abs(np.mean(vz**2) / (2 * KB * 300.0 / m) - 1) < 4 * 1.0 / math.sqrt(n)


def test_summarize_counts(oracle_mod):
    rng = np.random.default_rng(4)
    n, nj = 5000, 7
    res = {"status": rng.integers(1, 6, n).astype(np.uint8), "steps": rng.integers(0, 9999, n).astype(np.uint32),
           "ncoll": rng.integers(0, 50, n).astype(np.uint32), "job": rng.integers(0, nj, n).astype(np.uint32)}
    out = oracle_mod.summarize(res, nj, 1e-7)
    for j in range(nj):
        sel = res["job"] == j
        st = res["status"][sel]
        o = out[j]
        assert o.n == sel.sum()
        assert (o.crashed, o.returned, o.transmitted, o.left, o.timeout) == tuple(int((st == k).sum()) for k in range(1, 6))
        assert abs(o.mean_time - res["steps"][sel].mean() * 1e-7) < 1e-18
        assert abs(o.mean_coll - res["ncoll"][sel].mean()) < 1e-12
