"""The reference's random streams (MQ_RNG_COMPAT): xorshift64 injection
(Engine::inject, channel.rs:342-374) and xoshiro128** per ion
(channel.metal:38-53). CPU only: the generators, the oracle's injection and
the product's host injection (mq_inject_reference), which must agree bit for
bit."""
import math

import numpy as np
import pytest

import mqhip
import synth

M64 = (1 << 64) - 1


def test_xoshiro128ss_published_vector(oracle_mod):
    # rand_xoshiro's known-answer test of Xoshiro128StarStar seeded with the words 1, 2, 3, 4
    assert oracle_mod.xoshiro128ss([1, 2, 3, 4], 10) == [11520, 0, 5927040, 70819200, 2031721883, 1637235492,
                                                         1287239034, 3734860849, 3729100597, 4258142804]


def test_xorshift64_first_output_and_linearity(oracle_mod):
    # state 1: x ^= x << 13 -> 8193; x ^= x >> 7 -> 8257; x ^= x << 17 -> 8257 + 8257 * 2^17
    assert oracle_mod.xorshift64(1, 1) == [8257 + 8257 * (1 << 17)]
    # the step is linear over GF(2): next(a ^ b) = next(a) ^ next(b)
    rng = np.random.default_rng(4)
    for _ in range(50):
        a, b = (int(v) for v in rng.integers(1, 1 << 63, 2))
        assert oracle_mod.xorshift64(a ^ b, 1)[0] == oracle_mod.xorshift64(a, 1)[0] ^ oracle_mod.xorshift64(b, 1)[0]


def _physics(kind):
    if kind == "legacy":
        return mqhip.Physics.legacy(1e-3 * 133.322).copy(rng=mqhip.RNG_COMPAT)
    if kind == "thermal":
        return mqhip.Physics.correct(0.1).copy(rng=mqhip.RNG_COMPAT)
    return mqhip.Physics.correct(0.133322).copy(dt=1e-7, injection=mqhip.INJ_GAUSSIAN, inj_sigma_mm=0.1,
                                                rng=mqhip.RNG_COMPAT)


@pytest.mark.parametrize("kind", ["legacy", "thermal", "gaussian"])
def test_product_host_injection_equals_oracle(oracle_mod, kind):
    g = synth.c1_grid()
    j = mqhip.job(5 * 499.0, 5.0, 0.0, 100.0, 500e3, 0.5, 100e3, 500, 1499)
    ph = _physics(kind)
    a = mqhip.inject_reference(g, j, ph)
    b = oracle_mod.inject_reference(g, j, ph)
    for k, kb in (("x", "x"), ("y", "y"), ("z", "z"), ("vx", "vx"), ("vy", "vy"), ("vz", "vz"), ("cd", "cd"),
                  ("phi_ac0", "phi_ac")):
        assert np.array_equal(a[k], b[kb]), k
    assert np.array_equal(a["rng"], b["rng"])
    assert np.all(a["rng"][:, 0] & 1 == 1)  # the first seed word has its low bit set


def test_first_draw_of_the_stream(oracle_mod):
    # the first ion's kinetic energy is 0.1 (0.75 + 0.75 u) z eV with u the first
    # xorshift64 draw of the job stream, seeded seed * 0x9E3779B97F4A7C15 ^ 0xD1B54A32D192ED03
    seed, z, mass = 1499, 5.0, 5 * 499.0
    st = ((seed * 0x9E3779B97F4A7C15) & M64) ^ 0xD1B54A32D192ED03
    st ^= (st << 13) & M64
    st ^= st >> 7
    st ^= (st << 17) & M64
    u = (st >> 11) / 2.0**53
    st ^= (st << 13) & M64
    st ^= st >> 7
    st ^= (st << 17) & M64
    u_ang = (st >> 11) / 2.0**53
    g = synth.c1_grid()
    j = mqhip.job(mass, z, 0.0, 100.0, 500e3, 0.5, 100e3, 3, seed)
    r = oracle_mod.inject_reference(g, j, _physics("legacy"))
    amu, qe = 1.6605390666e-27, 1.602176634e-19
    kin = 0.1 * (0.75 + 0.75 * u) * z
    vin = math.sqrt(2 * kin * qe / (mass * amu))
    ang = math.radians(-10 + 20 * u_ang)
    assert r["vz"][0] == pytest.approx(vin * math.cos(ang), rel=1e-6)   # f32-rounded
    assert r["vx"][0] == pytest.approx(0.707 * vin * math.sin(ang), rel=1e-6, abs=1e-6)
    # the position is the axis on the reference's 62.5 pm fixed-point grid, truncated
    assert r["x"][0] == int(g["axis_x"] * 1.6e10) / 1.6e10 and r["z"][0] == int(1e-3 * 1.6e10) / 1.6e10


def test_reference_injection_distributions(oracle_mod):
    g = synth.c1_grid()
    ph = _physics("thermal")
    j = mqhip.job(1726.946, 2.0, 425.0, 70.0, 500e3, 0.35, 84.25e3, 20000, 4242)
    r = oracle_mod.inject_reference(g, j, ph)
    rad = np.hypot(r["x"] - g["axis_x"], r["y"] - g["axis_y"])
    # uniform disc of radius r_inject = 2 mm: r^2 is uniform, mean r^2 = R^2 / 2
    assert np.max(rad) <= 2e-3 + 1e-10
    assert np.mean(rad**2) == pytest.approx((2e-3) ** 2 / 2, rel=0.03)
    # transverse Maxwellian at 300 K and the flux-weighted (Rayleigh) axial speed: <vz^2> = 2 kT/m
    kT_m = 1.380649e-23 * 300.0 / (1726.946 * 1.6605390666e-27)
    assert np.var(r["vx"]) == pytest.approx(kT_m, rel=0.05)
    assert np.mean(r["vz"] ** 2) == pytest.approx(2 * kT_m, rel=0.05)
    assert np.all(r["vz"] > 0)
    # AC phase in [0, 1), countdown >= 1, distinct seeds per ion
    assert np.all((r["phi_ac"] >= 0) & (r["phi_ac"] <= 1)) and np.all(r["cd"] >= 1)
    assert len({tuple(w) for w in r["rng"][:1000]}) == 1000


def test_compat_rejected_where_not_supported():
    g = synth.c1_grid()
    with pytest.raises(mqhip.MqError, match="dt"):
        mqhip.inject_reference(g, mqhip.job(499.0, 1.0, 0.0, 100.0, 500e3, 0.5, 100e3, 3, 1),
                               _physics("thermal").copy(dt=0.0))
