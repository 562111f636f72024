"""Scan API host logic (no GPU): Rust notation of the cache keys, the key
layout the spectrum model's reader parses (spectrum_trans.rs:28-48), and the
adaptive refinement rule (scan.rs:59-67)."""
import math

import numpy as np
import pytest

import mqhip

# Rust {:?} / {} of f64 (core::fmt float: shortest round-trip digits; Debug is
# decimal with a ".0" minimum for 1e-4 <= |x| < 1e16 and exponential outside).
FLOATS = [
    (5.0, "5.0", "5"), (0.5, "0.5", "0.5"), (-5.0, "-5.0", "-5"), (0.0, "0.0", "0"),
    (1e-7, "1e-7", "0.0000001"), (4e-8, "4e-8", "0.00000004"), (1.5e-7, "1.5e-7", "0.00000015"),
    (500000.0, "500000.0", "500000"), (84250.0, "84250.0", "84250"), (1726.946, "1726.946", "1726.946"),
    (0.0001, "0.0001", "0.0001"), (9e-5, "9e-5", "0.00009"), (1e15, "1000000000000000.0", "1000000000000000"),
    (1e16, "1e16", "10000000000000000"), (1.2345e20, "1.2345e20", "123450000000000000000"),
    (0.1 + 0.2, "0.30000000000000004", "0.30000000000000004"), (1.0 / 3.0, "0.3333333333333333", "0.3333333333333333"),
]


@pytest.mark.parametrize("x,dbg,disp", FLOATS)
def test_rust_float_notation(x, dbg, disp):
    assert mqhip.format_f64(x) == dbg
    assert mqhip.format_f64(x, display=True) == disp


def test_shortest_digits_match_python_repr():
    rng = np.random.default_rng(8)
    for x in np.concatenate([rng.random(500) * 10.0 ** rng.integers(-3, 15, 500), rng.random(200)]):
        d = mqhip.format_f64(float(x))
        assert float(d) == float(x)
        r = repr(float(x))
        digits = lambda s: s.replace("-", "").split("e")[0].replace(".", "").strip("0")
        assert digits(d) == digits(r), (x, d, r)


def test_key_layout_matches_reference_reader():
    ph = mqhip.Physics.legacy(1e-3 * 133.322)
    j = mqhip.job(2495.0, 5.0, 0.0, 100.0, 500e3, 0.5, 100e3, 1000, 1499)
    k = mqhip.scan_key("Geom { de: 6.35 }", j, [0.0, 5.0, -5.0, 0.0], ph)
    p = k.split("|")
    assert len(p) == 7
    assert p[0] == "Geom { de: 6.35 }"
    assert p[1] == "Species { mass: 2495.0, z: 5.0, ccs_n2: 0.0 }"
    assert p[2] == "RfAc { v_rf: 100.0, f_rf: 500000.0, v_ac: 0.5, f_ac: 100000.0 }"
    assert (p[3], p[4]) == ("1000", "1499")
    assert p[5] == "[0.0, 5.0, -5.0, 0.0]"
    pr = repr(1e-3 * 133.322)
    assert p[6] == ("Physics { legacy_force: true, legacy_collisions: true, legacy_fields: true, legacy_injection: true, "
                    f"legacy_criteria: true, pressure_pa: {pr}, temp_k: 300.0, dt: 1e-7, t_max: 1.0, r_inject_mm: 0.0 }}")
    # the reader's number extraction (spectrum_trans.rs:18-22) finds every field
    def num(s, key):
        i = s.index(key) + len(key)
        t = ""
        while i < len(s) and (s[i].isdigit() or s[i] in ".-eE+"):
            t += s[i]
            i += 1
        return float(t)
    assert (num(p[1], "mass: "), num(p[1], "z: "), num(p[1], "ccs_n2: ")) == (2495.0, 5.0, 0.0)
    assert (num(p[2], "v_rf: "), num(p[2], "f_rf: "), num(p[2], "v_ac: "), num(p[2], "f_ac: ")) == (100.0, 500e3, 0.5, 100e3)
    # extensions appear only when used
    j2 = mqhip.job(2495.0, 5.0, 0.0, 100.0, 500e3, 0.5, 100e3, 1000, 1499, dc_scale=0.5, pressure_pa=0.2)
    k2 = mqhip.scan_key("G", j2, [0.0, 5.0, -5.0, 0.0], ph.copy(injection=mqhip.INJ_GAUSSIAN, inj_sigma_mm=0.1,
                                                                 max_steps=20000))
    assert "dc_scale: 0.5, pressure_pa: 0.2 }" in k2 and k2.endswith("injection: 1, inj_sigma_mm: 0.1, max_steps: 20000 }")


def fine_points_reference(fs, t1, n_lo, base, f_step):
    """scan.rs:57-67 restated."""
    tmax = max(t1)
    fine = []
    for f, t in zip(fs, t1):
        sig = math.sqrt(max(t, 1.0 / n_lo) / n_lo)
        if t >= 0.35 * tmax and t - base > 4.0 * sig:
            g = f - f_step
            while g <= f + f_step + 1e-9:
                fine.append(round(g * 1e4) / 1e4)
                g += f_step / 4.0
    return sorted(set(fine))


def test_adaptive_refinement_rule():
    rng = np.random.default_rng(2)
    for trial in range(50):
        fs = [70.0 + 2.0 * i for i in range(21)]
        peak = rng.uniform(72, 108)
        t1 = [round(0.6 * math.exp(-((f - peak) / 3.0) ** 2) * 200) / 200 for f in fs]
        base = rng.uniform(0, 0.05)
        got = mqhip.adaptive_fine_points(fs, t1, 200, base, 2.0)
        assert got == fine_points_reference(fs, t1, 200, base, 2.0)
    # a flat scan at the baseline refines nothing
    assert mqhip.adaptive_fine_points([1.0, 2.0, 3.0], [0.0, 0.0, 0.0], 100, 0.0, 1.0) == []


def test_key_separates_streams_and_paths():
    """A request for the reference's own random streams (rng COMPAT) or for the FP64 plain path must not
    find a line that a default (Philox, FP32 batched) run wrote, and the reverse; the default key stays the
    reference's format (scan.rs:9-11), with no extension field."""
    ph = mqhip.Physics.legacy(1e-3 * 133.322)
    j = mqhip.job(2495.0, 5.0, 0.0, 100.0, 500e3, 0.5, 100e3, 1000, 1499)
    dc = [0.0, 5.0, -5.0, 0.0]
    default = mqhip.scan_key("G", j, dc, ph)
    assert default == mqhip.scan_key("G", j, dc, ph, mqhip.PATH_BATCHED)
    assert "rng" not in default and "path" not in default
    compat = mqhip.scan_key("G", j, dc, ph.copy(rng=mqhip.RNG_COMPAT))
    plain = mqhip.scan_key("G", j, dc, ph, mqhip.PATH_PLAIN)
    both = mqhip.scan_key("G", j, dc, ph.copy(rng=mqhip.RNG_COMPAT), mqhip.PATH_PLAIN)
    assert len({default, compat, plain, both}) == 4
    assert compat.endswith("r_inject_mm: 0.0, rng: 1 }")
    assert plain.endswith("r_inject_mm: 0.0, path: 1 }")
    assert both.endswith("r_inject_mm: 0.0, rng: 1, path: 1 }")
    # everything before the Physics extension fields is the default key
    for k in (compat, plain, both):
        assert k.split("|")[:6] == default.split("|")[:6]
        assert k.startswith(default[:-2])
