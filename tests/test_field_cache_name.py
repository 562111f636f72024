"""The field-cache file name (cache_path, channel.rs:128-138): ch_<h>.bin with h the 64-bit FNV-1a of
format!("{:?}{:?}", geom, basis) from the reference's own offset constant 1469598103934665603 (one digit short of
the FNV suite's 14695981039346656037) and prime 1099511628211. The expected Debug string is written here from the
field order of `struct Geom` (channel.rs:18-36) and Rust's Debug rules (floats always with a fraction or an
exponent, bools as true/false, integers plain, arrays as [a, b]); host only."""
import pytest

import mqhip


def rust_f64(x: float) -> str:
    """Rust {:?} of an f64 in the ranges a Geom holds (shortest round-trip digits, '.0' on integral values)."""
    if x == int(x) and abs(x) < 1e16:
        return f"{x:.1f}"
    return repr(x)


def geom_debug(g: mqhip.Geom) -> str:
    f = rust_f64
    mis = ", ".join("[" + ", ".join(f(g.mis[q][k]) for k in range(4)) + "]" for q in range(4))
    return (f"Geom {{ de: {f(g.de)}, coef: {f(g.coef)}, q1_start: {f(g.q1_start)}, q1_len: {f(g.q1_len)}, "
            f"alpha: {f(g.alpha)}, gap: {f(g.gap)}, plate: {f(g.plate)}, hole_d: {f(g.hole_d)}, "
            f"q2_len: {f(g.q2_len)}, stem_d: {f(g.stem_d)}, q2_flare_bug: {'true' if g.q2_flare_bug else 'false'}, "
            f"grow: {f(g.grow)}, tilt_mode: {g.tilt_mode}, ac_mode: {g.ac_mode}, mis: [{mis}], n_xy: {g.n_xy}, "
            f"h_mm: {f(g.h_mm)} }}")


def fnv_ref(s: str) -> int:
    h = 1469598103934665603
    for c in s.encode():
        h ^= c
        h = (h * 1099511628211) % (1 << 64)
    return h


BASES = ["Rf", "Ac", "DcQ1", "DcPlate", "DcDown", "DcBox"]


def test_default_geom_debug_string():
    g = mqhip.Geom.default()
    assert geom_debug(g) == ("Geom { de: 6.35, coef: 0.323, q1_start: 3.0, q1_len: 12.0, alpha: 0.02, gap: 1.0, "
                             "plate: 1.6, hole_d: 7.0, q2_len: 3.0, stem_d: 4.6, q2_flare_bug: true, grow: 0.0, "
                             "tilt_mode: 0, ac_mode: 0, mis: [[0.0, 0.0, 0.0, 0.0], [0.0, 0.0, 0.0, 0.0], "
                             "[0.0, 0.0, 0.0, 0.0], [0.0, 0.0, 0.0, 0.0]], n_xy: 160, h_mm: 0.1 }")


@pytest.mark.parametrize("variant", ["default", "s14", "inset", "asbuilt4", "misaligned"])
def test_cache_name_is_fnv_of_the_debug_strings(variant):
    if variant == "default":
        g = mqhip.Geom.default()
    elif variant == "s14":
        g = mqhip.Geom.s14()
    elif variant == "inset":   # s14_geom "inset" (channel.rs:609)
        g = mqhip.Geom.default(grow=0.4, q2_flare_bug=0, tilt_mode=1, ac_mode=1, alpha=0.525 / 12.0)
    elif variant == "asbuilt4":   # channel.rs:613
        g = mqhip.Geom.default(grow=0.4, q2_flare_bug=0, de=6.34, coef=(8.30 - 6.34) / 6.34, alpha=0.021,
                               hole_d=5.75, ac_mode=1)
    else:
        g = mqhip.Geom.default()
        g.mis[1][0] = 0.05
        g.mis[2][3] = -0.001
    for b, name in enumerate(BASES):
        want = f"ch_{fnv_ref(geom_debug(g) + name):016x}.bin"
        assert mqhip.field_cache_name(g, b) == want, (variant, name)
