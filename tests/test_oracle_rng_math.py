"""Pins of the oracle's random-number generator and elementary functions.

Philox4x32-10 is pinned by the Random123 known-answer vectors (Salmon et al.,
SC'11, kat_vectors) and cross-checked against an independent implementation
(rocRAND's philox4x32_10 engine, compiled from its ROCm header). sin/cos/log are
pinned against the platform libm, and their polynomial constants are recomputed
from their mathematical definitions with 60-digit decimal arithmetic, in both
the oracle source and the device header.
"""
import math
import os
import re
import shutil
import subprocess
from decimal import Decimal, getcontext

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

KATS = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", KATS)
def test_philox_kat(oracle_mod, ctr, key, want):
    assert tuple(oracle_mod.philox(ctr, key)) == want


ROCRAND_SRC = r"""
#include <hip/hip_runtime.h>
#include <rocrand/rocrand_kernel.h>
#include <cstdio>
struct E : rocrand_device::philox4x32_10_engine {
  using philox4x32_10_engine::philox4x32_10_engine;
  uint4 tr(uint4 c, uint2 k) { return ten_rounds(c, k); }
};
int main() {
  E e(0, 0, 0);
  unsigned a[6];
  while (scanf("%u %u %u %u %u %u", a, a + 1, a + 2, a + 3, a + 4, a + 5) == 6) {
    uint4 r = e.tr(uint4{a[0], a[1], a[2], a[3]}, uint2{a[4], a[5]});
    printf("%u %u %u %u\n", r.x, r.y, r.z, r.w);
  }
}
"""


@pytest.mark.skipif(shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="hipcc not available")
def test_philox_matches_rocrand(oracle_mod, tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    src = tmp_path / "kat.cpp"
    src.write_text(ROCRAND_SRC)
    exe = tmp_path / "kat"
    subprocess.run([hipcc, "-x", "hip", "--offload-arch=gfx950", "-O1", str(src), "-o", str(exe)], check=True,
                   capture_output=True)
    rng = np.random.default_rng(5)
    rows = rng.integers(0, 2**32, size=(300, 6), dtype=np.uint64)
    inp = "\n".join(" ".join(str(int(v)) for v in r) for r in rows)
    out = subprocess.run([str(exe)], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    for r, line in zip(rows, out):
        want = [int(v) for v in line.split()]
        got = oracle_mod.philox([int(v) for v in r[:4]], [int(v) for v in r[4:]])
        assert got == want


def test_sincos_against_libm(oracle_mod):
    rng = np.random.default_rng(1)
    ps = np.concatenate([rng.random(4000), np.linspace(0, 1, 257), [0.125, 0.375, 1e-9, 0.9999999999]])
    # reference in x87 extended precision: the argument 2 pi p must not be rounded to double first
    two_pi = 2 * np.longdouble("3.14159265358979323846264338327950288")
    for p in ps:
        s, c = oracle_mod.sincos2pi(float(p))
        arg = two_pi * np.longdouble(float(p))
        assert abs(s - float(np.sin(arg))) <= 2.5e-16, p
        assert abs(c - float(np.cos(arg))) <= 2.5e-16, p
    # exact special values
    assert oracle_mod.sincos2pi(0.0) == (0.0, 1.0)
    assert oracle_mod.sincos2pi(0.25) == (1.0, -0.0) or oracle_mod.sincos2pi(0.25)[0] == 1.0
    assert oracle_mod.sincos2pi(0.5)[1] == -1.0


def test_log_against_libm(oracle_mod):
    rng = np.random.default_rng(2)
    xs = np.concatenate([rng.random(3000) * (1 - 2**-33) + 2**-33, rng.random(1000) * 1e6 + 1.0,
                         [2**-33, 0.5, 1.0, 2.0, math.sqrt(0.5), 1e-7]])
    for x in xs:
        got, want = oracle_mod.log(float(x)), math.log(x)
        assert abs(got - want) <= 4 * 2.2e-16 * max(abs(want), 1e-300) + 1e-300, (x, got, want)
    assert oracle_mod.log(1.0) == 0.0


def _hexfloats(text):
    return [float.fromhex(m) for m in re.findall(r"-?0x1\.[0-9a-f]+p[-+]\d+", text)]


def test_polynomial_constants_from_definition():
    getcontext().prec = 60
    pi = Decimal("3.14159265358979323846264338327950288419716939937510582097494459")
    ln2 = Decimal("0.693147180559945309417232121458176568075500134360255254120680009")
    h = pi / 2
    fact = math.factorial
    S = [float((-1) ** k * h ** (2 * k + 1) / fact(2 * k + 1)) for k in range(10)]
    Cc = [float((-1) ** k * h ** (2 * k) / fact(2 * k)) for k in range(10)]
    L = [float(Decimal(1) / (2 * k + 1)) for k in range(1, 13)]
    hi = float.fromhex("0x1.62e4200000000p-1")
    lo = float(ln2 - Decimal(hi))
    want = set(S + Cc + L + [hi, lo])
    for path in ("oracle/oracle.c", "multiq-workbench_amd/csrc/mq_math.h"):
        text = open(os.path.join(ROOT, path)).read()
        consts = set(_hexfloats(text)) - {float.fromhex("0x1.6a09e667f3bcdp-1"), 1.0}
        missing = want - set(_hexfloats(text)) - {1.0}
        assert not missing, (path, missing)
        assert consts <= want | {1.0}, (path, consts - want)
    # sqrt(1/2) threshold is the double nearest to sqrt(0.5)
    assert float.fromhex("0x1.6a09e667f3bcdp-1") == math.sqrt(0.5)
    # LN2_HI has its low 32 mantissa bits clear so e * LN2_HI is exact for |e| < 2^21
    assert hi + lo == float(ln2)
