"""GPU field build (include/mq_fieldsolve.h) against the CPU oracle.

* Rasterization: electrode codes identical node for node (integer parity).
* Laplace solve: the GPU conjugate-gradient solution and the oracle's SOR
  solution of the same discrete problem agree to the solver tolerance; both
  reproduce discretely harmonic polynomials; the GPU residual meets tol.
* E field: identical to the oracle's central differences.
* Fields::build on a coarse channel: every basis potential matches the oracle
  solve of the oracle's rasterization; the ch_*.bin cache round-trips.
"""
import os
import sys

import numpy as np
import pytest

import mqhip

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


GEOMS = {
    "default": {},
    "s14": {"grow": 0.4, "q2_flare_bug": 0},
    "tilt1_ac1": {"tilt_mode": 1, "ac_mode": 1, "alpha": 0.525 / 12.0, "grow": 0.4},
    "tilt2_coarse": {"tilt_mode": 2, "n_xy": 80, "h_mm": 0.2},
    "fine_005": {"n_xy": 320, "h_mm": 0.05, "grow": 0.4, "q2_flare_bug": 0},
}


@pytest.mark.parametrize("name", list(GEOMS))
def test_classify_matches_oracle(oracle_mod, name):
    kw = GEOMS[name]
    say(f"[gpu] classify {name}")
    g = mqhip.Geom.default(**kw)
    got = mqhip.geom_classify(g)
    want = oracle_mod.geom_classify(oracle_mod.geom_default(**kw))
    assert got.shape == want.shape
    assert np.array_equal(got, want)


def test_classify_with_misaligned_rods(oracle_mod):
    g = mqhip.Geom.default(grow=0.4)
    w = oracle_mod.geom_default(grow=0.4)
    vals = [[0.1, -0.05, 0.004, 0.0], [0.0, 0.2, 0.0, -0.01], [-0.07, 0.0, 0.002, 0.003], [0.0, 0.0, 0.0, 0.0]]
    for q in range(4):
        for k in range(4):
            g.mis[q][k] = vals[q][k]
            w.mis[q][k] = vals[q][k]
    say("[gpu] classify misaligned")
    assert np.array_equal(mqhip.geom_classify(g), oracle_mod.geom_classify(w))


def _blob_problem(shape, seed):
    rng = np.random.default_rng(seed)
    mask = np.zeros(shape, np.uint8)
    fixed = np.zeros(shape)
    for _ in range(5):
        c = [rng.integers(3, s - 3) for s in shape]
        r = rng.integers(1, 3)
        sl = tuple(slice(ci - r, ci + r + 1) for ci in c)
        mask[sl] = 1
        fixed[sl] = rng.uniform(-2, 2)
    return mask, fixed


@pytest.mark.parametrize("outer", [0.0, 1.0])
def test_laplace_matches_oracle(oracle_mod, outer):
    mask, fixed = _blob_problem((25, 21, 23), 7)
    say(f"[gpu] laplace vs oracle outer={outer}")
    got, st = mqhip.laplace_solve(mask, fixed, outer, 1e-12)
    want, sw = oracle_mod.laplace_sor(mask, fixed, outer, 1e-12)
    assert st["converged"] and sw["converged"]
    assert st["residual"] <= 1e-12
    # both meet the residual tolerance of the same SPD system: they agree far below it
    assert np.abs(got - want).max() < 1e-9
    assert np.array_equal(got[mask == 1], fixed[mask == 1])


def test_laplace_reproduces_harmonic_polynomial():
    nz, ny, nx = 19, 23, 21
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x, y, z = (i - 10.0) / 10, (j - 11.0) / 10, (k - 9.0) / 10
    f = x * x + y * y - 2 * z * z + x * y * z
    mask = np.zeros(f.shape, np.uint8)
    mask[0], mask[-1], mask[:, 0], mask[:, -1], mask[:, :, 0], mask[:, :, -1] = 1, 1, 1, 1, 1, 1
    mask[8:11, 5:7, 12:15] = 1
    say("[gpu] laplace harmonic polynomial")
    phi, st = mqhip.laplace_solve(mask, f, 0.0, 1e-13)
    assert st["converged"]
    assert np.abs(phi - f).max() < 1e-11


def test_laplace_reports_non_convergence():
    mask, fixed = _blob_problem((25, 21, 23), 9)
    say("[gpu] laplace max_iter")
    phi, st = mqhip.laplace_solve(mask, fixed, 0.0, 1e-14, max_iter=3)
    assert not st["converged"] and st["iterations"] <= 32


def test_efield_matches_oracle(oracle_mod):
    rng = np.random.default_rng(3)
    phi = rng.normal(size=(11, 9, 13))
    say("[gpu] efield")
    assert np.array_equal(mqhip.efield(phi, 1e-4), oracle_mod.efield(phi, 1e-4))


def test_fields_build_matches_oracle_and_cache(oracle_mod, tmp_path):
    kw = {"n_xy": 40, "h_mm": 0.4, "grow": 0.4, "q2_flare_bug": 0}
    g = mqhip.Geom.default(**kw)
    say("[gpu] fields build, coarse channel")
    grid = mqhip.channel_grid(g, tol=1e-10, cache_dir=str(tmp_path), potentials=True)
    w = oracle_mod.geom_default(**kw)
    codes = oracle_mod.geom_classify(w)
    mask = (codes != 0).astype(np.uint8)
    assert np.array_equal(grid["mask"].reshape(codes.shape), mask)
    h = g.h_mm * 1e-3
    for b in range(6):
        st = grid["stats"][b]
        assert st["converged"] and st["residual"] <= 1e-10, (b, st)
        fixed = oracle_mod.basis_values(codes, b, g.ac_mode)
        want, _ = oracle_mod.laplace_sor(mask, fixed, 1.0 if b == 5 else 0.0, 1e-12)
        got = grid["phi"][b].reshape(codes.shape)
        assert np.abs(got - want).max() < 1e-7, b
        e = {0: grid["rf"], 1: grid["ac"], 2: grid["dc"][0], 3: grid["dc"][1], 4: grid["dc"][2], 5: grid["dc"][3]}[b]
        assert np.array_equal(e.reshape(codes.shape + (3,)), mqhip.efield(got, h))
        assert os.path.exists(tmp_path / mqhip.field_cache_name(g, b))
    # second build reads the cache: no solve, identical output
    again = mqhip.channel_grid(g, tol=1e-10, cache_dir=str(tmp_path), potentials=True)
    for b in range(6):
        assert again["stats"][b]["iterations"] == 0
        assert np.array_equal(again["phi"][b], grid["phi"][b])
    assert np.array_equal(again["rf"], grid["rf"])
    # geometry-derived scalars
    assert grid["nx"] == 40 and grid["nz"] == int(24.2 / 0.4 + 1e-6)
    assert grid["z_exit"] == pytest.approx(18.6e-3) and grid["axis_x"] == pytest.approx(20 * 0.4e-3)


def test_fields_build_rejects_bad_geometry():
    g = mqhip.Geom.default(h_mm=0.0)
    with pytest.raises(mqhip.MqError, match="h_mm"):
        mqhip.channel_grid(g)
