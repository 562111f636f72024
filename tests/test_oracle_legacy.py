"""The oracle of the legacy Java potentials (oracle/oracle_legacy.c), pinned
without running the sketch (CPU only):

* the grid and check point the sketch's constants fix, and the size
  Fields::from_theirs hard-codes (channel.rs:191);
* the float32 rasterization of GEOMETRY.pde against an independent one: the
  Rust model's Geom::classify and basis values for Geom::default, which the
  reference states is "the same rasterization as theirs" (channel.rs:688),
  restated separately in FP64 in oracle_fields.c. Every one of the 6.2M nodes
  must agree in electrode flag and in value (RF; DC = plate - downstream, the
  basis from_theirs puts in the plate slot, channel.rs:188-189; AC);
* the sweep: its fixed point is the discrete Laplace solution (a dense direct
  solve on small grids), and one sweep equals the sketch's loop run node by
  node in its order (z outer, x, y inner), which a sweep with an axis run
  backwards or a Jacobi step does not (any order that visits a node's three
  lower neighbours before it gives the same sweep, the property the GPU's
  wavefronts rest on);
* the stop rule's arithmetic on hand-evaluated cases, including the sketch's
  quirks (a check point at 0 never stops the loop before the cap).
"""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
RF, DC, AC = 0, 1, 2


def test_dimensions_and_check_point(oracle_mod):
    # xDevice = yDevice = 16 mm, zDevice = 3 + 12 + 1 + 1.6 + 1 + 3 + 1 + 1.6 = 24.2 mm at 10 nodes/mm;
    # check point (9, 8.5, 14) mm -> nodes (90, 85, 140)
    assert oracle_mod.legacy_dims(10.0) == (160, 160, 242, (90, 85, 140))
    assert oracle_mod.legacy_dims(2.5) == (40, 40, 60, (22, 21, 35))


@pytest.fixture(scope="module")
def rust_model(oracle_mod):
    g = oracle_mod.geom_default()
    codes = oracle_mod.geom_classify(g)
    vals = {b: oracle_mod.basis_values(codes, b, 0) for b in range(6)}
    return codes, vals


@pytest.mark.parametrize("basis", [RF, DC, AC])
def test_rasterization_equals_the_rust_models(oracle_mod, rust_model, basis):
    codes, vals = rust_model
    phi, el = oracle_mod.legacy_rasterize(basis, 10.0)
    assert phi.shape == codes.shape == (242, 160, 160)
    assert np.array_equal(el, (codes != 0).astype(np.uint8))
    want = {RF: vals[0], DC: vals[3] - vals[4], AC: vals[1]}[basis]
    assert np.array_equal(phi, want.astype(np.float32))
    assert np.all(phi[el == 0] == 0)


def test_rasterization_landmarks(oracle_mod):
    phi, el = oracle_mod.legacy_rasterize(DC, 10.0)
    assert not el[:31].any()                 # z <= 3 mm: no electrode
    assert el[241].all() and np.all(phi[241] == -1)   # the last plate, downstream: -1
    # the plate (16, 17.6] mm, planes 161..176: electrode outside the 7 mm hole, +1
    assert el[165, 80, 80] == 0 and el[165, 80, 80 + 36] == 1 and phi[165, 80, 80 + 36] == 1
    assert el[165, 80, 80 + 34] == 0         # 3.4 mm from the axis: in the hole


def _laplace_direct(phi0, el):
    """Dense solve of the discrete Laplace problem: free interior nodes average their 6 neighbours."""
    nz, ny, nx = phi0.shape
    idx = -np.ones(phi0.shape, int)
    free = [(k, j, i) for k in range(1, nz - 1) for j in range(1, ny - 1) for i in range(1, nx - 1) if not el[k, j, i]]
    for n, (k, j, i) in enumerate(free):
        idx[k, j, i] = n
    A = np.zeros((len(free), len(free)))
    b = np.zeros(len(free))
    for n, (k, j, i) in enumerate(free):
        A[n, n] = 6.0
        for dk, dj, di in ((0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0), (-1, 0, 0), (1, 0, 0)):
            m = idx[k + dk, j + dj, i + di]
            if m >= 0:
                A[n, m] -= 1.0
            else:
                b[n] += float(phi0[k + dk, j + dj, i + di])
    x = np.linalg.solve(A, b)
    out = phi0.astype(np.float64).copy()
    for n, (k, j, i) in enumerate(free):
        out[k, j, i] = x[n]
    return out


def _random_problem(seed, shape=(7, 6, 8)):
    rng = np.random.default_rng(seed)
    el = (rng.random(shape) < 0.15).astype(np.uint8)
    phi = np.where(el == 1, rng.uniform(-1, 1, shape), 0.0).astype(np.float32)
    # faces keep their start value: give them one too
    for s in (np.s_[0], np.s_[-1]):
        phi[s] = rng.uniform(-1, 1, phi[s].shape)
        phi[:, s] = rng.uniform(-1, 1, phi[:, s].shape)
        phi[:, :, s] = rng.uniform(-1, 1, phi[:, :, s].shape)
    return phi, el


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_fixed_point_is_the_discrete_laplace_solution(oracle_mod, seed):
    phi0, el = _random_problem(seed)
    check = (3, 3, 3)
    el[3, 3, 3] = 0
    p, nit, dif, _ = oracle_mod.legacy_sweeps(phi0, el, check, 0.0, 4000)
    want = _laplace_direct(phi0, el)
    assert np.max(np.abs(p - want)) < 2e-6
    assert np.array_equal(p[el == 1], phi0[el == 1])      # electrodes never move
    assert np.array_equal(p[0], phi0[0]) and np.array_equal(p[:, :, -1], phi0[:, :, -1])  # nor the faces


def _sweep_in_order(phi, el, order, backward=()):
    """One in-place sweep visiting interior nodes in the given loop order (outermost first), the axes in
    `backward` from high to low index."""
    p = phi.copy()
    nz, ny, nx = p.shape
    rng = {"k": range(1, nz - 1), "j": range(1, ny - 1), "i": range(1, nx - 1)}
    for ax in backward:
        rng[ax] = rng[ax][::-1]
    a, b, c = order
    for u in rng[a]:
        for v in rng[b]:
            for w in rng[c]:
                ix = dict(zip(order, (u, v, w)))
                k, j, i = ix["k"], ix["j"], ix["i"]
                if el[k, j, i]:
                    continue
                s = np.float32(p[k, j, i - 1]) + np.float32(p[k, j, i + 1])
                s = np.float32(s + p[k, j - 1, i])
                s = np.float32(s + p[k, j + 1, i])
                s = np.float32(s + p[k - 1, j, i])
                s = np.float32(s + p[k + 1, j, i])
                p[k, j, i] = np.float32(s / np.float32(6.0))
    return p


def test_one_sweep_is_the_sketchs_loop_order(oracle_mod):
    phi0, el = _random_problem(7, (6, 7, 9))
    got = phi0.copy()
    oracle_mod.legacy_sweep(got, el)
    # the sketch: k (z) outer, i (x), j (y) inner
    assert np.array_equal(got, _sweep_in_order(phi0, el, ("k", "i", "j")))
    # every order that visits the three lower neighbours of a node before it gives the same sweep (what lets
    # the GPU run it as wavefronts), e.g. j before i
    assert np.array_equal(got, _sweep_in_order(phi0, el, ("k", "j", "i")))
    # an axis run backwards, or a Jacobi step, gives other values: the test sees the direction
    for ax in ("i", "j", "k"):
        assert not np.array_equal(got, _sweep_in_order(phi0, el, ("k", "i", "j"), backward=(ax,)))
    jac = phi0.copy()
    inner = np.s_[1:-1, 1:-1, 1:-1]
    avg = (phi0[1:-1, 1:-1, :-2] + phi0[1:-1, 1:-1, 2:] + phi0[1:-1, :-2, 1:-1] + phi0[1:-1, 2:, 1:-1]
           + phi0[:-2, 1:-1, 1:-1] + phi0[2:, 1:-1, 1:-1]) / np.float32(6)
    jac[inner] = np.where(el[inner] == 1, phi0[inner], avg)
    assert not np.array_equal(got, jac)


def test_stop_rule_arithmetic(oracle_mod):
    # dif = |100 ((prev + 1e-10) - now) / (prev + 1e-10)| in float
    assert oracle_mod.legacy_dif(1.0, 0.5) == 50.0
    assert oracle_mod.legacy_dif(2.0, 2.0) == 0.0            # 2 + 1e-10 rounds to 2 in float
    assert oracle_mod.legacy_dif(0.0, 0.0) == pytest.approx(100.0, rel=1e-6)   # 1e-10 / 1e-10
    assert oracle_mod.legacy_dif(-0.5, -0.25) == 50.0


def test_stop_rule_loop(oracle_mod):
    phi0, el = _random_problem(11)
    # a check point on a face never changes; at 0 the change stays 100 % and the loop runs to the cap
    phi0[0, 2, 2] = 0.0
    _, nit, dif, _ = oracle_mod.legacy_sweeps(phi0, el, (2, 2, 0), 0.001, 37)
    assert nit == 37 and dif == pytest.approx(100.0, rel=1e-6)
    # at a nonzero constant the first sweep changes it from the start value 1, the second by 0
    phi0[0, 2, 2] = 0.25
    _, nit, dif, _ = oracle_mod.legacy_sweeps(phi0, el, (2, 2, 0), 0.001, 37)
    assert nit == 2 and dif == 0.0
    # max_iter = 0: no sweep
    p, nit, _, _ = oracle_mod.legacy_sweeps(phi0, el, (3, 3, 3), 0.001, 0)
    assert nit == 0 and np.array_equal(p, phi0)
    # a loose threshold stops early, at the first sweep whose change is at most it
    el[3, 3, 3] = 0
    _, nit, dif, trace = oracle_mod.legacy_sweeps(phi0, el, (3, 3, 3), 5.0, 1000)
    difs = [oracle_mod.legacy_dif(a, b) for a, b in zip([1.0] + list(trace[:-1]), trace)]
    assert difs[-1] == dif <= 5.0 and all(d > 5.0 for d in difs[:-1]) and nit == len(trace)


def test_full_size_fixture_is_consistent(oracle_mod):
    ref = json.load(open(os.path.join(HERE, "golden", "legacy_potentials.json")))
    assert (ref["scl"], ref["precision_pct"], ref["max_iter"]) == (10.0, 0.001, 6000)
    for b in ref["bases"]:
        assert b["dims"] == [160, 160, 242] and b["check"] == [90, 85, 140]
        assert 0 < b["sweeps"] <= 6000
        # the loop stopped by the rule (or the cap)
        assert b["dif_pct"] <= 0.001 or b["sweeps"] == 6000
