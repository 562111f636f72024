"""The oracle of the basis-field pipeline, pinned against the geometry's own
definition (channel.rs:18-119), the mathematics of the discrete Laplace
problem (fieldsolve/src/lib.rs:150-166) and closed forms of the central
difference (channel.rs:168-184). CPU only."""
import math

import numpy as np
import pytest

import mqhip

MM = 1.0


@pytest.fixture(scope="module")
def default_codes(oracle_mod):
    g = oracle_mod.geom_default()
    return g, oracle_mod.geom_classify(g)


def test_default_geometry_dimensions(oracle_mod):
    g = oracle_mod.geom_default()
    assert (g.n_xy, g.h_mm, g.de, g.coef, g.alpha, g.hole_d, g.q2_flare_bug, g.grow) == \
        (160, 0.1, 6.35, 0.323, 0.02, 7.0, 1, 0.0)
    # layers: 3 | Q1 12 | gap 1 | plate 1.6 | gap 1 | Q2 3 | gap 1 | slab 1.6 -> 24.2 mm, 242 planes
    assert oracle_mod.load().or_geom_nz(g) == 242


def test_planes_without_electrodes_and_full_planes(default_codes):
    g, c = default_codes
    assert not c[:31].any()                        # z <= 3 mm (q1_start): open
    assert np.all(c[227:242] == mqhip_const("SLAB"))  # (22.6, 24.2] mm: slab over the whole plane


def mqhip_const(name):
    return {"NONE": 0, "Q1": 1, "STEM3": 5, "PLATE": 9, "STEM5": 10, "Q2": 14, "STEM7": 18, "SLAB": 22}[name]


def test_layer_boundaries_are_inclusive_above(default_codes):
    g, c = default_codes
    # k = 150 is z = 15.0 mm = end of Q1 (z <= l2 + eps): still Q1; k = 151 is the stem section
    fam = lambda k: set(np.unique(c[k])) - {0}
    assert fam(150) == {1, 2, 3, 4}
    assert fam(151) == {5, 6, 7, 8}
    assert fam(160) == {5, 6, 7, 8} and fam(161) == {9}
    assert fam(176) == {9} and fam(177) == {10, 11, 12, 13}
    assert fam(186) == {10, 11, 12, 13} and fam(187) == {14, 15, 16, 17}
    assert fam(216) == {14, 15, 16, 17} and fam(217) == {18, 19, 20, 21}
    assert fam(226) == {18, 19, 20, 21} and fam(227) == {22}


def test_rod_cross_sections_match_their_areas(default_codes):
    g, c = default_codes
    h = g.h_mm
    for k, code, d in [(90, 1, g.de), (155, 5, g.stem_d), (200, 14, g.de), (220, 18, g.stem_d)]:
        for q in range(4):
            n = int((c[k] == code + q).sum())
            area = math.pi * (d / 2) ** 2 / h ** 2
            assert abs(n - area) < 2.5 * math.pi * d / h, (k, q, n, area)  # within a perimeter band


def test_rod_centres_follow_the_squeeze(default_codes):
    g, c = default_codes
    h = g.h_mm
    a = 80 * h
    cc = g.coef * g.de / 2 + g.de / 2  # 4.2005 mm off axis in x and y
    for k in (31, 60, 90, 120, 150):
        z = k * h
        sq = g.alpha * (15.0 - z)
        for q, (sx, sy) in enumerate([(-1, -1), (1, -1), (-1, 1), (1, 1)]):
            jj, ii = np.nonzero(c[k] == 1 + q)
            cx, cy = ii.mean() * h, jj.mean() * h
            # tilt mode 0: all four rods shifted in y by +-sq (lower pair up, upper pair down)
            assert abs(cx - (a + sx * cc)) < 0.02, (k, q, cx)
            assert abs(cy - (a + sy * cc - sy * sq)) < 0.02, (k, q, cy, sq)


def test_q2_flare_quirk_moves_rods_outward(oracle_mod):
    h = 0.1
    for bug in (1, 0):
        g = oracle_mod.geom_default(q2_flare_bug=bug)
        c = oracle_mod.geom_classify(g)
        jj, ii = np.nonzero(c[215] == 14)   # LL rod near the end of Q2, z = 21.5 mm
        a = 8.0
        cy = jj.mean() * h
        sq = 0.02 * (15.0 - 21.5) if bug else 0.0   # negative: the rods move away from the axis
        assert abs(cy - (a - 4.2005 + sq)) < 0.02


def test_plate_hole_and_growth(oracle_mod):
    for grow in (0.0, 0.4):
        g = oracle_mod.geom_default(grow=grow)
        c = oracle_mod.geom_classify(g)
        hole = int((c[170] == 0).sum())
        r = 3.5 - grow * 0.1
        assert abs(hole - math.pi * (r / 0.1) ** 2) < 2.5 * math.pi * 2 * r / 0.1
        rod = int((c[90] == 1).sum())
        rr = 3.175 + grow * 0.1
        assert abs(rod - math.pi * (rr / 0.1) ** 2) < 2.5 * math.pi * 2 * rr / 0.1
    # growth adds a ring of about 2 pi r (0.4 h) / h^2 nodes per rod cross-section
    c0 = oracle_mod.geom_classify(oracle_mod.geom_default(grow=0.0))
    c4 = oracle_mod.geom_classify(oracle_mod.geom_default(grow=0.4))
    d = int((c4[90] == 1).sum()) - int((c0[90] == 1).sum())
    assert abs(d - 2 * math.pi * 31.75 * 0.4) < 60


def test_axis_planes_belong_to_no_rod(oracle_mod):
    g = oracle_mod.geom_default()
    # x == axis or y == axis: no quadrant, never a rod node
    for k in (50, 155, 200):
        for t in range(0, 160, 7):
            assert oracle_mod.classify_node(g, 80, t, k) == 0
            assert oracle_mod.classify_node(g, t, 80, k) == 0


def test_tilt_modes_and_misalignment(oracle_mod):
    h = 0.1
    k = 40
    sq = 0.02 * (15.0 - k * h)
    # mode 1: only the upper pair converges in x
    c = oracle_mod.geom_classify(oracle_mod.geom_default(tilt_mode=1))
    for q, (ex, ey) in enumerate([(8 - 4.2005, 8 - 4.2005), (8 + 4.2005, 8 - 4.2005),
                                  (8 - 4.2005 + sq, 8 + 4.2005), (8 + 4.2005 - sq, 8 + 4.2005)]):
        jj, ii = np.nonzero(c[k] == 1 + q)
        assert abs(ii.mean() * h - ex) < 0.02 and abs(jj.mean() * h - ey) < 0.02
    # mode 2: symmetric radial squeeze by sq
    c = oracle_mod.geom_classify(oracle_mod.geom_default(tilt_mode=2))
    jj, ii = np.nonzero(c[k] == 1)
    assert abs(ii.mean() * h - (8 - 4.2005 + sq / math.sqrt(2))) < 0.02
    # misalignment of the LL rod: shift + tilt about the Q1 midpoint (z = 9 mm)
    g = oracle_mod.geom_default(grow=0.0)
    g.mis[0][0], g.mis[0][2] = 0.3, 0.01
    c = oracle_mod.geom_classify(g)
    for kk in (40, 140):
        jj, ii = np.nonzero(c[kk] == 1)
        assert abs(ii.mean() * h - (8 - 4.2005 + 0.3 + 0.01 * (kk * h - 9.0))) < 0.02
    # mis applies to Q1 only
    jj, ii = np.nonzero(c[200] == 14)
    assert abs(ii.mean() * h - (8 - 4.2005)) < 0.02


def test_basis_values(oracle_mod):
    bv = lambda b, code, ac=0: oracle_mod.load().or_basis_value(b, code, ac)
    RF, AC, DQ1, DPL, DDN, DBOX = range(6)
    # RF: LL, RU -1; RL, LU +1 on every rod family; plate and slab 0
    for base in (1, 5, 10, 14, 18):
        assert [bv(RF, base + q) for q in range(4)] == [-1, 1, 1, -1]
    assert bv(RF, 9) == 0 and bv(RF, 22) == 0 and bv(RF, 0) == 0
    # AC: left -1, right +1 on Q1 and the stems next to it; ac_mode 1 drops the upper pair
    for base in (1, 5):
        assert [bv(AC, base + q) for q in range(4)] == [-1, 1, -1, 1]
        assert [bv(AC, base + q, 1) for q in range(4)] == [-1, 1, 0, 0]
    for code in (9, 10, 14, 18, 22):
        assert bv(AC, code) == 0
    assert [bv(DQ1, c) for c in (1, 5, 8, 9, 10, 22)] == [1, 1, 1, 0, 0, 0]
    assert [bv(DPL, c) for c in (1, 8, 9, 10, 22)] == [0, 0, 1, 0, 0]
    assert [bv(DDN, c) for c in (1, 8, 9, 10, 13, 14, 18, 21, 22)] == [0, 0, 0, 1, 1, 1, 1, 1, 1]
    assert all(bv(DBOX, c) == 0 for c in range(23))


# ---------------------------------------------------------------- the discrete Laplace problem

def _resid(phi, free):
    r = np.zeros_like(phi)
    r[1:-1, 1:-1, 1:-1] = (phi[2:, 1:-1, 1:-1] + phi[:-2, 1:-1, 1:-1] + phi[1:-1, 2:, 1:-1] + phi[1:-1, :-2, 1:-1]
                           + phi[1:-1, 1:-1, 2:] + phi[1:-1, 1:-1, :-2] - 6 * phi[1:-1, 1:-1, 1:-1])
    return np.abs(r[free]).max() if free.any() else 0.0


def _free(mask):
    f = mask == 0
    f[0], f[-1], f[:, 0], f[:, -1], f[:, :, 0], f[:, :, -1] = False, False, False, False, False, False
    return f


def _blob_problem(shape=(21, 17, 19), seed=0):
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    mask = np.zeros(shape, np.uint8)
    fixed = np.zeros(shape)
    for _ in range(4):
        c = [rng.integers(3, s - 3) for s in shape]
        r = rng.integers(1, 3)
        sl = tuple(slice(ci - r, ci + r + 1) for ci in c)
        mask[sl] = 1
        fixed[sl] = rng.uniform(-2, 2)
    return mask, fixed


@pytest.mark.parametrize("poly", ["x2-y2", "xy", "x2+y2-2z2", "xyz"])
def test_sor_reproduces_discretely_harmonic_polynomials(oracle_mod, poly):
    # Quadratic harmonics (and xyz) have a zero 7-point Laplacian on a uniform
    # grid, so with the polynomial as Dirichlet data on the faces and on some
    # interior electrode nodes the discrete solution IS the polynomial.
    nz, ny, nx = 13, 15, 17
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x, y, z = (i - 7.0) / 8, (j - 6.0) / 8, (k - 5.0) / 8
    f = {"x2-y2": x * x - y * y, "xy": x * y, "x2+y2-2z2": x * x + y * y - 2 * z * z, "xyz": x * y * z}[poly]
    mask, _ = _blob_problem((nz, ny, nx), 3)
    mask[0], mask[-1], mask[:, 0], mask[:, -1], mask[:, :, 0], mask[:, :, -1] = 1, 1, 1, 1, 1, 1
    phi, st = oracle_mod.laplace_sor(mask, f, 0.0, 1e-13)
    assert st["converged"]
    assert np.abs(phi - f).max() < 1e-11


def test_sor_residual_linearity_and_maximum_principle(oracle_mod):
    mask, fixed = _blob_problem()
    free = _free(mask)
    tol = 1e-10
    a, st = oracle_mod.laplace_sor(mask, fixed, 0.5, tol)
    scale = max(np.abs(fixed[mask == 1]).max(), 0.5)
    assert st["converged"] and _resid(a, free) / scale <= tol * 1.0001
    # fixed nodes and faces hold their values
    assert np.array_equal(a[mask == 1], fixed[mask == 1])
    face = ~free & (mask == 0)
    assert np.all(a[face] == 0.5)
    # maximum principle: free values lie within the range of the boundary data
    bd = np.concatenate([fixed[mask == 1], [0.5]])
    assert a[free].min() >= bd.min() - 1e-9 and a[free].max() <= bd.max() + 1e-9
    # linearity: solution(2 f1 - f2) = 2 solution(f1) - solution(f2)
    m2, f2 = _blob_problem(seed=1)
    mask_u = mask | m2
    f1 = np.where(mask_u == 1, fixed, 0)
    g2 = np.where(mask_u == 1, np.random.default_rng(5).uniform(-1, 1, fixed.shape), 0)
    s1, _ = oracle_mod.laplace_sor(mask_u, f1, 0.0, 1e-12)
    s2, _ = oracle_mod.laplace_sor(mask_u, g2, 0.0, 1e-12)
    s3, _ = oracle_mod.laplace_sor(mask_u, 2 * f1 - g2, 0.0, 1e-12)
    assert np.abs(s3 - (2 * s1 - s2)).max() < 1e-9


def test_sor_outer_only_gives_constant(oracle_mod):
    # no electrodes, faces at 1 V: the solution is 1 everywhere
    mask = np.zeros((9, 10, 11), np.uint8)
    phi, st = oracle_mod.laplace_sor(mask, np.zeros(mask.shape), 1.0, 1e-13)
    assert st["converged"] and np.abs(phi - 1).max() < 1e-12


def test_sor_symmetric_problem_gives_symmetric_solution(oracle_mod):
    nz, ny, nx = 15, 13, 13
    mask = np.zeros((nz, ny, nx), np.uint8)
    fixed = np.zeros(mask.shape)
    mask[7, 3:5, 3:5] = 1
    fixed[7, 3:5, 3:5] = 1.0
    mask[7, 8:10, 8:10] = 1       # mirror image through the centre in x and y
    fixed[7, 8:10, 8:10] = 1.0
    phi, _ = oracle_mod.laplace_sor(mask, fixed, 0.0, 1e-13)
    assert np.abs(phi - phi[:, ::-1, ::-1]).max() < 1e-11


# ---------------------------------------------------------------- E = -grad(phi)

def test_efield_exact_for_linear_and_quadratic_potentials(oracle_mod):
    nz, ny, nx = 6, 7, 8
    h = 2.5e-4
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x, y, z = i * h, j * h, k * h
    lin = 3.0 * x - 2.0 * y + 0.5 * z + 1.0
    e = oracle_mod.efield(lin, h)
    assert np.allclose(e[..., 0], -3.0, rtol=1e-12, atol=0) and np.allclose(e[..., 1], 2.0, rtol=1e-12)
    assert np.allclose(e[..., 2], -0.5, rtol=1e-12)
    q = x * x - y * z
    e = oracle_mod.efield(q, h)
    inner = (slice(1, -1),) * 3
    assert np.allclose(e[..., 0][inner], (-2 * x)[inner], rtol=1e-10, atol=1e-15)
    assert np.allclose(e[..., 1][inner], z[inner], rtol=1e-10, atol=1e-15)
    assert np.allclose(e[..., 2][inner], y[inner], rtol=1e-10, atol=1e-15)
    # at the x faces the difference is one-sided over one spacing
    assert np.allclose(e[:, :, 0, 0], -(q[:, :, 1] - q[:, :, 0]) / h, rtol=1e-12)
    assert np.allclose(e[:, :, -1, 0], -(q[:, :, -1] - q[:, :, -2]) / h, rtol=1e-12)


def test_geom_struct_layout_matches_the_header(oracle_mod):
    # the oracle fills the same struct the product reads: same default values through both
    a = oracle_mod.geom_default()
    b = mqhip.Geom.default()
    assert bytes(a) == bytes(b)
