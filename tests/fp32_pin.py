"""Helpers of the FP32 per-ion pins (tests/test_gpu_fp32_pin.py, tests/test_fp32_pin_cpu.py).

* smooth_grid(): a grid whose three bases are sums of plane waves with wavelengths of 8 to 24 cells, so every
  coefficient of a cell's trilinear polynomial (including the cross terms tx ty, ty tz, tx tz, tx ty tz) is far
  from zero, unlike the ideal quadrupole whose RF and AC fields are affine.
* field_bound(): the FP32 error bound of the batched path's field at a point. The cell record holds the 8
  polynomial coefficients c_k of each component, each rounded once from FP64 (relative error u = 2^-24); the
  worker evaluates f = c0 + c1 tx + ... + c7 tx ty tz with 4 products and 7 fused multiply-adds in FP32, from cell
  coordinates rounded to FP32. Each coefficient is a signed sum of at most 8 corner values, so
  sum_k |c_k| <= 27 max|corner| and every partial sum is bounded by it; the 8 coefficient roundings, 4 products,
  7 FMAs and 3 coordinate roundings contribute at most 22 u relative to that sum, and the drive combination
  a_ac E_ac + (s_dc E_dc + a_rf E_rf) two more. The bound is 32 u sum_b |a_b| 27 max|corner_b|.
* drive_response(): the sensitivity of each ion's end state to the field's magnitude, from the oracle itself:
  the oracle is rerun with the amplitude of each basis (V_rf, V_exc, dc_scale) raised by kappa = 2^-18 = 64 u,
  and |state(kappa) - state| is summed over the three. The FP32 path differs from the FP64 path by per-step field
  errors of at most a few u (field_bound) and by the FP32 rounding of the velocity (u |v| per step, below u times
  the per-step kick for these fields); the same tangent map carries both to the end state, so their effect stays
  below the response to a systematic 64 u error of every basis.
* Emulation (CPU only, test infrastructure): record build, polynomial evaluation and the collision-free step of
  the batched path restated in numpy float32, used to show on the CPU that the bounds fail when one coefficient
  of the record is wrong.
"""
from __future__ import annotations

import numpy as np

from gridutil import make_grid

U = 2.0 ** -24
KAPPA = 2.0 ** -18


def smooth_grid(nx=96, ny=96, nz=160, h=1.0e-4, seed=5, amp=(100.0, 100.0, 100.0)):
    """Node-major grid of smooth random fields (E per volt) for the three bases, no electrodes, return and exit
    planes outside the box. amp = (RF, DC, AC) field amplitudes [V/m per volt]."""
    rng = np.random.default_rng(seed)
    L = np.array([nx - 1, ny - 1, nz - 1]) * h

    def field(a):
        waves = []
        for _ in range(4):
            lam = rng.uniform(8, 24) * h
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            waves.append((2 * np.pi / lam * d, rng.uniform(0, 2 * np.pi), rng.normal(size=3) * a / 2))

        def fn(X, Y, Z):
            out = [np.zeros_like(X) for _ in range(3)]
            for k, ph, c in waves:
                s = np.cos(k[0] * X + k[1] * Y + k[2] * Z + ph)
                for d in range(3):
                    out[d] = out[d] + c[d] * s
            return out
        return fn

    g = make_grid(nx, ny, nz, h, rf=field(amp[0]), dc1=field(amp[1]), ac=field(amp[2]),
                  axis=(L[0] / 2, L[1] / 2), z_inject=L[2] / 2)
    return g


def smooth_jobs(g, n, half):
    """Two jobs of n ions at the centre of the smooth grid: a 0.3 mm spot at 30 K and drives of 5-6 V, so that
    256 steps keep every ion inside the box (planes outside it: every ion ends at the step cap)."""
    import mqhip
    ph = mqhip.Physics.correct(0.0).copy(dt=1e-7, max_steps=256, injection=mqhip.INJ_GAUSSIAN, inj_sigma_mm=0.3,
                                         temp_k=30.0, legacy_force=1 if half else 0)
    jobs = [mqhip.job(499.0, 1.0, 0.0, 5.0, 500e3, 0.4, 100e3, n, 91),
            mqhip.job(640.0, 2.0, 0.0, 6.0, 450e3, 0.2, 80e3, n, 92, 0.3)]
    return jobs, ph


def corners(F, nx, ny, nz, h, x, y, z):
    """The 8 corner values (n, 8, 3) of the cell holding each point (cell index clamped as the kernels do) and
    the cell coordinates t (n, 3)."""
    g = np.stack([x / h, y / h, z / h], 1)
    n = np.array([nx, ny, nz])
    g = np.clip(g, 0.0, n - 1.0)
    c = np.minimum(g.astype(np.int64), n - 2)
    t = g - c
    F = F.reshape(nz, ny, nx, 3)
    out = np.empty((len(x), 8, 3))
    for q in range(8):
        dx, dy, dz = q & 1, (q >> 1) & 1, (q >> 2) & 1
        out[:, q] = F[c[:, 2] + dz, c[:, 1] + dy, c[:, 0] + dx]
    return out, t, c


TINY = 1e-30  # V/m per unit amplitude: field values below FP32's normal range round to 0 or a subnormal


def field_bound(grid, dc, x, y, z, a_rf, s_dc, a_ac):
    """32 u sum_b |a_b| 27 max|corner_b| per point (module docstring), plus sum_b |a_b| TINY for the values that
    FP32 cannot hold at all (the tails of a Gaussian barrier fall below 1e-38)."""
    nx, ny, nz, h = grid["nx"], grid["ny"], grid["nz"], grid["h"]
    dcc = combined_dc(grid, dc)
    tot = np.zeros(len(x))
    for F, a in ((grid["rf"], a_rf), (dcc, s_dc), (grid["ac"], a_ac)):
        cr, _, _ = corners(F, nx, ny, nz, h, x, y, z)
        tot += abs(a) * (27.0 * 32.0 * U * np.abs(cr).max(axis=(1, 2)) + TINY)
    return tot


def combined_dc(grid, dc):
    out = np.zeros(3 * grid["nx"] * grid["ny"] * grid["nz"])
    for b in range(4):
        if grid["dc"][b] is not None and dc[b] != 0.0:
            out = out + dc[b] * grid["dc"][b]
    return out


def drive_response(oracle_mod, grid, jobs, dc, ph, base, nthreads):
    """Sum over the three bases of |end state(amplitude x (1 + kappa)) - end state| per ion: position [m]
    (max over components) and velocity [m/s]."""
    import mqhip
    dx = np.zeros(len(base["x"]))
    dv = np.zeros(len(base["x"]))
    for which in ("v_rf", "v_ac", "dc_scale"):
        js = []
        for j in jobs:
            j2 = mqhip.job(j.sp.mass, j.sp.z, j.sp.ccs_n2, j.drive.v_rf, j.drive.f_rf, j.drive.v_ac, j.drive.f_ac,
                           j.n, j.seed, j.dc_scale, j.pressure_pa)
            if which == "dc_scale":
                j2.dc_scale = j.dc_scale * (1.0 + KAPPA)
            else:
                setattr(j2.drive, which, getattr(j.drive, which) * (1.0 + KAPPA))
            js.append(j2)
        r = oracle_mod.run_batch(grid, js, dc, ph, nthreads)
        dx += np.max([np.abs(r[k] - base[k]) for k in ("x", "y", "z")], axis=0)
        dv += np.max([np.abs(r[k] - base[k]) for k in ("vx", "vy", "vz")], axis=0)
    return dx, dv


def state_errors(got, want):
    ex = np.max([np.abs(got[k] - want[k]) for k in ("x", "y", "z")], axis=0)
    ev = np.max([np.abs(got[k] - want[k]) for k in ("vx", "vy", "vz")], axis=0)
    return ex, ev


# ------------------------------------------------------------------ CPU emulation (float32)

def fma32(a, b, c):
    """fmaf on float32 arrays: the product of two float32 values is exact in float64."""
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(np.float32)


def record_coeffs(cr, wrong=None):
    """FP32 trilinear coefficients c0..c7 (n, 8, 3) from FP64 corner values (n, 8, 3), corner order
    dx + 2 dy + 4 dz; wrong = index of one coefficient to compute wrongly (a planted mistake)."""
    f000, f100, f010, f110, f001, f101, f011, f111 = (cr[:, q] for q in range(8))
    c = [f000, f100 - f000, f010 - f000, f001 - f000, (f110 - f100) - (f010 - f000), (f011 - f010) - (f001 - f000),
         (f101 - f100) - (f001 - f000), ((f111 - f110) - (f101 - f100)) - ((f011 - f010) - (f001 - f000))]
    if wrong == 7:
        c[7] = ((f111 - f110) - (f101 - f100)) - ((f011 - f010) + (f001 - f000))
    elif wrong == 4:
        c[4] = (f110 - f100) - (f010 + f000)
    return np.stack(c, 1).astype(np.float32)


def poly32(c, t):
    tx, ty, tz = (t[:, d].astype(np.float32)[:, None] for d in range(3))
    pxy, pyz, pxz = tx * ty, ty * tz, tx * tz
    pxyz = pxy * tz
    f = c[:, 0]
    for k, m in ((1, tx), (2, ty), (3, tz), (4, pxy), (5, pyz), (6, pxz), (7, pxyz)):
        f = fma32(np.broadcast_to(m, f.shape), c[:, k], f)
    return f


def emulated_field(grid, dc, x, y, z, a_rf, s_dc, a_ac, wrong=None, nearest=False, dcc=None):
    """The batched path's field restated in float32: (n, 3). dcc: the combined DC table, if already built."""
    nx, ny, nz, h = grid["nx"], grid["ny"], grid["nz"], grid["h"]
    if dcc is None:
        dcc = combined_dc(grid, dc)
    e = []
    for F in (grid["rf"], dcc, grid["ac"]):
        cr, t, _ = corners(F, nx, ny, nz, h, x, y, z)
        if nearest:
            t = np.zeros_like(t)
        e.append(poly32(record_coeffs(cr, wrong), t))
    er, ed, ea = e
    a_rf, s_dc, a_ac = (np.broadcast_to(np.asarray(v, np.float32), (len(x),))[:, None] for v in (a_rf, s_dc, a_ac))
    return fma32(np.broadcast_to(a_ac, ea.shape), ea, fma32(np.broadcast_to(s_dc, ed.shape), ed, a_rf * er))


def emulated_run(grid, dc, ions, jp, nsteps, wrong=None, nearest=False):
    """Collision-free steps of the batched path in float32 (velocity, field) and float64 (position, phase):
    v += kq E; x += v dt. ions: dict of float64 start arrays (x..vz, phi_ac0); jp: per-ion dict of kq, v_rf,
    w_rf, v_ac, w_ac, s_dc, dt. Returns the end state after nsteps (no termination)."""
    x, y, z = ions["x"].copy(), ions["y"].copy(), ions["z"].copy()
    dcc = combined_dc(grid, dc)
    v = [ions[k].astype(np.float32) for k in ("vx", "vy", "vz")]
    kq = jp["kq"].astype(np.float32)
    for s in range(nsteps):
        pr = jp["w_rf"] * s
        pr = pr - np.floor(pr)
        pa = jp["w_ac"] * s + ions["phi_ac0"]
        pa = pa - np.floor(pa)
        a_rf = (jp["v_rf"].astype(np.float32) * np.sin(2 * np.pi * pr.astype(np.float32)).astype(np.float32))
        a_ac = (jp["v_ac"].astype(np.float32) * np.sin(2 * np.pi * pa.astype(np.float32)).astype(np.float32))
        e = emulated_field(grid, dc, x, y, z, a_rf, jp["s_dc"], a_ac, wrong, nearest, dcc)
        for d in range(3):
            v[d] = fma32(kq, e[:, d], v[d])
        x = x + v[0].astype(np.float64) * jp["dt"]
        y = y + v[1].astype(np.float64) * jp["dt"]
        z = z + v[2].astype(np.float64) * jp["dt"]
    return {"x": x, "y": y, "z": z, "vx": v[0].astype(np.float64), "vy": v[1].astype(np.float64),
            "vz": v[2].astype(np.float64)}
