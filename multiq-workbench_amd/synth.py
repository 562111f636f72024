"""Synthetic basis-field grids and workload generators.

The reference builds its grids with a multigrid Laplace solve of the channel
geometry (channel.rs:143-217, simulation/fieldsolve). These synthetic grids
stand in for them: an ideal quadrupole whose RF potential has the reference's
orientation phi = -2 X Y / r0^2 (channel.rs:510-522, r0_profile), a uniform
AC dipole across the left/right rod pairs (ac_sign, channel.rs:105), one DC
basis (the plate slot, index 1) shaped as a Gaussian axial barrier with its
Laplace-consistent radial term, centred where the reference's gate plate sits
(16.0-17.6 mm, before the exit plane at 18.6 mm: Geom::layers, channel.rs:50-54),
and a conductor mask on the hyperbolic rod
surfaces |X Y| >= r0^2 / 2 plus a cylindrical wall at 1.1 r0. All grids are
node-major (i fastest), E per volt, FP64.
"""
from __future__ import annotations

import numpy as np

R0_REFERENCE = 2.765e-3   # r0 of the reference channel (channel.rs:58, Geom::default)


def ideal_channel(nx: int, ny: int, nz: int, h: float, r0: float = R0_REFERENCE, z_gate: float = 16.8e-3,
                  w_gate: float = 1.0e-3, z_exit: float = 18.6e-3, wall: float = 1.1) -> dict:
    ax = (nx // 2) * h
    ay = (ny // 2) * h
    x = np.arange(nx) * h - ax
    y = np.arange(ny) * h - ay
    z = np.arange(nz) * h
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")  # shape (nz, ny, nx): i fastest
    N = nx * ny * nz
    rf = np.empty((nz, ny, nx, 3))
    rf[..., 0] = 2.0 * Y / r0**2
    rf[..., 1] = 2.0 * X / r0**2
    rf[..., 2] = 0.0
    ac = np.zeros((nz, ny, nx, 3))
    ac[..., 0] = -1.0 / r0
    u = Z - z_gate
    g = np.exp(-(u / w_gate) ** 2)
    g1 = -2.0 * u / w_gate**2 * g
    g2 = (4.0 * u**2 / w_gate**4 - 2.0 / w_gate**2) * g
    g3 = (12.0 * u / w_gate**4 - 8.0 * u**3 / w_gate**6) * g
    r2 = X**2 + Y**2
    dc1 = np.empty((nz, ny, nx, 3))
    dc1[..., 0] = g2 * X / 2.0
    dc1[..., 1] = g2 * Y / 2.0
    dc1[..., 2] = -g1 + g3 * r2 / 4.0
    mask = ((np.abs(X * Y) >= r0**2 / 2.0) | (r2 >= (wall * r0) ** 2)).astype(np.uint8)
    return {
        "nx": nx, "ny": ny, "nz": nz, "h": h,
        "axis_x": ax, "axis_y": ay, "z_inject": 1.0e-3,
        "z_return": 0.5e-3, "z_exit": z_exit,
        "z_return_legacy": 1.0e-3, "z_exit_legacy": 23.7e-3,
        "crash_z0_legacy": 3.0e-3, "crash_z1_legacy": 15.0e-3,
        "rf": rf.reshape(3 * N), "ac": ac.reshape(3 * N),
        "dc": [None, dc1.reshape(3 * N), None, None],
        "mask": mask.reshape(N),
        "r0": r0,
    }


def c1_grid() -> dict:
    """64 x 64 x 256 nodes at 0.1 mm (the Fig-4 single-channel configuration)."""
    return ideal_channel(64, 64, 256, 1.0e-4)


def c2_grid() -> dict:
    """128 x 128 x 512 nodes at 0.05 mm (the batched mass scan)."""
    return ideal_channel(128, 128, 512, 0.5e-4)


def c3_grid() -> dict:
    """192 x 192 x 768 nodes at 1/30 mm (the operating-setting sweep)."""
    return ideal_channel(192, 192, 768, 1.0e-4 / 3.0)


def c1_jobs(mod, n: int = 1000, mz: float = 499.0, z: float = 1.0, ccs: float = 230.0, seed: int | None = None):
    """Fig-4 point: 100 V RF at 500 kHz, 0.5 V AC at 100 kHz, seed 1000 + m/z."""
    s = int(1000 + mz) if seed is None else seed
    return [mod.job(mz * z, z, ccs, 100.0, 500e3, 0.5, 100e3, n, s)]


def c2_jobs(mod, n: int = 1000, z: float = 1.0, ccs: float = 230.0):
    """Mass scan m/z 475..523 step 2 (25) x 64 RF frequencies 450..550 kHz x n ions."""
    jobs = []
    for mz in range(475, 524, 2):
        for k in range(64):
            f_rf = 450e3 + 100e3 * k / 63.0
            jobs.append(mod.job(mz * z, z, ccs, 100.0, f_rf, 0.5, 100e3, n, 1000 + mz + 100000 * k))
    return jobs


def lhs(n: int, d: int, seed: int) -> np.ndarray:
    """Latin-hypercube sample of n points in [0,1)^d (stratified, seeded)."""
    rng = np.random.default_rng(seed)
    out = np.empty((n, d))
    for k in range(d):
        out[:, k] = (rng.permutation(n) + rng.random(n)) / n
    return out


def c3_jobs(mod, n_settings: int = 4096, n: int = 256, seed: int = 3):
    """Operating-setting sweep: V_rf 50-150 V, f_exc 80-120 kHz, V_exc 0.1-1 V,
    V_dc 0-10 V (per-job DC scale of the plate basis at 1 V), N2 0.5-2 mTorr
    (per-job pressure), n ions of m/z 499, z = 1 per setting."""
    u = lhs(n_settings, 5, seed)
    jobs = []
    for i in range(n_settings):
        v_rf = 50.0 + 100.0 * u[i, 0]
        f_exc = 80e3 + 40e3 * u[i, 1]
        v_exc = 0.1 + 0.9 * u[i, 2]
        v_dc = 10.0 * u[i, 3]
        p = (0.5 + 1.5 * u[i, 4]) * 0.133322
        jobs.append(mod.job(499.0, 1.0, 230.0, v_rf, 500e3, v_exc, f_exc, n, 7000 + i, v_dc, p))
    return jobs


# ---------------------------------------------------------------- tandem batch

AMU = 1.66053906660e-27
QE = 1.602176634e-19
UBQ_MASS = 8565.0                       # ubiquitin, Da
UBQ_CHARGES = (6.0, 7.0, 8.0, 9.0)      # m/z 1428.5, 1224.6, 1071.6, 952.7
DRIVEN_PORTS = (0, 2, 6, 8)             # the four driven ports of trapsim's default tones (lib.rs:13)


def secular_frequency(mass_da: float, z: float, v_rf: float, f_rf: float, r0: float = R0_REFERENCE,
                      terms: int = 30) -> float:
    """Radial secular frequency [Hz] of an ideal quadrupole, a = 0, full kick:
    q = 4 z e V / (m r0^2 Omega^2), f = beta f_rf / 2 with beta from the
    Mathieu continued fraction (March, Quadrupole Ion Trap Mass Spectrometry)."""
    om = 2.0 * np.pi * f_rf
    q = 4.0 * z * QE * v_rf / (mass_da * AMU * r0 * r0 * om * om)
    beta = np.sqrt(q * q / 2.0)
    for _ in range(300):
        def cf(b, sgn):
            v = 0.0
            for k in range(terms, 0, -1):
                m = b + sgn * 2 * k
                v = q * q / (m * m - v)
            return v
        nb = np.sqrt(max(cf(beta, 1) + cf(beta, -1), 1e-18))
        if abs(nb - beta) < 1e-14:
            break
        beta = nb
    return beta * f_rf / 2.0


def tandem_cube(nx: int, ny: int, nz: int, h: float, seam_z: float, r0_rf: float = 6.0e-3) -> dict:
    """Splitter/combiner grid of the tandem batch (a stand-in for trapsim's
    486-post cube, geometry.rs): a confining RF quadrupole E = 2 (Y, X) / r0_rf^2
    per volt about the box axis, no excitation, and two DC bases: slot 0 a
    uniform axial push E_z = 1 / seam_z per volt (+V drives ions toward the
    port board), slot 1 a uniform transverse push E_y = 1 / L per volt, L the
    distance from the axis to the y = (ny - 1) h face where the entrance
    aperture is (+V drives ions to that face). No electrodes inside the box;
    its faces are the walls."""
    ax = (nx // 2) * h
    ay = (ny // 2) * h
    x = np.arange(nx) * h - ax
    y = np.arange(ny) * h - ay
    Z, Y, X = np.meshgrid(np.arange(nz) * h, y, x, indexing="ij")
    N = nx * ny * nz
    rf = np.empty((nz, ny, nx, 3))
    rf[..., 0] = 2.0 * Y / r0_rf**2
    rf[..., 1] = 2.0 * X / r0_rf**2
    rf[..., 2] = 0.0
    push = np.zeros((nz, ny, nx, 3))
    push[..., 2] = 1.0 / seam_z
    exit_push = np.zeros((nz, ny, nx, 3))
    exit_push[..., 1] = 1.0 / ((ny - 1) * h - ay)
    return {
        "nx": nx, "ny": ny, "nz": nz, "h": h,
        "axis_x": ax, "axis_y": ay, "z_inject": 1.0e-3,
        "z_return": 0.5e-3, "z_exit": seam_z,
        "z_return_legacy": 1.0e-3, "z_exit_legacy": seam_z,
        "crash_z0_legacy": 0.0, "crash_z1_legacy": 0.0,
        "rf": rf.reshape(3 * N), "ac": None,
        "dc": [push.reshape(3 * N), exit_push.reshape(3 * N), None, None],
        "mask": None,
    }


def tandem_setup(mod, size: str = "c5", n_per_charge: int | None = None, v_rf: float = 115.0,
                 v_push: float = 2.0, v_exit: float = 100.0, v_gate: float = 1.0, v_ac: float = 0.5,
                 pitch: float = 2.0e-3, ap_r: float = 0.8e-3, pressure_pa: float = 0.133322, max_steps: int = 30000):
    """The tandem configuration. size "c5": cube and ports 128 x 128 x 512 at
    0.05 mm, 4 x 262144 ions (1M); size "small": cube 48 x 48 x 176 at
    0.15 mm, ports 64 x 64 x 256 at 0.1 mm, for parity checks.

    Ions: ubiquitin (8565 Da) 6+ to 9+, one job per charge state, injected 1 mm
    above the splitter floor (Gaussian spot, 300 K). Ports: nine mouths on a
    3 x 3 lattice of the given pitch about the cube axis; ports 0, 2, 6, 8 are
    driven at the secular frequencies of 6+, 7+, 8+, 9+ (full kick); the
    others are idle with the gate raised by 4 V (trapsim lib.rs:196-203).
    The splitter pushes ions to the board (v_push on slot 0); the combiner
    pushes the forwarded ions to the y face (v_exit on slot 1), where the
    entrance aperture sits 1 mm below the board in line with the mouths of
    ports 2, 5 and 8, so that ions forwarded by ports 2 and 8 can be detected.
    Returns (cube, port, geom, drive, jobs, physics)."""
    if size == "c5":
        cube = tandem_cube(128, 128, 512, 0.5e-4, seam_z=24.0e-3)
        port = ideal_channel(128, 128, 512, 0.5e-4)
        n = 262144 if n_per_charge is None else n_per_charge
    elif size == "small":
        cube = tandem_cube(48, 48, 176, 1.5e-4, seam_z=24.0e-3)
        port = ideal_channel(64, 64, 256, 1.0e-4)
        n = 256 if n_per_charge is None else n_per_charge
    else:
        raise ValueError(size)
    ax, ay = cube["axis_x"], cube["axis_y"]
    mouths = [(ax + (c - 1) * pitch, ay + (r - 1) * pitch) for r in range(3) for c in range(3)]
    # the entrance aperture on the cube's y face 1 mm below the board, in line with the mouths of ports 2, 5
    # and 8, with a 2 mm tube (trapsim's l_ext)
    geom = mod.tandem_geom(24.0e-3, mouths, ap_r, (ax + pitch, 23.0e-3), 2.0e-3)
    tones = [0.0] * 9
    for p, z in zip(DRIVEN_PORTS, UBQ_CHARGES):
        tones[p] = secular_frequency(UBQ_MASS, z, v_rf, 500e3)
    drive = mod.tandem_drive(v_rf, 500e3, v_ac, tones, [v_push, 0, 0, 0], [0, v_exit, 0, 0], [0, v_gate, 0, 0],
                             [0, 4.0, 0, 0])
    ccs = {6.0: 1000.0, 7.0: 1150.0, 8.0: 1300.0, 9.0: 1450.0}
    jobs = [mod.job(UBQ_MASS, z, ccs[z], v_rf, 500e3, v_ac, 0.0, n, 6000 + int(z)) for z in UBQ_CHARGES]
    ph = mod.Physics.correct(pressure_pa).copy(injection=mod.INJ_GAUSSIAN, inj_sigma_mm=1.0, max_steps=max_steps)
    return cube, port, geom, drive, jobs, ph
