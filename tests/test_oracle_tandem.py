"""The tandem oracle (trapsim Model::advance / enter / leave,
simulation/trapsim/src/lib.rs:176-385), pinned by closed forms and flow
invariants. CPU only.

A single ion at rest on the cube axis (0 K, zero spot size, no gas, no RF) in
uniform fields moves in closed form under kick-drift:
    v_n = n a dt,   z_n = z_0 + a dt^2 n (n + 1) / 2.
That fixes the step and the position of every crossing, so each counter can
be checked against an independent prediction. Random multi-ion runs are
checked against the flow balance every counter set must satisfy.
"""
import math

import numpy as np
import pytest

import mqhip
import synth

AMU, QE = 1.66053906660e-27, 1.602176634e-19
H = 0.25e-3


def uniform_grid(nx, ny, nz, h, e_bases, z_ret=0.5e-3, z_exit=None):
    """Grid dict whose RF and AC are zero and whose DC basis b is the uniform
    field e_bases[b] (a 3-vector per volt) or None."""
    N = nx * ny * nz
    zero = np.zeros(3 * N)
    dc = [None if e is None else np.tile(np.asarray(e, float), N) for e in e_bases]
    return {"nx": nx, "ny": ny, "nz": nz, "h": h, "axis_x": (nx // 2) * h, "axis_y": (ny // 2) * h,
            "z_inject": 1.0e-3, "z_return": z_ret, "z_exit": z_exit if z_exit else (nz - 3) * h,
            "z_return_legacy": z_ret, "z_exit_legacy": 0.0, "crash_z0_legacy": 0.0, "crash_z1_legacy": 0.0,
            "rf": zero, "ac": zero, "dc": dc, "mask": None}


def setup(cube_e, port_e, pitch=1.5e-3, ap_r=0.5e-3, seam=20.0e-3, entrance=None, l_ext=1.0e-3):
    """Cube 6 x 6 x 25 mm and port 5 x 5 x 20 mm of uniform fields; nine mouths on a 3 x 3 lattice about the
    cube axis; the entrance aperture on the cube's y = 6 mm face, by default at x = 3 mm, z = 1 mm (the
    injection height)."""
    cube = uniform_grid(25, 25, 101, H, cube_e)                 # 6 x 6 x 25 mm
    port = uniform_grid(21, 21, 81, H, port_e, z_exit=18.0e-3)  # 5 x 5 x 20 mm
    ax, ay = cube["axis_x"], cube["axis_y"]
    mouths = [(ax + (c - 1) * pitch, ay + (r - 1) * pitch) for r in range(3) for c in range(3)]
    geom = mqhip.tandem_geom(seam, mouths, ap_r, entrance or (ax, 1.0e-3), l_ext)
    return cube, port, geom


def rest_physics(max_steps, dt=4e-8):
    return mqhip.Physics.correct(0.0).copy(dt=dt, temp_k=1e-30, injection=mqhip.INJ_GAUSSIAN, inj_sigma_mm=0.0,
                                          max_steps=max_steps)


def drive(dc_split=(1, 0, 0, 0), dc_combine=(1, 0, 0, 0), dc_port=(1, 0, 0, 0), idle=(0, 0, 0, 0), tones=None):
    return mqhip.tandem_drive(0.0, 500e3, 0.0, tones or [0.0] * 9, dc_split, dc_combine, dc_port, idle)


def one_ion(z=7.0, mass=8565.0):
    return [mqhip.job(mass, z, 0.0, 0.0, 500e3, 0.0, 0.0, 1, 11)]


def accel(z, mass, e):
    return z * QE / (mass * AMU) * e


def crossing_step(z0, a, dt, plane):
    """Smallest n with z0 + a dt^2 n (n + 1) / 2 > plane."""
    c = 2 * (plane - z0) / (a * dt * dt)
    n = int(math.floor((-1 + math.sqrt(1 + 4 * c)) / 2))
    while z0 + a * dt * dt * n * (n + 1) / 2 <= plane:
        n += 1
    return n


def run(oracle_mod, cube, port, geom, jobs, d, ph):
    return oracle_mod.tandem_run(cube, port, geom, jobs, d, ph, 1)


def test_entry_step_and_position(oracle_mod):
    ez = 40.0  # V/m per volt, +z
    cube, port, geom = setup([(0, 0, ez), None, None, None], [(0, 0, ez), None, None, None])
    dt = 4e-8
    a = accel(7.0, 8565.0, ez)
    n_s = crossing_step(1.0e-3, a, dt, 20.0e-3)
    z_n = 1.0e-3 + a * dt * dt * n_s * (n_s + 1) / 2
    assert z_n - 20.0e-3 > 1e-9 and 20.0e-3 - (z_n - a * dt * dt * n_s) > 1e-9  # not borderline
    # one step before the crossing: still in the splitter
    c, t, s = run(oracle_mod, cube, port, geom, one_ion(), drive(), rest_physics(n_s - 1))
    assert s["region"][0] == 0 and s["status"][0] == 5 and c[4]["entered"] == 0
    assert t["split"] == 1 and t["n"] == 1
    # at the crossing step: in port 4 (the central mouth), at z_ret + depth, same velocity
    c, t, s = run(oracle_mod, cube, port, geom, one_ion(), drive(), rest_physics(n_s))
    assert s["region"][0] == 1 + 4 and s["last"][0] == 255 and s["status"][0] == 5
    assert c[4]["entered"] == 1 and sum(p["entered"] for p in c) == 1 and t["channel"] == 1
    assert s["z"][0] == pytest.approx(0.5e-3 + (z_n - 20.0e-3), rel=1e-9, abs=1e-15)
    assert s["x"][0] == pytest.approx(port["axis_x"], abs=1e-15)
    assert s["vz"][0] == pytest.approx(a * dt * n_s, rel=1e-12)
    assert s["steps"][0] == n_s


def test_lateral_drift_selects_the_mouth(oracle_mod):
    # a transverse push moves the ion one pitch in +x by the seam: mouth 5 (row 1, col 2)
    ez = 40.0
    dt = 4e-8
    a = accel(7.0, 8565.0, ez)
    n_s = crossing_step(1.0e-3, a, dt, 20.0e-3)
    k = a * dt * dt * n_s * (n_s + 1) / 2
    for frac_pitch, port_hit in [(1.0, 5), (-1.0, 3)]:
        ex = ez * frac_pitch * 1.5e-3 / k
        cube, port, geom = setup([(ex, 0, ez), None, None, None], [(0, 0, ez), None, None, None])
        c, t, s = run(oracle_mod, cube, port, geom, one_ion(), drive(), rest_physics(n_s + 2))
        if True:
            assert s["region"][0] == 1 + port_hit and c[port_hit]["entered"] == 1
            # the drift puts the ion on the mouth centre at the seam, so it lands on the
            # port axis and moves on only by its transverse speed over the two extra steps
            vx = accel(7.0, 8565.0, ex) * dt * n_s
            assert abs(s["x"][0] - port["axis_x"]) < 3 * abs(vx) * dt + 1e-9


def test_above_the_board_between_mouths(oracle_mod):
    """Above the seam outside every aperture the ion keeps moving (lib.rs:213-229): half a pitch off, it is lost
    only when it passes the top of the cube; 0.6 pitch off, its lateral drift carries it into mouth 5's
    aperture above the seam and it enters port 5 there. Kick-drift from rest in uniform fields:
    x_n - x_0 = (E_x / E_z) (z_n - z_0)."""
    ez, dt = 40.0, 4e-8
    a = accel(7.0, 8565.0, ez)
    n_s = crossing_step(1.0e-3, a, dt, 20.0e-3)
    k = a * dt * dt * n_s * (n_s + 1) / 2
    top = 100 * H  # the cube's top, 25 mm
    # half a pitch off: 0.75 mm from both neighbours at the seam, then further away from mouth 4 and
    # closer to mouth 5 (reached 1.0 mm off the axis, at z - z0 = 25.3 mm: above the top)
    ex = ez * 0.5 * 1.5e-3 / k
    cube, port, geom = setup([(ex, 0, ez), None, None, None], [(0, 0, ez), None, None, None])
    n_top = crossing_step(1.0e-3, a, dt, top)
    c, t, s = run(oracle_mod, cube, port, geom, one_ion(), drive(), rest_physics(n_top - 1))
    assert s["status"][0] == 5 and s["region"][0] == 0 and s["z"][0] > 20.0e-3
    c, t, s = run(oracle_mod, cube, port, geom, one_ion(), drive(), rest_physics(n_top + 5))
    assert s["status"][0] == 1 and s["steps"][0] == n_top and s["region"][0] == 0 and t["lost"] == 1
    assert sum(p["entered"] for p in c) == 0 and s["z"][0] > top
    # 0.6 pitch off (0.9 mm at the seam): mouth 5's aperture begins 1.0 mm off the axis
    ex = ez * 0.6 * 1.5e-3 / k
    ax_ratio = accel(7.0, 8565.0, ex) / a
    cube, port, geom = setup([(ex, 0, ez), None, None, None], [(0, 0, ez), None, None, None])
    n_e = n_s
    while ax_ratio * a * dt * dt * n_e * (n_e + 1) / 2 <= 1.0e-3:
        n_e += 1
    z_e = 1.0e-3 + a * dt * dt * n_e * (n_e + 1) / 2
    assert 20.0e-3 < z_e < top and n_e > n_s + 10
    c, t, s = run(oracle_mod, cube, port, geom, one_ion(), drive(), rest_physics(n_e))
    assert s["region"][0] == 1 + 5 and c[5]["entered"] == 1 and s["steps"][0] == n_e
    assert s["z"][0] == pytest.approx(0.5e-3 + (z_e - 20.0e-3), rel=1e-9, abs=1e-15)
    c, t, s = run(oracle_mod, cube, port, geom, one_ion(), drive(), rest_physics(n_e - 1))
    assert s["region"][0] == 0 and s["status"][0] == 5 and sum(p["entered"] for p in c) == 0


def test_back_through_the_entrance(oracle_mod):
    """A +y push on the splitter ion: through the entrance aperture on the y = 6 mm face (lib.rs:231-236), along
    the tube, back once it is l_ext beyond the face (lib.rs:237-243); with the aperture 1 mm off, lost on the
    wall at the face. y_n = y_0 + a dt^2 n (n + 1) / 2 from rest."""
    ey, dt = 40.0, 4e-8
    a = accel(7.0, 8565.0, ey)
    face, l_ext = 24 * H, 1.0e-3
    for ent_z, lost in ((1.0e-3, False), (2.0e-3, True)):
        cube, port, geom = setup([(0, ey, 0), None, None, None], [(0, 0, ey), None, None, None],
                                 entrance=(3.0e-3, ent_z), l_ext=l_ext)
        y0 = cube["axis_y"]
        n_face = crossing_step(y0, a, dt, face)
        n_back = crossing_step(y0, a, dt, face + l_ext)
        assert n_back > n_face + 5
        c, t, s = run(oracle_mod, cube, port, geom, one_ion(), drive(), rest_physics(n_back + 50))
        if lost:
            assert s["status"][0] == 1 and s["steps"][0] == n_face and t["lost"] == 1 and t["back"] == 0
        else:
            assert s["status"][0] == 2 and s["steps"][0] == n_back and t["back"] == 1
            assert s["y"][0] > face + l_ext and s["region"][0] == 0
            # inside the tube one step earlier
            c, t, s = run(oracle_mod, cube, port, geom, one_ion(), drive(), rest_physics(n_back - 1))
            assert s["status"][0] == 5 and face < s["y"][0] <= face + l_ext


def test_detected_through_the_entrance(oracle_mod):
    """The combiner ends where the splitter began: a forwarded ion pushed +y in the combiner leaves through the
    entrance aperture and is detected l_ext beyond the face (lib.rs:237-243). From its state at the combiner
    entry (taken from the oracle at that step) the combiner's uniform +y push gives the face step in closed
    form; the entrance is centred on where the ion meets the face."""
    ez, dt = 40.0, 4e-8
    ey = 40.0
    bases = [(0, 0, ez), (0, ey, 0), None, None]
    d = drive(dc_split=(1, 0, 0, 0), dc_combine=(0, 15.0, 0, 0))
    l_ext = 0.2e-3
    cube, port, geom = setup(bases, [(0, 0, ez), None, None, None], l_ext=l_ext)
    # find the forward step and the combiner entry state
    n = 1
    while True:
        c, t, s = run(oracle_mod, cube, port, geom, one_ion(z=8.0), d, rest_physics(n))
        if s["region"][0] == 10:
            break
        n = n * 2 if n < 4096 else n + 4096
        assert n < 400000
    lo, hi = n // 2, n
    while hi - lo > 1:
        mid = (lo + hi) // 2
        c, t, s = run(oracle_mod, cube, port, geom, one_ion(z=8.0), d, rest_physics(mid))
        lo, hi = (mid, hi) if s["region"][0] != 10 else (lo, mid)
    c, t, s = run(oracle_mod, cube, port, geom, one_ion(z=8.0), d, rest_physics(hi))
    assert s["region"][0] == 10 and c[4]["forwarded"] == 1
    x0, y0, z0, vz, vy0 = s["x"][0], s["y"][0], s["z"][0], s["vz"][0], s["vy"][0]
    assert vz < 0 and abs(vy0) < 1e-9  # at rest across (the 0 K start leaves a 1e-15 m/s Maxwellian)
    ay = accel(8.0, 8565.0, 15.0 * ey)
    face = 24 * H
    k_face = crossing_step(y0, ay, dt, face)
    z_face = z0 + vz * dt * k_face
    assert z_face > 0.6e-3
    k_det = crossing_step(y0, ay, dt, face + l_ext)
    z_det = z0 + vz * dt * k_det
    assert abs(z_det - z_face) < 0.4e-3  # still inside the 0.5 mm aperture when detected
    cube, port, geom = setup(bases, [(0, 0, ez), None, None, None], entrance=(x0, z_face), l_ext=l_ext)
    c, t, s = run(oracle_mod, cube, port, geom, one_ion(z=8.0), d, rest_physics(hi + k_det + 100))
    assert s["status"][0] == 3 and t["detected"] == 1 and s["steps"][0] == hi + k_det
    assert s["region"][0] == 10 and s["y"][0] > face + l_ext


def test_return_then_reentry_counts_a_revisit(oracle_mod):
    # splitter pushes up, the port pushes back down: the ion bounces between
    # splitter and port 4; each re-entry is a revisit with retry_from[4]
    ez = 40.0
    cube, port, geom = setup([(0, 0, ez), None, None, None], [(0, 0, -ez), None, None, None])
    dt = 4e-8
    a = accel(7.0, 8565.0, ez)
    n_s = crossing_step(1.0e-3, a, dt, 20.0e-3)
    # in the port the ion decelerates uniformly and comes back after ~2 v/a more steps
    c, t, s = run(oracle_mod, cube, port, geom, one_ion(), drive(), rest_physics(n_s + 2 * n_s + 5))
    assert c[4]["entered"] == 1 and c[4]["returned"] == 1
    assert s["region"][0] == 0 and s["last"][0] == 4 and s["vz"][0] < 0
    # the same symmetric bounce repeats: entries and returns alternate, all from port 4
    c, t, s = run(oracle_mod, cube, port, geom, one_ion(), drive(), rest_physics(20 * n_s))
    e, r = c[4]["entered"], c[4]["returned"]
    assert e >= 3 and r in (e - 1, e)
    assert c[4]["revisits"] == e - 1 and c[4]["retry_from"][4] == e - 1 and c[4]["from_other"] == 0
    assert sum(sum(p["retry_from"]) for p in c) == e - 1


def test_forward_combine_and_floor(oracle_mod):
    # port accelerates the ion past z_exit: forwarded (by job); the combiner pushes it down onto the cube's
    # floor, away from the entrance face: lost there
    ez = 40.0
    cube, port, geom = setup([(0, 0, ez), None, None, None], [(0, 0, ez), None, None, None])
    d = drive(dc_combine=(-1, 0, 0, 0))
    c, t, s = run(oracle_mod, cube, port, geom, one_ion(z=8.0), d, rest_physics(200000))
    assert c[4]["forwarded"] == 1 and c[4]["by_job"][0] == 1 and c[4]["returned"] == 0
    assert s["status"][0] == 1 and s["region"][0] == 10 and t["lost"] == 1 and s["last"][0] == 255
    assert s["z"][0] < 0.0 and s["vz"][0] < 0


def test_reverse_entry_from_the_combiner(oracle_mod):
    # combiner push up, three times the splitter's: the forwarded ion (which gained its
    # energy over about 37 mm at 1 V) stops within about 12 mm and is turned back into
    # the port it came from, in reverse
    ez = 40.0
    cube, port, geom = setup([(0, 0, ez), None, None, None], [(0, 0, ez), None, None, None])
    d = drive(dc_combine=(3, 0, 0, 0))
    c, t, s = run(oracle_mod, cube, port, geom, one_ion(), d, rest_physics(60000))
    assert c[4]["forwarded"] >= 1 and c[4]["reverse_entries"] >= 1
    assert c[4]["reverse_entries"] in (c[4]["forwarded"], c[4]["forwarded"] - 1)


def test_return_to_another_port_counts_from_other(oracle_mod):
    # the port sends the ion back; a transverse field in the splitter carries it to mouth 5
    ez = 40.0
    dt = 4e-8
    a = accel(7.0, 8565.0, ez)
    n_s = crossing_step(1.0e-3, a, dt, 20.0e-3)
    # small transverse field r = Ex/Ez = 0.006: the first crossing is 19 mm * r = 0.11 mm off
    # the central mouth; after the port's round trip (x drifts 76 mm * r at the gained speed)
    # and the splitter's (152 mm * r) the next crossing is ~247 mm * r = 1.5 mm off: mouth 5
    cube, port, geom = setup([(0.24, 0, ez), None, None, None], [(0, 0, -ez), None, None, None])
    c, t, s = run(oracle_mod, cube, port, geom, one_ion(), drive(), rest_physics(40 * n_s))
    entered = [p["entered"] for p in c]
    assert sum(entered) >= 2
    total_revisits = sum(p["revisits"] for p in c)
    assert total_revisits == sum(entered) - 1
    assert sum(p["from_other"] for p in c) == sum(sum(p["retry_from"][j] for j in range(9) if j != i)
                                                  for i, p in enumerate(c))
    assert any(p["from_other"] > 0 for p in c), entered


def test_idle_ports_see_the_bias_and_driven_ports_the_tone(oracle_mod):
    # the idle bias is added to the port's DC; an idle port with a blocking bias (net -2 V:
    # the ion, which gained its energy over 19 mm at 1 V, stops within 9.5 mm) returns it,
    # the same port driven (tone > 0, zero amplitude) accelerates and forwards it
    ez = 40.0
    cube, port, geom = setup([(0, 0, ez), None, None, None], [(0, 0, ez), None, None, None])
    tones = [0.0] * 9
    d = drive(dc_combine=(-1, 0, 0, 0), idle=(-3, 0, 0, 0), tones=tones)
    c, t, s = run(oracle_mod, cube, port, geom, one_ion(), d, rest_physics(30000))
    assert c[4]["returned"] >= 1 and c[4]["forwarded"] == 0
    tones[4] = 100e3
    d = drive(dc_combine=(-1, 0, 0, 0), idle=(-3, 0, 0, 0), tones=tones)
    c, t, s = run(oracle_mod, cube, port, geom, one_ion(), d, rest_physics(30000))
    assert c[4]["forwarded"] == 1 and c[4]["returned"] == 0 and t["combine"] + t["lost"] == 1


def flow_checks(c, t, s, njobs):
    n = t["n"]
    assert n == t["split"] + t["channel"] + t["combine"] + t["lost"] + t["back"] + t["detected"]
    for p in range(9):
        q = c[p]
        # ions still travelling in port p at the end = entries - exits (returned, forwarded,
        # lost); a lost ion keeps its region, so only unfinished ions (status 5) count here
        inside = int(np.sum((s["region"] == 1 + p) & (s["status"] == 5)))
        assert q["entered"] + q["reverse_entries"] - q["returned"] - q["forwarded"] - q["lost"] == inside
        assert q["revisits"] == sum(q["retry_from"])
        assert q["from_other"] == q["revisits"] - q["retry_from"][p]
        assert sum(q["by_job"][:njobs]) == q["forwarded"] and sum(q["by_job"][njobs:]) == 0
    assert int(np.sum(s["status"] == 5)) == t["split"] + t["channel"] + t["combine"]
    assert int(np.sum(s["steps"])) == t["ion_steps"]


def test_random_run_flow_balance(oracle_mod):
    cube, port, geom, d, jobs, ph = synth.tandem_setup(mqhip, "small", n_per_charge=64, max_steps=15000)
    c, t, s = oracle_mod.tandem_run(cube, port, geom, jobs, d, ph, 8)
    assert t["n"] == 256
    flow_checks(c, t, s, len(jobs))
    assert sum(p["entered"] for p in c) > 0 and sum(p["returned"] for p in c) > 0


def closed_form_case(name, oracle_mod=None):
    """(cube, port, geom, drive, jobs, physics) of one single-ion case of this file, for the GPU's FP64 path
    to reproduce bit for bit (tests/test_gpu_tandem.py)."""
    ez, dt = 40.0, 4e-8
    a = accel(7.0, 8565.0, ez)
    n_s = crossing_step(1.0e-3, a, dt, 20.0e-3)
    k = a * dt * dt * n_s * (n_s + 1) / 2
    jobs = one_ion()
    if name in ("entry", "bounce", "forward_floor", "reverse", "from_other"):
        port_e = -ez if name in ("bounce", "from_other") else ez
        cube_e = (0.24, 0, ez) if name == "from_other" else (0, 0, ez)
        cube, port, geom = setup([cube_e, None, None, None], [(0, 0, port_e), None, None, None])
        dcc = {"forward_floor": (-1, 0, 0, 0), "reverse": (3, 0, 0, 0)}.get(name, (1, 0, 0, 0))
        if name == "forward_floor":
            jobs = one_ion(z=8.0)
        return cube, port, geom, drive(dc_combine=dcc), jobs, rest_physics(60000 if name != "entry" else n_s + 3)
    if name in ("board_top", "board_mouth"):
        ex = ez * (0.5 if name == "board_top" else 0.6) * 1.5e-3 / k
        cube, port, geom = setup([(ex, 0, ez), None, None, None], [(0, 0, ez), None, None, None])
        return cube, port, geom, drive(), jobs, rest_physics(crossing_step(1.0e-3, a, dt, 100 * H) + 5)
    if name in ("back", "wall"):
        cube, port, geom = setup([(0, ez, 0), None, None, None], [(0, 0, ez), None, None, None],
                                 entrance=(3.0e-3, 1.0e-3 if name == "back" else 2.0e-3), l_ext=1.0e-3)
        return cube, port, geom, drive(), jobs, rest_physics(2000)
    if name == "detect":
        bases = [(0, 0, ez), (0, ez, 0), None, None]
        d = drive(dc_split=(1, 0, 0, 0), dc_combine=(0, 15.0, 0, 0))
        cube, port, geom = setup(bases, [(0, 0, ez), None, None, None], l_ext=0.2e-3)
        # centre the entrance where the forwarded ion meets the y face (found with the oracle, as in
        # test_detected_through_the_entrance)
        ph = rest_physics(200000)
        c, t, s = run(oracle_mod, cube, port, geom, one_ion(z=8.0), d, ph)
        hit = (s["x"][0], s["z"][0])
        cube, port, geom = setup(bases, [(0, 0, ez), None, None, None], entrance=hit, l_ext=0.2e-3)
        return cube, port, geom, d, one_ion(z=8.0), ph
    raise KeyError(name)


def test_closed_form_case_table_runs(oracle_mod):
    # the end of the cases whose closed forms the tests above check
    want = {"entry": (5, 5), "forward_floor": (1, 10), "board_top": (1, 0), "board_mouth": (5, 6), "back": (2, 0),
            "wall": (1, 0), "detect": (3, 10)}
    for name, (status, region) in want.items():
        cube, port, geom, d, jobs, ph = closed_form_case(name, oracle_mod)
        c, t, s = run(oracle_mod, cube, port, geom, jobs, d, ph)
        assert (s["status"][0], s["region"][0]) == (status, region), name
