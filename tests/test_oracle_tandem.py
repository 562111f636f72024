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


def setup(cube_e, port_e, pitch=1.5e-3, ap_r=0.5e-3, seam=20.0e-3, z_back=0.5e-3):
    cube = uniform_grid(25, 25, 101, H, cube_e)                 # 6 x 6 x 25 mm
    port = uniform_grid(21, 21, 81, H, port_e, z_exit=18.0e-3)  # 5 x 5 x 20 mm
    ax, ay = cube["axis_x"], cube["axis_y"]
    mouths = [(ax + (c - 1) * pitch, ay + (r - 1) * pitch) for r in range(3) for c in range(3)]
    geom = mqhip.tandem_geom(seam, mouths, ap_r, z_back)
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
    for frac_pitch, port_hit in [(1.0, 5), (-1.0, 3), (0.5, None)]:
        ex = ez * frac_pitch * 1.5e-3 / k
        cube, port, geom = setup([(ex, 0, ez), None, None, None], [(0, 0, ez), None, None, None])
        c, t, s = run(oracle_mod, cube, port, geom, one_ion(), drive(), rest_physics(n_s + 2))
        if port_hit is None:
            # half a pitch off: 0.75 mm from both neighbours, outside the 0.5 mm apertures: the board
            assert s["status"][0] == 1 and s["region"][0] == 0 and t["lost"] == 1
            assert sum(p["entered"] for p in c) == 0
        else:
            assert s["region"][0] == 1 + port_hit and c[port_hit]["entered"] == 1
            # the drift puts the ion on the mouth centre at the seam, so it lands on the
            # port axis and moves on only by its transverse speed over the two extra steps
            vx = accel(7.0, 8565.0, ex) * dt * n_s
            assert abs(s["x"][0] - port["axis_x"]) < 3 * abs(vx) * dt + 1e-9


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


def test_forward_combine_and_detect(oracle_mod):
    # port accelerates the ion past z_exit: forwarded (by job), combiner push down: detected
    ez = 40.0
    cube, port, geom = setup([(0, 0, ez), None, None, None], [(0, 0, ez), None, None, None])
    d = drive(dc_combine=(-1, 0, 0, 0))
    c, t, s = run(oracle_mod, cube, port, geom, one_ion(z=8.0), d, rest_physics(200000))
    assert c[4]["forwarded"] == 1 and c[4]["by_job"][0] == 1 and c[4]["returned"] == 0
    assert s["status"][0] == 3 and s["region"][0] == 10 and t["detected"] == 1 and s["last"][0] == 255
    assert s["z"][0] < 0.5e-3 and s["vz"][0] < 0


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
    assert c[4]["forwarded"] == 1 and c[4]["returned"] == 0 and t["detected"] == 1


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
