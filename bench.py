#!/usr/bin/env python3
"""Throughput of the batched ion-trajectory integrator on one MI355X.

One step = one run_batch of the whole workload (inject, integrate every ion to
termination, summarize), inputs resident in HBM. The default workload is the
batched mass scan: m/z 475..523 step 2 (25 masses) x 64 RF frequencies
(450..550 kHz) x 1000 ions = 1.6M trajectories on a 128x128x512 grid at
0.05 mm, Fig-4 drive (100 V RF, 0.5 V AC at 100 kHz, 5 V DC gate), half-kick
(1/2 a dt) push, thermal N2 at 1 mTorr / 300 K, Gaussian 0.1 mm spot with a
300 K Maxwell-Boltzmann start, dt = 100 ns, 20,000-step cap.

Prints one JSON line: value = FP32 batched ion-steps/s; plus the plain FP64
form on the same inputs, the outcome agreement between them, the roofline
object of the batched worker kernel and the CPU baseline (the FP64 oracle on
the host's cores over a bounded sample of the same jobs). Sub-objects: the
differentiable form (gradient evaluations/s), the tandem multi-domain batch
(1M ions), the reference channel geometry solved on the GPU with the two
numbers the reference recorded for it, the legacy Java potentials
regenerated on the GPU with the published Fig. 4 columns on them, and the
drivers around the path (live population, Optimize searches, adaptive scan).
"""
from __future__ import annotations

import argparse
import csv
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "multiq-workbench_amd"))

import numpy as np  # noqa: E402

import mqhip  # noqa: E402
import synth  # noqa: E402

BYTES_PER_STEP_FP32 = 292.0   # 8 corners x 3 bases x 3 components x 4 B + 4 B mask word
BYTES_PER_STEP_FP64 = 580.0
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload(name: str):
    if name == "c2":
        grid = synth.c2_grid()
        jobs = synth.c2_jobs(mqhip, n=1000)
        ph = mqhip.Physics.correct(0.133322).copy(dt=1e-7, max_steps=20000, injection=mqhip.INJ_GAUSSIAN,
                                                  inj_sigma_mm=0.1, legacy_force=1)
        desc = ("mass scan: 25 m/z (475-523) x 64 f_rf (450-550 kHz) x 1000 ions = 1.6M trajectories; "
                "128x128x512 grid @ 0.05 mm (FP32 cell tables); 100 V RF, 0.5 V AC @ 100 kHz, 5 V DC gate; "
                "half-kick push; thermal N2 1 mTorr 300 K; Gaussian 0.1 mm spot, 300 K MB; dt 100 ns; cap 20000 steps")
        return grid, jobs, [0.0, 5.0, 0.0, 0.0], ph, desc
    if name == "c3":
        grid = synth.c3_grid()
        jobs = synth.c3_jobs(mqhip)
        ph = mqhip.Physics.correct(0.133322).copy(dt=1e-7, max_steps=20000, injection=mqhip.INJ_GAUSSIAN,
                                                  inj_sigma_mm=0.1, legacy_force=1)
        desc = ("operating-setting sweep: 4096 LHS settings (V_rf 50-150 V, f_exc 80-120 kHz, V_exc 0.1-1 V, "
                "V_dc 0-10 V, 0.5-2 mTorr) x 256 ions = 1M trajectories; 192x192x768 grid @ 1/30 mm")
        return grid, jobs, [0.0, 1.0, 0.0, 0.0], ph, desc
    if name == "c1":
        grid = synth.c1_grid()
        jobs = synth.c1_jobs(mqhip, n=1000)
        ph = mqhip.Physics.correct(0.133322).copy(dt=1e-7, max_steps=20000, injection=mqhip.INJ_GAUSSIAN,
                                                  inj_sigma_mm=0.1, legacy_force=1)
        return grid, jobs, [0.0, 5.0, 0.0, 0.0], ph, "Fig-4 single channel, 1000 ions, 64x64x256 @ 0.1 mm"
    raise SystemExit(f"unknown config {name}")


def pulls(a, b):
    """Per-job, per-class binomial pulls between two outcome lists."""
    out = []
    for x, y in zip(a, b):
        for k in ("crashed", "returned", "transmitted", "left", "timeout"):
            p1, p2 = getattr(x, k) / max(x.n, 1), getattr(y, k) / max(y.n, 1)
            se = np.sqrt(p1 * (1 - p1) / max(x.n, 1) + p2 * (1 - p2) / max(y.n, 1))
            out.append(0.0 if se == 0 and p1 == p2 else (p1 - p2) / max(se, 1e-12))
    return np.array(out)


def cpu_baseline(grid, jobs, dc, ph, seconds: float):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # the FP64 oracle: CPU baseline only
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    threads = max(1, min(16, ncpu))
    # calibrate on `threads` evenly spaced jobs (one job alone leaves most threads idle behind
    # its longest-lived ions), then size the sample to about `seconds`
    cal = [jobs[i] for i in np.linspace(0, len(jobs) - 1, min(len(jobs), threads)).astype(int)]
    t0 = time.perf_counter()
    r = oracle.run_batch(grid, cal, dc, ph, threads)
    t1 = time.perf_counter() - t0
    rate0 = r["steps"].sum() / max(t1, 1e-9)
    per_job = r["steps"].sum() / len(cal)
    njobs = int(max(1, min(len(jobs), seconds * rate0 / max(per_job, 1))))
    pick = [jobs[i] for i in np.linspace(0, len(jobs) - 1, njobs).astype(int)]
    log(f"cpu baseline: oracle on {njobs} jobs x {pick[0].n} ions with {threads} threads")
    t0 = time.perf_counter()
    r = oracle.run_batch(grid, pick, dc, ph, threads)
    dt = time.perf_counter() - t0
    steps = int(r["steps"].sum())
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": steps / dt, "unit": "ion-steps/s", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle (C restatement of the reference step) on {njobs} evenly spaced jobs of the "
                      f"workload ({njobs * pick[0].n} trajectories, {steps} ion-steps, {dt:.1f} s); host: {model}, "
                      f"{ncpu} CPUs visible"}


def grad_bench(fd: bool, evals: int = 3):
    """Differentiable form: 65,536 collision-free ions, m/z 499, 20,000 steps, 0.05 mm grid with
    FP32-rounded basis values; one gradient evaluation = forward + backward (adjoint).
    Operating point: the Fig-4 drive with a 0.5 V gate. The adjoint is the exact derivative of
    the discrete map, ion by ion. A few ions end near r_leak on orbits whose sensitivity grows
    over the 20,000 steps (|dJ_i/dV_rf| up to ~2e5 per volt): a step of 1e-6 relative moves them
    far outside the linear range, so central differences are compared with the adjoint over all
    ions and over the ions the step keeps in its linear range (fd_check), and again over all
    ions on a 2,000-step horizon."""
    grid = synth.c2_grid()
    for k in ("rf", "ac"):
        grid[k] = grid[k].astype(np.float32).astype(np.float64)
    grid["dc"] = [d.astype(np.float32).astype(np.float64) if d is not None else None for d in grid["dc"]]
    eng = mqhip.Engine(grid)
    ph = mqhip.Physics.correct(0.0).copy(dt=1e-7, injection=mqhip.INJ_GAUSSIAN, inj_sigma_mm=0.1, legacy_force=1)
    job = mqhip.job(499.0, 1.0, 0.0, 100.0, 500e3, 0.5, 100e3, 65536, 4499, dc_scale=0.5)
    obj = mqhip.Objective(20000, 0, 18.6e-3, 2e-5, 1.0e-3, 0.3)
    dc = [0.0, 1.0, 0.0, 0.0]
    log("gradient: forward + backward of 65536 ions x 20000 steps")
    # one untimed evaluation first: it builds the float cell records (once per grid and DC set)
    eng.forward(job, dc, ph, obj)
    eng.backward(1.0)
    t0 = time.perf_counter()
    for _ in range(evals):
        J = eng.forward(job, dc, ph, obj)
        fwd = eng.stats()
        grad = eng.backward(1.0)
        bwd = eng.stats()
    dt = (time.perf_counter() - t0) / evals
    log(f"gradient: {1.0 / dt:.3f} evaluations/s (forward {fwd['kernel_ms']:.1f} ms, backward {bwd['kernel_ms']:.1f} ms)")
    # the backward recomputes the checkpoint segments (2 passes) unless the whole trajectory was kept
    whole = bwd["ion_steps"] == 65536 * 20000
    out = {"evals_per_s": 1.0 / dt, "seconds_per_eval": dt, "evals_timed": evals, "J": J,
           "grad": dict(zip(["V_rf", "f_rf", "V_exc", "f_exc", "V_dc"], grad.tolist())),
           "forward_ms": fwd["kernel_ms"], "backward_ms": bwd["kernel_ms"],
           "ion_steps_per_s": (fwd["ion_steps"] + bwd["ion_steps"]) / dt,
           "config": "65536 ions, m/z 499 z=1, 20000 collision-free steps, 128x128x512 @ 0.05 mm (FP32-rounded), "
                     "100 V RF @ 500 kHz, 0.5 V AC @ 100 kHz, 0.5 V gate; objective mean sigmoid((z-18.6 mm)/0.02 mm)"
                     " - 0.3 sigmoid((r-1 mm)/0.02 mm); " +
                     ("the whole trajectory kept in device memory (31 GB), no recompute" if whole
                      else "sqrt(N) checkpoints, segments recomputed")}
    if fd:
        axis = (grid["axis_x"], grid["axis_y"])
        out["fd_check_h_rel_1e-6"] = fd_check(eng, job, dc, ph, obj, grad, axis)
        # smaller steps keep even the most sensitive trajectories in their linear range at 20,000 steps
        # (|dJ_i/dV_rf| ~ 2e5 per volt times a 1e-8 V step); FP64 leaves room for them: a phase f dt n of
        # order 1e3 carries 1e-13 of rounding against a shift of 1e-7 at h = 1e-10
        for hr in (1e-9, 1e-10, 1e-11, 1e-12):
            out[f"fd_check_h_rel_{hr:g}"] = fd_check(eng, job, dc, ph, obj, grad, axis, hr)
        # the same operating point over 2,000 steps, where every trajectory stays in the
        # linear range of the step: all ions, no selection
        obj_s = mqhip.Objective(2000, 0, 18.6e-3, 2e-5, 1.0e-3, 0.3)
        eng.forward(job, dc, ph, obj_s)
        grad_s = eng.backward(1.0)
        out["fd_check_2000_steps_h_rel_1e-6"] = fd_check(eng, job, dc, ph, obj_s, grad_s, axis)
    eng.close()
    return out


def fd_check(eng, job, dc, ph, obj, grad, axis, h_rel=1e-6):
    """Adjoint vs central differences (h = h_rel relative, default 1e-6) and their Richardson extrapolation
    (h, h/2), over all ions and over the ions that the step keeps in its linear range: an ion
    whose final z or radius moves by more than 1 % of the objective's length scale between
    theta - h and theta + h is outside it (the step then samples the sigmoid's curvature, or
    a trajectory whose sensitivity grew over the run, rather than its slope). The split uses
    forward runs only."""
    n = int(job.n)
    gi = eng.backward_per_ion(n)   # per-ion adjoint contributions (dJ_i/dtheta / n)
    names = [("drive", "v_rf", "V_rf"), ("drive", "f_rf", "f_rf"), ("drive", "v_ac", "V_exc"),
             ("drive", "f_ac", "f_exc"), (None, "dc_scale", "V_dc")]

    def central(grp, name, h):
        base = getattr(job.drive if grp else job, name)
        vals, per, fin = [], [], []
        for sgn in (+1, -1):
            j2 = mqhip.Job()
            j2.sp, j2.drive, j2.n, j2.seed, j2.dc_scale = job.sp, job.drive, job.n, job.seed, job.dc_scale
            setattr(j2.drive if grp else j2, name, base + sgn * h)
            vals.append(eng.forward(j2, dc, ph, obj))
            per.append(eng.forward_per_ion() / n)
            fin.append(eng.forward_final())
        return (vals[0] - vals[1]) / (2 * h), (per[0] - per[1]) / (2 * h), fin

    def moved(fin):
        (xp, yp, zp), (xm, ym, zm) = fin
        rp = np.hypot(xp - axis[0], yp - axis[1])
        rm = np.hypot(xm - axis[0], ym - axis[1])
        return np.maximum(np.abs(zp - zm), np.abs(rp - rm))

    rel = {}
    for k, (grp, name, label) in enumerate(names):
        h = h_rel * abs(getattr(job.drive if grp else job, name))
        d1, p1, f1 = central(grp, name, h)
        d2, p2, _ = central(grp, name, h / 2)
        rich, prich = (4 * d2 - d1) / 3, (4 * p2 - p1) / 3
        g = float(grad[k])
        lin = moved(f1) <= 0.01 * obj.scale
        adj_l, fd_l = float(gi[k][lin].sum()), float(prich[lin].sum())
        mag_l = float(np.abs(prich[lin]).sum())
        top = np.argsort(np.abs(gi[k]))[::-1][:3]
        rel[label] = {"adjoint": g, "central_h": d1, "central_h2": d2, "richardson": rich,
                      "rel_err_richardson_all_ions": abs(rich - g) / max(abs(g), 1e-300),
                      "linear_range_ions": float(lin.mean()),
                      "adjoint_over_linear": adj_l, "richardson_over_linear": fd_l,
                      "rel_err_over_linear": abs(adj_l - fd_l) / max(abs(fd_l), 1e-300),
                      "err_over_linear_per_abs_sum": abs(adj_l - fd_l) / max(mag_l, 1e-300),
                      "outside_share_of_adjoint": float(gi[k][~lin].sum() / g) if g != 0 else 0.0,
                      "largest_ions": [{"ion": int(i), "adjoint": float(gi[k][i] * n),
                                        "central_h": float(p1[i] * n), "central_h2": float(p2[i] * n),
                                        "linear": bool(lin[i])} for i in top]}
        log(f"gradient check {label} (h {h_rel:g}): adjoint {g:.8g} richardson {rich:.8g}; over the {lin.mean():.5f} of ions "
            f"in the step's linear range: adjoint {adj_l:.10g} richardson {fd_l:.10g}")
    return rel


def tandem_bench(steps: int = 2):
    """Tandem multi-domain batch (C5): ubiquitin 6+..9+ (4 x 262144 ions) through a
    splitter, nine ports (four driven) and a combiner on one clock; FP32 batched
    vs FP64 plain on the same inputs, per-port counts and the retry matrix."""
    cube, port, geom, d, jobs, ph = synth.tandem_setup(mqhip, "c5")
    t = mqhip.Tandem(cube, port, geom)
    t.run(jobs, d, ph, mqhip.PATH_BATCHED)  # warmup: cell tables
    n_steps = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        cb, tb = t.run(jobs, d, ph, mqhip.PATH_BATCHED)
        n_steps += t.stats()["ion_steps"]
    wall = time.perf_counter() - t0
    kms = t.stats()["kernel_ms"]
    t0 = time.perf_counter()
    cp, tp = t.run(jobs, d, ph, mqhip.PATH_PLAIN)
    pwall = time.perf_counter() - t0
    psteps = t.stats()["ion_steps"]
    t.close()
    log(f"tandem: batched {n_steps / wall:.3e} ion-steps/s ({n_steps // steps} ion-steps per run), "
        f"plain {psteps / pwall:.3e} ({psteps} ion-steps)")

    def pull(a, b):
        return (a - b) / np.sqrt(max(a + b, 1))

    keys = ("entered", "returned", "forwarded", "lost", "revisits", "from_other")
    port_pulls = [pull(cb[p][k], cp[p][k]) for p in range(9) for k in keys]
    end_pulls = [pull(tb[k], tp[k]) for k in ("split", "channel", "combine", "lost", "back", "detected")]
    return {
        "config": "ubiquitin (8565 Da) 6+..9+, 4 x 262144 = 1048576 ions; splitter/combiner grid 128x128x512 @ "
                  "0.05 mm (confining RF, axial push), nine ports on one 128x128x512 @ 0.05 mm channel grid; "
                  "115 V RF @ 500 kHz shared clock; ports 0, 2, 6, 8 driven at the 6+..9+ secular frequencies "
                  f"{[round(x / 1e3, 1) for x in d.tones if x > 0]} kHz (0.5 V), idle ports +4 V gate; "
                  "combiner push of 100 V to the y-face exit (1 mm below the board, in line with ports 2, 5, 8); "
                  "thermal N2 1 mTorr; dt 40 ns; cap 30000 steps",
        "value": n_steps / wall, "unit": "ion-steps/s", "kernel_ion_steps_per_s": n_steps / max(kms * 1e-3 * steps, 1e-9),
        "plain_fp64": psteps / pwall, "speedup_vs_plain_fp64": (n_steps / wall) / (psteps / pwall),
        "ms_per_run": 1e3 * wall / steps,
        "totals_batched": tb, "totals_plain": tp,
        "ports_batched": [{k: cb[p][k] for k in keys + ("reverse_entries",)} | {"by_job": cb[p]["by_job"][:4]}
                          for p in range(9)],
        "ports_plain": [{k: cp[p][k] for k in keys + ("reverse_entries",)} | {"by_job": cp[p]["by_job"][:4]}
                        for p in range(9)],
        "retry_matrix_batched": [cb[p]["retry_from"] for p in range(9)],
        "retry_matrix_plain": [cp[p]["retry_from"] for p in range(9)],
        "count_pulls_within_4sigma": float(np.mean(np.abs(np.array(port_pulls + end_pulls)) <= 4.0)),
        "max_abs_count_pull": float(np.max(np.abs(np.array(port_pulls + end_pulls)))),
        "fp64_exactness": "the FP64 path's per-port counts and retry matrix equal the CPU oracle's exactly "
                          "(tests/test_gpu_tandem.py)",
    }


def reference_geometry_bench():
    """The reference's own channel geometry, solved on the GPU, and two numbers the
    reference recorded: REF_CHSIM_GPU (validate.rs:110-112) and the legacy Fig. 4 peak."""
    out = {}
    t0 = time.perf_counter()
    g14 = mqhip.channel_grid(mqhip.Geom.s14(), tol=1e-7)
    out["field_build_s"] = time.perf_counter() - t0
    out["field_build_cg_iterations"] = [st["iterations"] for st in g14["stats"]]
    e = mqhip.Engine(g14)
    ph = mqhip.Physics.correct(0.1).copy(t_max=3e-3)
    j = mqhip.job(2 * 863.473, 2.0, 425.0, 70.0, 500e3, 0.35, 84.25e3, 100000, 4242)
    o = e.run_batch([j], [18, 24, 15, 20], ph, mqhip.PATH_BATCHED)[0]
    st = e.stats()
    e.close()
    ref = {"transmitted": 19669, "crashed_plus_left": 3412, "returned": 75797, "timeout": 1122}
    got = {"transmitted": o.transmitted, "crashed_plus_left": o.crashed + o.left, "returned": o.returned,
           "timeout": o.timeout}
    p = {k: (got[k] - ref[k]) / 1e5 / np.sqrt((got[k] / 1e5 * (1 - got[k] / 1e5) + ref[k] / 1e5 * (1 - ref[k] / 1e5))
                                             / 1e5) for k in ref}
    out["s14_ref_chsim_gpu"] = {"batched": got, "reference": ref, "pulls": p,
                                "ion_steps_per_s": st["ion_steps"] / (st["wall_ms"] * 1e-3)}
    gd = mqhip.channel_grid(mqhip.Geom.default(), tol=1e-7)
    e = mqhip.Engine(gd)
    mzs = [475.0 + 2 * i for i in range(25)]
    jobs = [mqhip.job(mz * 5, 5.0, 0.0, 100.0, 500e3, 0.5, 100e3, 1000, 1000 + int(mz)) for mz in mzs]
    outs = e.run_batch(jobs, [0, 5, -5, 0], mqhip.Physics.legacy(1e-3 * 133.322), mqhip.PATH_BATCHED)
    e.close()
    tr = [o.transmitted for o in outs]
    k = int(np.argmax(tr))
    out["fig4_legacy_converged"] = {"transmitted_per_1000": tr, "peak_mz": mzs[k], "peak": tr[k],
                                    "reference_rust_legacy_peak": 742}
    log(f"reference geometry: S14 pulls {p}, Fig. 4 peak {tr[k]} at m/z {mzs[k]}")
    return out


def drivers_bench():
    """The drivers around the path, timed on the engine (each is parity-tested in tests/test_gpu_*):
    * live: mqviz's block loop (engine.rs:448-538) on the Fig-4 grid at a 0.02 V gate, two species, Poisson
      injection at 5e8 ions/s, 128-step blocks of 100 ns steps; simulated time per wall second and ion-steps/s over
      40 timed blocks after 200 warm ones.
    * optimize_grid: one Optimize experiment (research.rs:133-163) on the corrected S14 channel: 4 AC amplitudes x
      AC-frequency scans of 70-100 kHz at 0.5 kHz x 500 ions per point (+ the Na adduct's 500 for the leakage),
      fits, rescans of narrow peaks, and the confirmation with 3x the ions.
    * optimize_grad: the gradient search over f_exc on the differentiable form (16,384 ions, 2,000 steps), started
      2 kHz from the optimum.
    * adaptive_scan: scan::adaptive_scan (scan.rs:52-78) on the Fig-4 grid at a 0.02 V gate, 80-120 kHz at 1 kHz
      coarse with 2,000 ions, refined at 0.25 kHz with 8,000 ions where T >= 0.35 Tmax and above the baseline, into
      a fresh result cache."""
    import tempfile
    out = {}
    # live population
    g1 = synth.c1_grid()
    e = mqhip.Engine(g1)
    try:
        ph = mqhip.Physics.correct(0.133322).copy(dt=1e-7, max_steps=20000, injection=mqhip.INJ_GAUSSIAN,
                                                  inj_sigma_mm=0.1, legacy_force=1)
        sp = [mqhip.Species(499.0, 1.0, 230.0), mqhip.Species(505.0, 1.0, 235.0)]
        live = mqhip.Live(e, sp, [0.6, 0.4], mqhip.RfAc(100.0, 500e3, 0.5, 100e3), [0, 0.02, 0, 0], ph,
                          block_steps=128, capacity=1 << 20, inj_rate=5e8, seed=5)
        try:
            live.run(200)                     # warm: the population approaches its steady size
            s0 = live.run(0)
            t0 = time.perf_counter()
            s1 = live.run(40)
            wall = time.perf_counter() - t0
        finally:
            live.close()
        sim = (s1["step"] - s0["step"]) * ph.dt
        out["live"] = {"blocks_per_s": 40 / wall, "simulated_s_per_wall_s": sim / wall,
                       "ion_steps_per_s": (s1["ion_steps"] - s0["ion_steps"]) / wall, "alive_end": s1["alive"],
                       "injected_in_timed_blocks": s1["injected"] - s0["injected"]}
        # gradient search
        phg = mqhip.Physics.correct(0.0).copy(dt=1e-7, injection=mqhip.INJ_GAUSSIAN, inj_sigma_mm=0.1,
                                              legacy_force=1)
        obj = mqhip.Objective(2000, 0, 18.6e-3, 2e-4, 0.5e-3, -1.0)
        dcg = [0.0, 1.0, 0.0, 0.0]
        jg = mqhip.job(499.0, 1.0, 0.0, 100.0, 500e3, 0.5, 94e3, 16384, 4499, dc_scale=0.5)
        e.forward(jg, dcg, phg, obj)  # warm: tables of the differentiable form
        t0 = time.perf_counter()
        r = mqhip.optimize_grad(e, jg, dcg, phg, obj, mask=1 << 3, lo=[0, 0, 0, 90e3, 0],
                                hi=[1e3, 1e7, 10, 104e3, 10], step0=[1, 1, 1, 1e3, 1], max_iter=40, tol=1e-3)
        wall = time.perf_counter() - t0
        out["optimize_grad"] = {"wall_s": wall, "evaluations": r["n_eval"], "s_per_evaluation": wall / r["n_eval"],
                                "f_exc_khz": r["theta"][3] / 1e3, "J": r["J"]}
        # adaptive scan into a fresh cache
        with tempfile.TemporaryDirectory() as td:
            cache = os.path.join(td, "scan.tsv")
            t0 = time.perf_counter()
            pts = mqhip.adaptive_scan(e, "bench-fig4-grid", mqhip.Species(499.0, 1.0, 230.0),
                                      mqhip.RfAc(100.0, 500e3, 0.5, 100e3), [0, 0.02, 0, 0], ph, 80.0, 120.0, 1.0,
                                      2000, 8000, -1.0, 100, cache)
            wall = time.perf_counter() - t0
        ions = int(sum(p[3] for p in pts))
        out["adaptive_scan"] = {"wall_s": wall, "points": len(pts), "ions": ions,
                                "trajectories_per_s": ions / wall}
    finally:
        e.close()
    # one Optimize experiment on the S14 channel
    g14 = mqhip.channel_grid(mqhip.Geom.s14(), tol=1e-7)
    e = mqhip.Engine(g14)
    try:
        sp = mqhip.Species(2 * 863.473, 2.0, 425.0)
        na = mqhip.Species(2 * 863.473 + 21.982, 2.0, 430.0)
        rf = mqhip.RfAc(70.0, 500e3, 0.35, 84.25e3)
        pho = mqhip.Physics.correct(0.1).copy(t_max=3e-3)
        t0 = time.perf_counter()
        res = mqhip.optimize_grid(e, sp, rf, [18, 24, 15, 20], pho, [0.35, 0.2625, 0.175, 0.0875], 70.0, 100.0, 0.5,
                                  500, 5.0, 0xA1100101, leak=na, leak_n=500)
        wall = time.perf_counter() - t0
        rows = res["rows"] + [res["confirmation"]]
        npts = sum(int(round((r["hi_khz"] - r["lo_khz"]) / r["step_khz"])) + 1 for r in rows)
        out["optimize_grid"] = {"wall_s": wall, "rows": len(res["rows"]), "best_row": res["best"],
                                "confirmed": bool(res["confirmed"]), "scan_points_approx": npts,
                                "best_r": res["rows"][res["best"]]["r_mz"] if res["best"] >= 0 else None}
    finally:
        e.close()
    log(f"drivers: {out}")
    return out


LEGACY_BYTES_PER_UPDATE = 9  # read one other-colour value, write one, read one electrode flag (split layout)


def legacy_potentials_bench(cpu_sweeps: int = 40):
    """The legacy Java potentials regenerated on the GPU (160 x 160 x 242, the sketch's settings), the
    published Fig. 4 columns on them, and the sketch's sequential sweep timed on one CPU core."""
    import tempfile

    out = {}
    d = tempfile.mkdtemp(prefix="theirs_")
    t0 = time.perf_counter()
    st = mqhip.legacy_export(d)
    out["export_wall_s"] = time.perf_counter() - t0
    upd = sum(x["node_updates"] for x in st)
    ms = sum(x["ms"] for x in st)
    out["bases"] = {n: {"sweeps": x["sweeps"], "ms": x["ms"], "node_updates": x["node_updates"]}
                    for n, x in zip(("RF", "DC", "AC"), st)}
    out["node_updates_per_s"] = upd / (ms * 1e-3)
    out["gbs_at_9B_per_update"] = upd * LEGACY_BYTES_PER_UPDATE / (ms * 1e-3) / 1e9
    grid = mqhip.fields_from_theirs(d)
    e = mqhip.Engine(grid)
    mzs = [475.0 + 2 * i for i in range(25)]
    jobs = [mqhip.job(mz * 5, 5.0, 0.0, 100.0, 500e3, 0.5, 100e3, 1000, 1000 + int(mz)) for mz in mzs]
    outs = e.run_batch(jobs, [0, 5, 0, 0], mqhip.Physics.legacy(1e-3 * 133.322), mqhip.PATH_BATCHED)
    e.close()
    tr = np.array([o.transmitted for o in outs])
    rows = list(csv.DictReader(open(os.path.join(ROOT, "tests", "golden", "mass-scan.csv"))))
    zz = {}
    for col in ("rust_legacy", "java_original"):
        ref = np.array([int(r[col]) for r in rows], float)
        pp = (tr + ref) / 2000.0
        se = np.sqrt(pp * (1 - pp) * 2 / 1000.0)
        z = np.where(se > 0, (tr - ref) / 1000.0 / np.where(se > 0, se, 1), 0.0)
        zz[col] = {"max_abs_z": float(np.max(np.abs(z))), "chi2_25": float(np.sum(z ** 2))}
    k = int(np.argmax(tr))
    out["fig4_on_regenerated"] = {"transmitted_per_1000": tr.tolist(), "peak_mz": mzs[k], "peak": int(tr[k]),
                                  "vs_published": zz}
    # CPU baseline: the sketch's sequential sweep (the oracle's restatement) on one core, a bounded sample of
    # sweeps at the full size
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # CPU baseline only
    nx, ny, nz, check = oracle.legacy_dims(10.0)
    phi, el = oracle.legacy_rasterize(0, 10.0)
    free = int((el[1:-1, 1:-1, 1:-1] == 0).sum())
    t0 = time.perf_counter()
    for _ in range(cpu_sweeps):
        oracle.legacy_sweep(phi, el)
    cs = time.perf_counter() - t0
    out["cpu_baseline"] = {"node_updates_per_s": cpu_sweeps * free / cs, "cores": 1, "kind": "port",
                             "sample": f"{cpu_sweeps} sweeps of the RF basis at 160 x 160 x 242"}
    out["speedup_vs_cpu_core"] = out["node_updates_per_s"] / out["cpu_baseline"]["node_updates_per_s"]
    log(f"legacy potentials: sweeps {[x['sweeps'] for x in st]}, {ms:.0f} ms, "
        f"{out['node_updates_per_s']:.3g} node-updates/s; Fig. 4 peak {tr[k]} at m/z {mzs[k]}")
    return out


def read_profile_value(name: str, key: str):
    """A figure from the committed rocprofv3 PMC summary of the worker kernel on this workload, if any."""
    path = os.path.join(ROOT, "profiles", f"k_batched_{name}_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":", 1)[1])
    return None


def roofline(name: str, steps: int, kms: float, launches: int) -> dict:
    """Roofline object of the worker over a timed run (steps ion-steps, kms ms of worker-kernel time from HIP events
    on the engine's stream, `launches` launches): k_batched, and k_tail4 (four lanes per ion) for the launches once
    at most 16 ions per wavefront are live.

    achieved = HBM bytes per ion-step measured by rocprofv3 (FETCH_SIZE x 2.0 + WRITE_SIZE, scaled by a copy
    kernel of known size; profiles/k_batched_<config>_traffic.md) x this run's ion-steps / this run's kernel time.
    The gather demand of the step (8 corners x 3 bases x 3 components x 4 B + a 4-B mask word = 292 B per
    ion-step) is reported beside it: the worker keeps each ion's cell record in registers, so that demand is not
    what crosses HBM and its ratio to the peak can exceed 1."""
    src = f"profiles/k_batched_{name}_traffic.md"
    per_step = read_profile_value(name, "traffic_bytes_per_ion_step")
    l2_per_step = read_profile_value(name, "l2_read_bytes_per_ion_step")
    t = kms * 1e-3
    demand = steps * BYTES_PER_STEP_FP32 / t / 1e9
    r = {"bound": "hbm", "kernel": "k_batched (+ k_tail4 in the last launches)", "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "ion_steps_per_launch": steps / max(launches, 1), "avg_launch_ms": kms / max(launches, 1),
         "ion_steps_per_s_kernel": steps / t}
    if per_step:
        r["achieved"] = per_step * steps / t / 1e9
        r["frac"] = r["achieved"] / HBM_PEAK_GBS
        r["traffic"] = per_step * steps / max(launches, 1)
        r["traffic_bytes_per_ion_step"] = per_step
        r["traffic_source"] = (f"{src}: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE per ion-step of this workload, "
                               "times this run's ion-steps")
    else:
        r["achieved"] = None
        r["frac"] = None
        r["traffic"] = None
        r["traffic_source"] = f"no {src}"
    r["gather_demand"] = {"bytes_per_ion_step": BYTES_PER_STEP_FP32, "gbs": demand, "ratio_to_peak": demand / HBM_PEAK_GBS}
    if l2_per_step:
        r["l2_read"] = {"bytes_per_ion_step": l2_per_step, "gbs": l2_per_step * steps / t / 1e9,
                        "source": f"{src}: TCP_TCC_READ_REQ x 128 B per ion-step"}
    r["valu_busy"] = read_profile_value(name, "valu_busy")
    r["valu_busy_source"] = f"{src}: 4 x SQ_INSTS_VALU / (1024 SIMDs x GRBM_GUI_ACTIVE / 8)"
    return r


def config_bench(name: str, steps: int, warmup: int, plain_steps: int, gather: bool = True, gate=None):
    """K timed run_batch calls of one workload on the batched FP32 path, then the plain FP64 path on the same
    inputs with per-job, per-class pulls between them. gate: override of the plate-basis voltage."""
    grid, jobs, dc, ph, desc = workload(name)
    if gate is not None:
        dc = [dc[0], gate, dc[2], dc[3]]
        desc += f"; gate {gate} V instead"
    n_traj = int(sum(j.n for j in jobs))
    eng = mqhip.Engine(grid)
    log(f"{name}: engine ready: {grid['nx']}x{grid['ny']}x{grid['nz']} grid, {len(jobs)} jobs, {n_traj} trajectories")
    for w in range(warmup):
        eng.run_batch(jobs, dc, ph, mqhip.PATH_BATCHED)
        log(f"{name}: warmup {w + 1}/{warmup}: {eng.stats()['ion_steps']} ion-steps")
    tot = kms = launches = 0
    outs = None
    t0 = time.perf_counter()
    for s in range(steps):
        outs = eng.run_batch(jobs, dc, ph, mqhip.PATH_BATCHED)
        st = eng.stats()
        tot += st["ion_steps"]
        kms += st["kernel_ms"]
        launches += st["launches"]
        log(f"{name}: step {s + 1}/{steps}: {st['ion_steps']} ion-steps, worker kernels {st['kernel_ms']:.1f} ms, "
            f"wall {st['wall_ms']:.1f} ms")
    wall = time.perf_counter() - t0
    out = {"value": tot / wall, "unit": "ion-steps/s", "ms_per_step": 1e3 * wall / max(steps, 1),
           "workload": desc, "trajectories": n_traj, "jobs": len(jobs), "grid": [grid["nx"], grid["ny"], grid["nz"]],
           "h_mm": grid["h"] * 1e3, "dc": dc, "roofline": roofline(name, tot, kms, launches)}
    if gather:
        # achievable random-gather bandwidth of this config's cell table (304-B records at a 320-B stride,
        # worker occupancy): the whole table, a 128 MB window (Infinity Cache sized), a 2.5 MB window (L2 sized)
        ncell = (grid["nx"] - 1) * (grid["ny"] - 1) * (grid["nz"] - 1)
        gath = {"table": eng.gather_probe(0) / 1e9, "window_128MB": eng.gather_probe(128 * 2**20 // 320) / 1e9,
                "window_2.5MB": eng.gather_probe(8192) / 1e9}
        r = out["roofline"]
        r["achievable_gather_gbs"] = gath
        r["table_bytes"] = ncell * 320
        if r.get("achieved"):
            r["frac_of_achievable_window_128MB"] = r["achieved"] / gath["window_128MB"]
        log(f"{name}: gather probe {gath}")
    totals = lambda os_: {k: sum(getattr(o, k) for o in os_) for k in ("crashed", "returned", "transmitted", "left",
                                                                       "timeout")}
    out["outcomes_batched"] = totals(outs)
    if plain_steps > 0:
        psteps = pk = 0
        pouts = None
        t0 = time.perf_counter()
        for s in range(plain_steps):
            pouts = eng.run_batch(jobs, dc, ph, mqhip.PATH_PLAIN)
            st = eng.stats()
            psteps += st["ion_steps"]
            pk += st["kernel_ms"]
            log(f"{name}: plain FP64 step {s + 1}/{plain_steps}: {st['ion_steps']} ion-steps, kernels "
                f"{st['kernel_ms']:.1f} ms, {st['launches']} launches")
        pwall = time.perf_counter() - t0
        pv = psteps / pwall
        pl = pulls(outs, pouts)
        out["plain_fp64"] = {
            "value": pv, "unit": "ion-steps/s", "ms_per_step": 1e3 * pwall / plain_steps,
            "gather_demand_gbs": psteps * BYTES_PER_STEP_FP64 / (pk * 1e-3) / 1e9,
            "outcomes_plain": totals(pouts),
            "per_job_pulls_within_3sigma": float(np.mean(np.abs(pl) <= 3.0)),
            "max_abs_pull": float(np.max(np.abs(pl))),
            "class_total_pulls": class_total_pulls(outs, pouts),
        }
        out["speedup_vs_plain_fp64"] = out["value"] / pv
    eng.close()
    return out, (grid, jobs, dc, ph)


def class_total_pulls(a, b):
    """Pull of each class total between two outcome lists (binomial over all trajectories)."""
    n1, n2 = sum(o.n for o in a), sum(o.n for o in b)
    out = {}
    for k in ("crashed", "returned", "transmitted", "left", "timeout"):
        k1, k2 = sum(getattr(o, k) for o in a), sum(getattr(o, k) for o in b)
        p1, p2 = k1 / max(n1, 1), k2 / max(n2, 1)
        se = np.sqrt(p1 * (1 - p1) / max(n1, 1) + p2 * (1 - p2) / max(n2, 1))
        out[k] = 0.0 if se == 0 else float((p1 - p2) / se)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-grad", action="store_true")
    ap.add_argument("--grad-fd", type=int, default=1, help="check the adjoint against central differences")
    ap.add_argument("--skip-tandem", action="store_true")
    ap.add_argument("--skip-refgeom", action="store_true")
    ap.add_argument("--skip-legacy", action="store_true")
    ap.add_argument("--skip-drivers", action="store_true", help="skip the live / Optimize / adaptive-scan timings")
    ap.add_argument("--skip-c3", action="store_true", help="skip the 1M-trajectory sweep sub-object")
    ap.add_argument("--skip-transmit", action="store_true", help="skip the transmitting mass-scan sub-object")
    ap.add_argument("--legacy-only", action="store_true",
                    help="run only the legacy-potential regeneration sub-bench and print its JSON object")
    ap.add_argument("--grad-only", action="store_true",
                    help="run only the differentiable-form sub-bench (one forward + backward, plus the "
                         "finite-difference checks unless --grad-fd 0) and print its JSON object")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one torch elementwise copy of 1 GiB (read 1 GiB, write 1 GiB) first, as the "
                         "known-size reference for rocprofv3 FETCH_SIZE / WRITE_SIZE")
    a = ap.parse_args()
    if a.gpus != 1:
        sys.exit("bench.py runs on one GPU only: the integrator has no multi-GPU path in this build "
                 "(ions are independent; see the sharding design). Use --gpus 1.")

    if a.legacy_only:
        print(json.dumps({"legacy_potentials": legacy_potentials_bench()}), flush=True)
        return
    if a.grad_only:
        print(json.dumps({"grad": grad_bench(bool(a.grad_fd))}), flush=True)
        return
    if a.copy_calibration:
        import torch
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
        log("copy calibration: 1 GiB read + 1 GiB written by one elementwise kernel")
    log(f"building the {a.config} workload")
    head, (grid, jobs, dc, ph) = config_bench(a.config, a.steps, a.warmup, a.plain_steps)

    result = {
        "metric": "ion-steps/s, batched FP32 vs plain FP64, matched outcome counts; grad evals/s",
        "value": head["value"], "unit": "ion-steps/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (FP64 position/time/phase)", "data": "synthetic",
        "config": {"workload": head["workload"], "trajectories": head["trajectories"], "jobs": head["jobs"],
                   "grid": head["grid"], "h_mm": head["h_mm"], "parallelism": "single GPU"},
        "roofline": head["roofline"],
        "outcomes_batched": head["outcomes_batched"],
    }
    if "plain_fp64" in head:
        result["plain_fp64"] = head["plain_fp64"]
        result["speedup_vs_plain_fp64"] = head["speedup_vs_plain_fp64"]
    if a.config == "c2" and not a.skip_transmit:
        # the same mass scan with a 0.02 V gate: at 5 V no ion crosses the gate (300 K ions carry ~0.04 eV), so
        # this variant is where the transmitted class of the FP32 and FP64 paths is compared
        sub, _ = config_bench("c2", 1, 1, a.plain_steps, gather=False, gate=0.02)
        result["c2_transmitting"] = sub
    if a.config != "c3" and not a.skip_c3:
        sub, _ = config_bench("c3", a.steps, a.warmup, a.plain_steps, gather=True)
        result["c3"] = sub
    if not a.skip_grad:
        result["grad"] = grad_bench(bool(a.grad_fd))
    if not a.skip_tandem:
        result["tandem_c5"] = tandem_bench()
    if not a.skip_refgeom:
        result["reference_geometry"] = reference_geometry_bench()
    if not a.skip_legacy:
        result["legacy_potentials"] = legacy_potentials_bench()
    if not a.skip_drivers:
        result["drivers"] = drivers_bench()
    if not a.skip_cpu:
        result["cpu_baseline"] = cpu_baseline(grid, jobs, dc, ph, a.cpu_seconds)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
