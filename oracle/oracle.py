"""ctypes loader of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline. The
product never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "multiq-workbench_amd"))
import mqhip  # noqa: E402  (C struct definitions of include/mq_engine.h)


class OrIon(C.Structure):
    _fields_ = [("x", C.c_double * 3), ("v", C.c_double * 3), ("phi_ac", C.c_double),
                ("steps", C.c_uint32), ("ncoll", C.c_uint32), ("cd", C.c_uint32), ("status", C.c_uint8),
                ("rng", C.c_uint32 * 4)]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    u32p = C.POINTER(C.c_uint32)
    lib.or_philox4x32_10.argtypes = [u32p, u32p, u32p]
    lib.or_sincos2pi.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.or_log.argtypes = [C.c_double]
    lib.or_log.restype = C.c_double
    lib.or_inject.argtypes = [C.POINTER(mqhip.Fields), C.POINTER(mqhip.Job), C.POINTER(mqhip.Physics), C.c_uint32,
                              C.POINTER(OrIon)]
    lib.or_run_batch.argtypes = [C.POINTER(mqhip.Fields), C.POINTER(mqhip.Job), C.c_size_t, C.POINTER(C.c_double),
                                 C.POINTER(mqhip.Physics), C.c_int, C.POINTER(OrIon), u32p]
    lib.or_run_range.argtypes = [C.POINTER(mqhip.Fields), C.POINTER(mqhip.Job), C.POINTER(C.c_double),
                                 C.POINTER(mqhip.Physics), C.c_uint32, C.c_uint32, C.c_int, C.POINTER(OrIon)]
    lib.or_summarize.argtypes = [C.POINTER(OrIon), u32p, C.c_size_t, C.c_size_t, C.c_double,
                                 C.POINTER(mqhip.Outcome)]
    lib.or_grad_reference.argtypes = [C.POINTER(mqhip.Fields), C.POINTER(mqhip.Job), C.POINTER(C.c_double),
                                      C.POINTER(mqhip.Physics), C.POINTER(mqhip.Objective), C.c_int,
                                      C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.or_sample.argtypes = [C.POINTER(C.c_double), C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double), C.c_int,
                              C.POINTER(C.c_double)]
    lib.or_xorshift64.argtypes = [C.POINTER(C.c_uint64)]
    lib.or_xorshift64.restype = C.c_uint64
    lib.or_xoshiro128ss.argtypes = [C.POINTER(C.c_uint32)]
    lib.or_xoshiro128ss.restype = C.c_uint32
    lib.or_inject_reference.argtypes = [C.POINTER(mqhip.Fields), C.POINTER(mqhip.Job), C.POINTER(mqhip.Physics),
                                        C.POINTER(OrIon)]
    lib.or_geom_default.argtypes = [C.POINTER(mqhip.Geom)]
    lib.or_geom_nz.argtypes = [C.POINTER(mqhip.Geom)]
    lib.or_geom_nz.restype = C.c_int
    lib.or_geom_classify_node.argtypes = [C.POINTER(mqhip.Geom), C.c_int, C.c_int, C.c_int]
    lib.or_geom_classify_node.restype = C.c_int
    lib.or_geom_classify.argtypes = [C.POINTER(mqhip.Geom), C.POINTER(C.c_uint8)]
    lib.or_basis_value.argtypes = [C.c_int, C.c_int, C.c_int]
    lib.or_basis_value.restype = C.c_double
    lib.or_laplace_sor.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_uint8), C.POINTER(C.c_double),
                                   C.c_double, C.c_double, C.c_uint32, C.POINTER(C.c_double),
                                   C.POINTER(C.c_uint32), C.POINTER(C.c_double)]
    lib.or_efield.argtypes = [C.POINTER(C.c_double), C.c_int, C.c_int, C.c_int, C.c_double, C.POINTER(C.c_double)]
    lib.or_tandem_run.argtypes = [C.POINTER(mqhip.Fields), C.POINTER(mqhip.Fields), C.POINTER(mqhip.TandemGeom),
                                  C.POINTER(mqhip.Job), C.c_size_t, C.POINTER(mqhip.TandemDrive),
                                  C.POINTER(mqhip.Physics), C.c_int, C.POINTER(OrIon), u32p,
                                  C.POINTER(C.c_uint8), C.POINTER(C.c_uint8), C.POINTER(C.c_uint64),
                                  C.POINTER(C.c_uint64)]
    f32p = C.POINTER(C.c_float)
    u8p = C.POINTER(C.c_uint8)
    lib.or_legacy_dims.argtypes = [C.c_float, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                   C.POINTER(C.c_int)]
    lib.or_legacy_rasterize.argtypes = [C.c_float, C.c_int, f32p, u8p]
    lib.or_legacy_sweep.argtypes = [C.c_int, C.c_int, C.c_int, f32p, u8p]
    lib.or_legacy_dif.argtypes = [C.c_float, C.c_float]
    lib.or_legacy_dif.restype = C.c_float
    lib.or_legacy_sweeps.argtypes = [C.c_int, C.c_int, C.c_int, f32p, u8p, C.POINTER(C.c_int), C.c_float,
                                     C.c_uint32, f32p, f32p]
    lib.or_legacy_sweeps.restype = C.c_uint32
    _lib = lib
    return lib


def philox(ctr, key) -> list[int]:
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    load().or_philox4x32_10(c, k, o)
    return list(o)


def sincos2pi(p: float) -> tuple[float, float]:
    s, c = C.c_double(), C.c_double()
    load().or_sincos2pi(p, C.byref(s), C.byref(c))
    return s.value, c.value


def log(x: float) -> float:
    return load().or_log(x)


def _ions_to_dict(ions, job=None) -> dict:
    a = np.ctypeslib.as_array(ions)
    out = {
        "x": a["x"][:, 0].copy(), "y": a["x"][:, 1].copy(), "z": a["x"][:, 2].copy(),
        "vx": a["v"][:, 0].copy(), "vy": a["v"][:, 1].copy(), "vz": a["v"][:, 2].copy(),
        "phi_ac": a["phi_ac"].copy(), "steps": a["steps"].copy(), "ncoll": a["ncoll"].copy(),
        "cd": a["cd"].copy(), "status": a["status"].copy(), "rng": a["rng"].copy(),
    }
    if job is not None:
        out["job"] = np.ctypeslib.as_array(job).copy()
    return out


def run_batch(grid: dict, jobs, dc, ph, nthreads: int = 8) -> dict:
    f, keep = mqhip.make_fields(grid)
    ja = mqhip.jobs_array(jobs)
    n = int(sum(j.n for j in jobs))
    ions = (OrIon * max(n, 1))()
    job = (C.c_uint32 * max(n, 1))()
    rc = load().or_run_batch(C.byref(f), ja, len(jobs), (C.c_double * 4)(*dc), C.byref(ph), nthreads, ions, job)
    if rc != 0:
        raise RuntimeError("oracle run failed")
    del keep
    d = _ions_to_dict(ions, job)
    return {k: v[:n] for k, v in d.items()}


def run_range(grid: dict, job, dc, ph, i0: int, i1: int, nthreads: int = 1) -> dict:
    f, keep = mqhip.make_fields(grid)
    ions = (OrIon * max(i1 - i0, 1))()
    rc = load().or_run_range(C.byref(f), C.byref(job), (C.c_double * 4)(*dc), C.byref(ph), i0, i1, nthreads, ions)
    if rc != 0:
        raise RuntimeError("oracle run failed")
    del keep
    d = _ions_to_dict(ions)
    return {k: v[: i1 - i0] for k, v in d.items()}


def inject(grid: dict, job, ph, ion: int) -> dict:
    f, keep = mqhip.make_fields(grid)
    o = OrIon()
    load().or_inject(C.byref(f), C.byref(job), C.byref(ph), ion, C.byref(o))
    del keep
    return {"x": list(o.x), "v": list(o.v), "phi_ac": o.phi_ac, "cd": o.cd, "status": o.status}


def xorshift64(state: int, n: int) -> list[int]:
    s = C.c_uint64(state)
    return [load().or_xorshift64(C.byref(s)) for _ in range(n)]


def xoshiro128ss(state, n: int) -> list[int]:
    s = (C.c_uint32 * 4)(*state)
    return [load().or_xoshiro128ss(s) for _ in range(n)]


def inject_reference(grid: dict, job, ph) -> dict:
    """Engine::inject with the reference's xorshift64 stream: every ion of one job."""
    f, keep = mqhip.make_fields(grid)
    n = int(job.n)
    ions = (OrIon * max(n, 1))()
    load().or_inject_reference(C.byref(f), C.byref(job), C.byref(ph), ions)
    del keep
    d = _ions_to_dict(ions)
    return {k: v[:n] for k, v in d.items()}


def summarize(res: dict, njobs: int, dt: float) -> list:
    n = len(res["status"])
    a = (OrIon * max(n, 1))()
    arr = np.ctypeslib.as_array(a)
    arr["status"][:n] = res["status"]
    arr["steps"][:n] = res["steps"]
    arr["ncoll"][:n] = res["ncoll"]
    job = np.ascontiguousarray(res["job"], np.uint32)
    out = (mqhip.Outcome * njobs)()
    load().or_summarize(a, job.ctypes.data_as(C.POINTER(C.c_uint32)), n, njobs, dt, out)
    return list(out)


def sample(F: np.ndarray, nx: int, ny: int, nz: int, g, nearest: int = 0) -> np.ndarray:
    F = np.ascontiguousarray(F, np.float64)
    gg = (C.c_double * 3)(*g)
    out = (C.c_double * 3)()
    load().or_sample(F.ctypes.data_as(C.POINTER(C.c_double)), nx, ny, nz, gg, nearest, out)
    return np.array(list(out))


def grad_reference(grid: dict, job, dc, ph, obj, nthreads: int = 8):
    """Dual-number (forward-mode) FP64 gradient: (J, grad[5], grad_state (6, n))."""
    f, keep = mqhip.make_fields(grid)
    J = C.c_double()
    g = (C.c_double * 5)()
    n = int(job.n)
    gs = np.zeros(6 * n)
    rc = load().or_grad_reference(C.byref(f), C.byref(job), (C.c_double * 4)(*dc), C.byref(ph), C.byref(obj),
                                  nthreads, C.byref(J), g, gs.ctypes.data_as(C.POINTER(C.c_double)))
    if rc != 0:
        raise RuntimeError("oracle gradient failed")
    del keep
    return J.value, np.array(list(g)), gs.reshape(6, n)


def grad_ions(grid: dict, job, dc, ph, obj, ions, nthreads: int = 8):
    """Dual-number gradient of single ions of a job: (Ji (m,), gi (5, m)) for the m listed
    ion indices, per ion (not divided by job.n)."""
    f, keep = mqhip.make_fields(grid)
    ions = np.ascontiguousarray(ions, np.uint32)
    m = len(ions)
    Ji = np.zeros(m)
    gi = np.zeros(5 * m)
    lib = load()
    lib.or_grad_ions.argtypes = [C.POINTER(mqhip.Fields), C.POINTER(mqhip.Job), C.POINTER(C.c_double),
                                 C.POINTER(mqhip.Physics), C.POINTER(mqhip.Objective), C.c_int,
                                 C.POINTER(C.c_uint32), C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    rc = lib.or_grad_ions(C.byref(f), C.byref(job), (C.c_double * 4)(*dc), C.byref(ph), C.byref(obj), nthreads,
                          ions.ctypes.data_as(C.POINTER(C.c_uint32)), m, Ji.ctypes.data_as(C.POINTER(C.c_double)),
                          gi.ctypes.data_as(C.POINTER(C.c_double)))
    if rc != 0:
        raise RuntimeError(f"oracle per-ion gradient failed ({rc})")
    del keep
    return Ji, gi.reshape(5, m)


# ---------------------------------------------------------------- field pipeline

def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def geom_default(**kw):
    g = mqhip.Geom()
    load().or_geom_default(C.byref(g))
    for k, v in kw.items():
        setattr(g, k, v)
    return g


def geom_classify(g) -> np.ndarray:
    """Geom::classify restated on the CPU: codes of shape (nz, ny, nx)."""
    nz = load().or_geom_nz(C.byref(g))
    out = np.zeros(g.n_xy * g.n_xy * nz, np.uint8)
    load().or_geom_classify(C.byref(g), out.ctypes.data_as(C.POINTER(C.c_uint8)))
    return out.reshape(nz, g.n_xy, g.n_xy)


def classify_node(g, i, j, k) -> int:
    return load().or_geom_classify_node(C.byref(g), i, j, k)


def basis_values(codes: np.ndarray, basis: int, ac_mode: int) -> np.ndarray:
    lut = np.array([load().or_basis_value(basis, c, ac_mode) for c in range(23)])
    return lut[codes]


def laplace_sor(mask, fixed, outer_bc, tol, max_iter=200000):
    nz, ny, nx = mask.shape
    m = np.ascontiguousarray(mask, np.uint8)
    f = np.ascontiguousarray(fixed, np.float64)
    phi = np.zeros(m.shape)
    it = C.c_uint32()
    res = C.c_double()
    rc = load().or_laplace_sor(nx, ny, nz, m.ctypes.data_as(C.POINTER(C.c_uint8)), _dp(f), outer_bc, tol, max_iter,
                               _dp(phi), C.byref(it), C.byref(res))
    if rc < 0:
        raise RuntimeError("oracle SOR failed")
    return phi, {"iterations": it.value, "residual": res.value, "converged": rc == 0}


def efield(phi, h):
    nz, ny, nx = phi.shape
    p = np.ascontiguousarray(phi, np.float64)
    e = np.zeros(p.shape + (3,))
    load().or_efield(_dp(p), nx, ny, nz, h, _dp(e))
    return e


# ---------------------------------------------------------------- the Java sketch's potentials

def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _u8(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


def legacy_dims(scl: float = 10.0):
    nx, ny, nz = C.c_int(), C.c_int(), C.c_int()
    ch = (C.c_int * 3)()
    load().or_legacy_dims(scl, C.byref(nx), C.byref(ny), C.byref(nz), ch)
    return nx.value, ny.value, nz.value, tuple(ch)


def legacy_rasterize(basis: int, scl: float = 10.0):
    """define_RF / define_DC / define_AC of every node: (phi float32, electrode uint8), shape (nz, ny, nx)."""
    nx, ny, nz, _ = legacy_dims(scl)
    phi = np.zeros((nz, ny, nx), np.float32)
    el = np.zeros((nz, ny, nx), np.uint8)
    load().or_legacy_rasterize(scl, basis, _fp(phi), _u8(el))
    return phi, el


def legacy_sweep(phi: np.ndarray, el: np.ndarray) -> None:
    """One sequential sweep in place."""
    nz, ny, nx = phi.shape
    assert phi.dtype == np.float32 and phi.flags.c_contiguous and el.dtype == np.uint8
    load().or_legacy_sweep(nx, ny, nz, _fp(phi), _u8(el))


def legacy_dif(prev: float, now: float) -> float:
    return load().or_legacy_dif(prev, now)


def legacy_sweeps(phi: np.ndarray, el: np.ndarray, check, precision: float, max_iter: int):
    """calculatePotential's loop on a copy of phi: (phi, sweeps, last dif, check point trace)."""
    nz, ny, nx = phi.shape
    p = np.ascontiguousarray(phi, np.float32).copy()
    e = np.ascontiguousarray(el, np.uint8)
    ch = (C.c_int * 3)(*check)
    dif = C.c_float()
    trace = np.zeros(max(max_iter, 1), np.float32)
    nit = load().or_legacy_sweeps(nx, ny, nz, _fp(p), _u8(e), ch, precision, max_iter, C.byref(dif), _fp(trace))
    return p, nit, dif.value, trace[:nit]


# ---------------------------------------------------------------- tandem batch

PORT_KEYS = ("entered", "returned", "forwarded", "revisits", "from_other", "reverse_entries", "lost")
TOTAL_KEYS = ("n", "split", "channel", "combine", "lost", "back", "detected", "ion_steps")


def tandem_run(cube: dict, port: dict, geom, jobs, drive, ph, nthreads: int = 8):
    """(per-port counts as dicts, totals dict, final states dict) of the FP64 tandem restatement."""
    fc, kc = mqhip.make_fields(cube)
    fp, kp = mqhip.make_fields(port)
    ja = mqhip.jobs_array(jobs)
    n = int(sum(j.n for j in jobs))
    ions = (OrIon * max(n, 1))()
    job = (C.c_uint32 * max(n, 1))()
    region = np.zeros(max(n, 1), np.uint8)
    last = np.zeros(max(n, 1), np.uint8)
    counts = np.zeros(9 * 24, np.uint64)
    totals = np.zeros(8, np.uint64)
    u8 = C.POINTER(C.c_uint8)
    u64 = C.POINTER(C.c_uint64)
    rc = load().or_tandem_run(C.byref(fc), C.byref(fp), C.byref(geom), ja, len(jobs), C.byref(drive), C.byref(ph),
                              nthreads, ions, job, region.ctypes.data_as(u8), last.ctypes.data_as(u8),
                              counts.ctypes.data_as(u64), totals.ctypes.data_as(u64))
    if rc != 0:
        raise RuntimeError("oracle tandem run failed")
    del kc, kp
    d = _ions_to_dict(ions, job)
    d = {k: v[:n] for k, v in d.items()}
    d["region"] = region[:n].copy()
    d["last"] = last[:n].copy()
    cs = []
    for p in range(9):
        a = counts[24 * p: 24 * (p + 1)]
        c = {k: int(a[i]) for i, k in enumerate(PORT_KEYS)}
        c["retry_from"] = [int(v) for v in a[7:16]]
        c["by_job"] = [int(v) for v in a[16:24]]
        cs.append(c)
    return cs, {k: int(totals[i]) for i, k in enumerate(TOTAL_KEYS)}, d
