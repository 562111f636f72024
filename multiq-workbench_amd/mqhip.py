"""ctypes binding of libmqhip.so (the C ABI in include/mq_engine.h).

Python mirror of chsim's batch API (simulation/chsim/src/channel.rs:233-506):
`Engine(fields)` replaces `Engine::new`, `Engine.run_batch` replaces
`Engine::run_batch`, `Engine.run_batch_ions` replaces `Engine::run_batch_ions`,
`summarize` replaces `summarize`, and `Physics.legacy` / `Physics.correct`
mirror the presets of channel.rs:260-272. Every compute call goes to the HIP
library; there is no CPU fallback: a missing library or GPU raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmqhip.so")

ALIVE, CRASHED, RETURNED, TRANSMITTED, LEFT, TIMEOUT = range(6)
PATH_BATCHED, PATH_PLAIN = 0, 1
NOALLOC_MAX_IONS, NOALLOC_MAX_JOBS = 65536, 4096  # MQ_NOALLOC_MAX_IONS / _JOBS
INJ_REFERENCE, INJ_GAUSSIAN = 0, 1
RNG_PHILOX, RNG_COMPAT = 0, 1

AMU = 1.6605390666e-27
QE = 1.602176634e-19
KB = 1.380649e-23


class Species(C.Structure):
    _fields_ = [("mass", C.c_double), ("z", C.c_double), ("ccs_n2", C.c_double)]


class RfAc(C.Structure):
    _fields_ = [("v_rf", C.c_double), ("f_rf", C.c_double), ("v_ac", C.c_double), ("f_ac", C.c_double)]


class Job(C.Structure):
    _fields_ = [("sp", Species), ("drive", RfAc), ("n", C.c_uint64), ("seed", C.c_uint64), ("dc_scale", C.c_double),
                ("pressure_pa", C.c_double)]


class Physics(C.Structure):
    _fields_ = [
        ("legacy_force", C.c_int32), ("legacy_collisions", C.c_int32), ("legacy_fields", C.c_int32),
        ("legacy_injection", C.c_int32), ("legacy_criteria", C.c_int32), ("injection", C.c_int32),
        ("pressure_pa", C.c_double), ("temp_k", C.c_double), ("dt", C.c_double), ("t_max", C.c_double),
        ("r_inject_mm", C.c_double), ("inj_sigma_mm", C.c_double), ("max_steps", C.c_uint64),
        ("rng", C.c_int32), ("pad_", C.c_int32),
    ]

    @staticmethod
    def legacy(pressure_pa: float) -> "Physics":
        """Physics::legacy (channel.rs:261-264)."""
        return Physics(1, 1, 1, 1, 1, INJ_REFERENCE, pressure_pa, 300.0, 1e-7, 1.0, 0.0, 0.0, 0)

    @staticmethod
    def correct(pressure_pa: float, dt_div: float = 50.0) -> "Physics":
        """Physics::correct (channel.rs:268-272): dt = T_RF / dt_div at 500 kHz."""
        return Physics(0, 0, 0, 0, 0, INJ_REFERENCE, pressure_pa, 300.0, 2e-6 / dt_div, 0.02, 2.0, 0.0, 0)

    def copy(self, **kw) -> "Physics":
        p = Physics()
        C.pointer(p)[0] = self
        for k, v in kw.items():
            setattr(p, k, v)
        return p


class Outcome(C.Structure):
    _fields_ = [
        ("n", C.c_uint64), ("crashed", C.c_uint64), ("returned", C.c_uint64), ("transmitted", C.c_uint64),
        ("left", C.c_uint64), ("timeout", C.c_uint64), ("mean_coll", C.c_double), ("mean_time", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}

    def frac(self, k: int) -> float:
        return k / self.n if self.n else 0.0


class Fields(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("h", C.c_double),
        ("axis_x", C.c_double), ("axis_y", C.c_double), ("z_inject", C.c_double),
        ("z_return", C.c_double), ("z_exit", C.c_double),
        ("z_return_legacy", C.c_double), ("z_exit_legacy", C.c_double),
        ("crash_z0_legacy", C.c_double), ("crash_z1_legacy", C.c_double),
        ("rf", C.POINTER(C.c_double)), ("ac", C.POINTER(C.c_double)), ("dc", C.POINTER(C.c_double) * 4),
        ("mask", C.POINTER(C.c_uint8)),
    ]


class IonSoA(C.Structure):
    _fields_ = [
        ("x", C.POINTER(C.c_double)), ("y", C.POINTER(C.c_double)), ("z", C.POINTER(C.c_double)),
        ("vx", C.POINTER(C.c_double)), ("vy", C.POINTER(C.c_double)), ("vz", C.POINTER(C.c_double)),
        ("steps", C.POINTER(C.c_uint32)), ("ncoll", C.POINTER(C.c_uint32)), ("status", C.POINTER(C.c_uint8)),
        ("phi_ac0", C.POINTER(C.c_double)), ("cd", C.POINTER(C.c_uint32)),
    ]


class Geom(C.Structure):
    """mq_geom (include/mq_fieldsolve.h): Geom of channel.rs:18-36, lengths in mm."""
    _fields_ = [("de", C.c_double), ("coef", C.c_double), ("q1_start", C.c_double), ("q1_len", C.c_double),
                ("alpha", C.c_double), ("gap", C.c_double), ("plate", C.c_double), ("hole_d", C.c_double),
                ("q2_len", C.c_double), ("stem_d", C.c_double), ("q2_flare_bug", C.c_int32), ("grow", C.c_double),
                ("tilt_mode", C.c_int32), ("ac_mode", C.c_int32), ("mis", (C.c_double * 4) * 4),
                ("n_xy", C.c_int32), ("h_mm", C.c_double)]

    @staticmethod
    def default(**kw) -> "Geom":
        """Geom::default (channel.rs:38-43), with keyword overrides."""
        g = Geom()
        load().mq_geom_default(C.byref(g))
        for k, v in kw.items():
            setattr(g, k, v)
        return g

    @staticmethod
    def s14(**kw) -> "Geom":
        """The corrected S14 geometry (s14_geom "code" variant, channel.rs:605-619): grown rods, parallel Q2."""
        return Geom.default(grow=0.4, q2_flare_bug=0, **kw)

    def dims(self):
        nx, ny, nz = C.c_int32(), C.c_int32(), C.c_int32()
        lay = (C.c_double * 8)()
        _check(load().mq_geom_dims(C.byref(self), C.byref(nx), C.byref(ny), C.byref(nz), lay))
        return nx.value, ny.value, nz.value, list(lay)


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("residual", C.c_double), ("converged", C.c_int32), ("ms", C.c_double)]


class ChannelFields(C.Structure):
    _fields_ = [("rf", C.POINTER(C.c_double)), ("ac", C.POINTER(C.c_double)), ("dc", C.POINTER(C.c_double) * 4),
                ("phi", C.POINTER(C.c_double) * 6), ("mask", C.POINTER(C.c_uint8)), ("stats", SolveStats * 6)]


class TandemGeom(C.Structure):
    """mq_tandem_geom (include/mq_tandem.h), metres in the cube grid's frame."""
    _fields_ = [("seam_z", C.c_double), ("mouth", (C.c_double * 2) * 9), ("ap_r", C.c_double),
                ("entrance", C.c_double * 2), ("l_ext", C.c_double)]


class TandemDrive(C.Structure):
    """mq_tandem_drive: one RF, one excitation amplitude, a tone per port (0 = not driven), DC sets."""
    _fields_ = [("v_rf", C.c_double), ("f_rf", C.c_double), ("v_ac", C.c_double), ("tones", C.c_double * 9),
                ("dc_split", C.c_double * 4), ("dc_combine", C.c_double * 4), ("dc_port", C.c_double * 4),
                ("bias_idle", C.c_double * 4)]


class PortCounts(C.Structure):
    """mq_port_counts (Counts, trapsim lib.rs:28-39)."""
    _fields_ = [("entered", C.c_uint64), ("returned", C.c_uint64), ("forwarded", C.c_uint64),
                ("revisits", C.c_uint64), ("from_other", C.c_uint64), ("reverse_entries", C.c_uint64),
                ("lost", C.c_uint64), ("retry_from", C.c_uint64 * 9), ("by_job", C.c_uint64 * 8)]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_[:7]}
        d["retry_from"] = list(self.retry_from)
        d["by_job"] = list(self.by_job)
        return d


class TandemTotals(C.Structure):
    _fields_ = [("n", C.c_uint64), ("split", C.c_uint64), ("channel", C.c_uint64), ("combine", C.c_uint64),
                ("lost", C.c_uint64), ("back", C.c_uint64), ("detected", C.c_uint64), ("ion_steps", C.c_uint64)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


# numpy view of mq_chion (ChIon, channel.rs:292-294): 64 bytes per ion.
CHION_DTYPE = np.dtype([("p", "<i4", 4), ("v", "<f4", 4), ("aux", "<f4", 4), ("rng", "<u4", 4)])
POS_SCALE = 1.6e10


class RunStats(C.Structure):
    _fields_ = [("ion_steps", C.c_uint64), ("n_ions", C.c_uint64), ("launches", C.c_uint32),
                ("kernel_ms", C.c_double), ("wall_ms", C.c_double), ("quad_launches", C.c_uint32),
                ("plain_recs", C.c_uint32)]


class Tune(C.Structure):
    _fields_ = [("block", C.c_int32), ("waves_per_cu", C.c_int32), ("refill_period", C.c_int32),
                ("chunk", C.c_int32), ("chunk_cap", C.c_int32), ("resort", C.c_int32), ("sync_every", C.c_int32),
                ("graph", C.c_int32), ("coll_batch", C.c_int32), ("tail_lanes", C.c_int32),
                ("tail_quads", C.c_int32), ("plain_recs", C.c_int32)]


class Objective(C.Structure):
    _fields_ = [("nsteps", C.c_uint32), ("ckpt_every", C.c_uint32), ("z_exit", C.c_double), ("scale", C.c_double),
                ("r_leak", C.c_double), ("w_leak", C.c_double)]


class PeakFit(C.Structure):
    _fields_ = [("f0_khz", C.c_double), ("f0_err", C.c_double), ("fwhm_khz", C.c_double), ("fwhm_err", C.c_double),
                ("amp", C.c_double), ("base", C.c_double), ("r_f", C.c_double), ("r_f_err", C.c_double),
                ("chi2_red", C.c_double), ("converged", C.c_int32), ("npts", C.c_int32)]


class OptimizeOpts(C.Structure):
    _fields_ = [("f_lo_khz", C.c_double), ("f_hi_khz", C.c_double), ("f_step_khz", C.c_double), ("n", C.c_uint64),
                ("min_r", C.c_double), ("dlnf_dlnmz", C.c_double), ("seed", C.c_uint64), ("leak_n", C.c_uint64)]


class OptimizeRow(C.Structure):
    _fields_ = [("v_ac", C.c_double), ("lo_khz", C.c_double), ("hi_khz", C.c_double), ("step_khz", C.c_double),
                ("ions_per_point", C.c_uint64), ("fit", PeakFit), ("r_mz", C.c_double), ("r_mz_err", C.c_double),
                ("f_best_khz", C.c_double), ("t_best", C.c_double), ("t_best_err", C.c_double),
                ("t_k", C.c_uint64), ("t_n", C.c_uint64), ("leak_k", C.c_uint64), ("leak_n", C.c_uint64),
                ("leakage", C.c_double), ("leakage_err", C.c_double), ("leakage_ul95", C.c_double),
                ("refined", C.c_int32), ("issue", C.c_int32)]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "fit"}
        d["fit"] = {k: getattr(self.fit, k) for k, _ in PeakFit._fields_}
        return d


class GradOpts(C.Structure):
    _fields_ = [("mask", C.c_uint32), ("max_iter", C.c_uint32), ("lo", C.c_double * 5), ("hi", C.c_double * 5),
                ("step0", C.c_double * 5), ("tol", C.c_double)]


class LiveOpts(C.Structure):
    _fields_ = [("block_steps", C.c_uint32), ("capacity", C.c_uint32), ("inj_rate", C.c_double),
                ("recirculate", C.c_int32), ("max_tags", C.c_int32), ("residence_s", C.c_double),
                ("seed", C.c_uint64)]


class LiveState(C.Structure):
    _fields_ = [("step", C.c_uint64), ("alive", C.c_uint64), ("injected", C.c_uint64), ("reentered", C.c_uint64),
                ("dropped", C.c_uint64), ("finished", C.c_uint64), ("ion_steps", C.c_uint64)]


class LiveSweepCfg(C.Structure):
    _fields_ = [("f_lo_khz", C.c_double), ("f_hi_khz", C.c_double), ("step_khz", C.c_double),
                ("ions_per_point", C.c_uint32), ("sweeps", C.c_uint32)]


class LiveSweepState(C.Structure):
    _fields_ = [("npoints", C.c_uint32), ("point", C.c_uint32), ("sweeps_done", C.c_uint32), ("running", C.c_int32),
                ("draining", C.c_int32), ("pad_", C.c_int32), ("progress", C.c_double)]


LIVE_MAX_POINTS = 101


class ScanPoint(C.Structure):
    _fields_ = [("f_khz", C.c_double), ("transmitted", C.c_double), ("crashed", C.c_double), ("n", C.c_uint64)]


class LegacyOpts(C.Structure):
    """mq_legacy_opts: the Java sketch's potential settings (scl nodes/mm, stop threshold %, sweep cap)."""
    _fields_ = [("scl", C.c_float), ("precision_pct", C.c_float), ("max_iter", C.c_uint32), ("pad_", C.c_int32)]

    @classmethod
    def default(cls, **kw) -> "LegacyOpts":
        o = cls()
        load().mq_legacy_default(C.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        return o


class LegacyStats(C.Structure):
    _fields_ = [("sweeps", C.c_uint32), ("dif_pct", C.c_float), ("check", C.c_int32 * 3), ("pad_", C.c_int32),
                ("node_updates", C.c_uint64), ("ms", C.c_double)]

    def as_dict(self) -> dict:
        return {"sweeps": self.sweeps, "dif_pct": self.dif_pct, "check": list(self.check),
                "node_updates": self.node_updates, "ms": self.ms}


LEGACY_RF, LEGACY_DC, LEGACY_AC = 0, 1, 2


# Exported symbols of include/mq_engine.h with their ctypes signatures.
SIGNATURES = {
    "mq_last_error": (C.c_char_p, []),
    "mq_engine_create": (C.c_int, [C.POINTER(Fields), C.POINTER(C.c_void_p)]),
    "mq_engine_destroy": (None, [C.c_void_p]),
    "mq_run_batch": (C.c_int, [C.c_void_p, C.POINTER(Job), C.c_size_t, C.POINTER(C.c_double), C.POINTER(Physics),
                               C.c_int32, C.POINTER(Outcome)]),
    "mq_run_batch_ions": (C.c_int, [C.c_void_p, C.POINTER(Job), C.c_size_t, C.POINTER(C.c_double),
                                    C.POINTER(Physics), C.c_int32, C.POINTER(IonSoA), C.POINTER(C.c_uint32)]),
    "mq_run_batch_ions_noalloc": (C.c_int, [C.c_void_p, C.POINTER(Job), C.c_size_t, C.POINTER(C.c_double),
                                            C.POINTER(Physics), C.c_int32, C.POINTER(IonSoA),
                                            C.POINTER(C.c_uint32)]),
    "mq_field_probe": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_int32, C.c_double, C.c_double, C.c_double,
                                 C.c_size_t, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                 C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                 C.POINTER(C.c_uint8)]),
    "mq_fit_peak": (C.c_int, [C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_uint64), C.c_size_t,
                              C.POINTER(PeakFit)]),
    "mq_frac_sigma": (C.c_int, [C.c_uint64, C.c_uint64, C.POINTER(C.c_double)]),
    "mq_ratio": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.POINTER(C.c_double),
                           C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "mq_optimize_grid": (C.c_int, [C.c_void_p, C.POINTER(Species), C.POINTER(Species), C.POINTER(RfAc),
                                   C.POINTER(C.c_double),
                                   C.POINTER(Physics), C.c_int32, C.POINTER(C.c_double), C.c_size_t,
                                   C.POINTER(OptimizeOpts), C.POINTER(OptimizeRow), C.POINTER(C.c_int32),
                                   C.POINTER(OptimizeRow), C.POINTER(C.c_int32)]),
    "mq_optimize_grad": (C.c_int, [C.c_void_p, C.POINTER(Job), C.POINTER(C.c_double), C.POINTER(Physics),
                                   C.POINTER(Objective), C.POINTER(GradOpts), C.POINTER(C.c_double),
                                   C.POINTER(C.c_uint32), C.POINTER(C.c_double)]),
    "mq_live_create": (C.c_int, [C.c_void_p, C.POINTER(Species), C.POINTER(C.c_double), C.c_size_t,
                                 C.POINTER(RfAc), C.POINTER(C.c_double), C.POINTER(Physics), C.POINTER(LiveOpts),
                                 C.POINTER(C.c_void_p)]),
    "mq_live_destroy": (None, [C.c_void_p]),
    "mq_live_inject": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32]),
    "mq_live_run": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(LiveState)]),
    "mq_live_set_ac": (C.c_int, [C.c_void_p, C.c_double, C.c_int32]),
    "mq_live_tally": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Outcome)]),
    "mq_live_sweep_start": (C.c_int, [C.c_void_p, C.POINTER(LiveSweepCfg)]),
    "mq_live_sweep_stop": (C.c_int, [C.c_void_p]),
    "mq_live_sweep": (C.c_int, [C.c_void_p, C.POINTER(LiveSweepState), C.POINTER(C.c_double),
                                C.POINTER(C.c_uint64)]),
    "mq_live_slots": (C.c_int, [C.c_void_p, C.POINTER(IonSoA), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                C.POINTER(C.c_uint64)]),
    "mq_summarize": (C.c_int, [C.POINTER(C.c_uint8), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                               C.POINTER(C.c_uint32), C.c_size_t, C.c_size_t, C.c_double, C.POINTER(Outcome)]),
    "mq_last_stats": (C.c_int, [C.c_void_p, C.POINTER(RunStats)]),
    "mq_set_tune": (C.c_int, [C.c_void_p, C.POINTER(Tune)]),
    "mq_get_tune": (C.c_int, [C.c_void_p, C.POINTER(Tune)]),
    "mq_autotune": (C.c_int, [C.c_void_p, C.POINTER(Job), C.c_size_t, C.POINTER(C.c_double), C.POINTER(Physics),
                              C.c_uint64, C.c_char_p, C.c_int32, C.POINTER(Tune), C.POINTER(C.c_double)]),
    "mq_forward": (C.c_int, [C.c_void_p, C.POINTER(Job), C.POINTER(C.c_double), C.POINTER(Physics),
                             C.POINTER(Objective), C.POINTER(C.c_double)]),
    "mq_backward": (C.c_int, [C.c_void_p, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "mq_forward_per_ion": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "mq_backward_per_ion": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "mq_forward_final": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "mq_cached_batch": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(Job), C.c_size_t, C.POINTER(C.c_double),
                                  C.POINTER(Physics), C.c_int32, C.c_char_p, C.POINTER(Outcome),
                                  C.POINTER(C.c_size_t)]),
    "mq_adaptive_scan": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(Species), C.POINTER(RfAc), C.POINTER(C.c_double),
                                   C.POINTER(Physics), C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_uint64,
                                   C.c_double, C.c_uint64, C.c_int32, C.c_char_p, C.POINTER(ScanPoint), C.c_size_t,
                                   C.POINTER(C.c_size_t)]),
    "mq_adaptive_fine_points": (C.c_int, [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_size_t, C.c_uint64,
                                          C.c_double, C.c_double, C.POINTER(C.c_double), C.c_size_t,
                                          C.POINTER(C.c_size_t)]),
    "mq_scan_key": (C.c_int, [C.c_char_p, C.POINTER(Job), C.POINTER(C.c_double), C.POINTER(Physics), C.c_int32,
                              C.c_char_p, C.c_size_t]),
    "mq_format_f64": (C.c_int, [C.c_double, C.c_int32, C.c_char_p, C.c_size_t]),
    "mq_export_chion": (C.c_int, [C.POINTER(IonSoA), C.POINTER(C.c_uint32), C.c_size_t, C.POINTER(Job), C.c_size_t,
                                  C.POINTER(Physics), C.c_void_p]),
    "mq_summarize_chion": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.c_size_t, C.c_size_t,
                                     C.POINTER(Outcome)]),
    "mq_fnv1a64": (C.c_uint64, [C.c_void_p, C.c_size_t]),
    "mq_geom_default": (None, [C.POINTER(Geom)]),
    "mq_geom_dims": (C.c_int, [C.POINTER(Geom), C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                               C.POINTER(C.c_double)]),
    "mq_geom_fields_params": (C.c_int, [C.POINTER(Geom), C.POINTER(Fields)]),
    "mq_geom_classify": (C.c_int, [C.POINTER(Geom), C.POINTER(C.c_uint8)]),
    "mq_laplace_solve": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_uint8), C.POINTER(C.c_double),
                                   C.c_double, C.c_double, C.c_uint32, C.POINTER(C.c_double),
                                   C.POINTER(SolveStats)]),
    "mq_efield": (C.c_int, [C.POINTER(C.c_double), C.c_int32, C.c_int32, C.c_int32, C.c_double,
                            C.POINTER(C.c_double)]),
    "mq_fields_build": (C.c_int, [C.POINTER(Geom), C.c_double, C.c_char_p, C.POINTER(ChannelFields)]),
    "mq_field_cache_name": (C.c_int, [C.POINTER(Geom), C.c_int32, C.c_char_p, C.c_size_t]),
    "mq_inject_reference": (C.c_int, [C.POINTER(Fields), C.POINTER(Job), C.POINTER(Physics), C.POINTER(IonSoA),
                                      C.POINTER(C.c_uint32)]),
    "mq_gather_probe": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.POINTER(C.c_double)]),
    "mq_tandem_create": (C.c_int, [C.POINTER(Fields), C.POINTER(Fields), C.POINTER(TandemGeom),
                                   C.POINTER(C.c_void_p)]),
    "mq_tandem_destroy": (None, [C.c_void_p]),
    "mq_tandem_run": (C.c_int, [C.c_void_p, C.POINTER(Job), C.c_size_t, C.POINTER(TandemDrive), C.POINTER(Physics),
                                C.c_int32, C.POINTER(PortCounts), C.POINTER(TandemTotals)]),
    "mq_tandem_run_ions": (C.c_int, [C.c_void_p, C.POINTER(Job), C.c_size_t, C.POINTER(TandemDrive),
                                     C.POINTER(Physics), C.c_int32, C.POINTER(PortCounts), C.POINTER(TandemTotals),
                                     C.POINTER(IonSoA), C.POINTER(C.c_uint8), C.POINTER(C.c_uint8),
                                     C.POINTER(C.c_uint32)]),
    "mq_tandem_last_stats": (C.c_int, [C.c_void_p, C.POINTER(RunStats)]),
    "mq_tandem_set_plain_recs": (C.c_int, [C.c_void_p, C.c_int32]),
    "mq_legacy_default": (None, [C.POINTER(LegacyOpts)]),
    "mq_legacy_dims": (C.c_int, [C.POINTER(LegacyOpts), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                 C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "mq_legacy_rasterize": (C.c_int, [C.POINTER(LegacyOpts), C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_uint8)]),
    "mq_legacy_sweeps": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_uint8),
                                   C.POINTER(C.c_int32), C.c_float, C.c_uint32, C.POINTER(LegacyStats)]),
    "mq_legacy_potential": (C.c_int, [C.POINTER(LegacyOpts), C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_uint8),
                                      C.POINTER(LegacyStats)]),
    "mq_legacy_export": (C.c_int, [C.POINTER(LegacyOpts), C.c_char_p, C.POINTER(LegacyStats)]),
    "mq_fields_from_theirs": (C.c_int, [C.c_char_p, C.POINTER(ChannelFields)]),
}

_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load libmqhip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built (run `make -C {HERE}`)")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class MqError(RuntimeError):
    pass


def _check(rc: int) -> None:
    if rc != 0:
        msg = load().mq_last_error().decode(errors="replace")
        raise MqError(f"mq error {rc}: {msg}")


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_fields(grid: dict) -> tuple[Fields, list]:
    """Build an mq_fields struct from a grid dict (see synth.py); returns the
    struct and the list of arrays it points into (keep it alive)."""
    keep = []

    def arr(a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a

    f = Fields()
    f.nx, f.ny, f.nz = grid["nx"], grid["ny"], grid["nz"]
    f.h = grid["h"]
    for k in ("axis_x", "axis_y", "z_inject", "z_return", "z_exit", "z_return_legacy", "z_exit_legacy",
              "crash_z0_legacy", "crash_z1_legacy"):
        setattr(f, k, grid[k])
    f.rf = _dptr(arr(grid["rf"], np.float64))
    f.ac = _dptr(arr(grid["ac"], np.float64)) if grid.get("ac") is not None else C.POINTER(C.c_double)()
    for b in range(4):
        d = grid["dc"][b]
        f.dc[b] = _dptr(arr(d, np.float64)) if d is not None else C.POINTER(C.c_double)()
    m = grid.get("mask")
    f.mask = arr(m, np.uint8).ctypes.data_as(C.POINTER(C.c_uint8)) if m is not None else C.POINTER(C.c_uint8)()
    return f, keep


def jobs_array(jobs) -> C.Array:
    arr = (Job * len(jobs))()
    for i, j in enumerate(jobs):
        arr[i] = j
    return arr


def job(mass, z, ccs, v_rf, f_rf, v_ac, f_ac, n, seed, dc_scale=1.0, pressure_pa=0.0) -> Job:
    return Job(Species(mass, z, ccs), RfAc(v_rf, f_rf, v_ac, f_ac), n, seed, dc_scale, pressure_pa)


class Engine:
    """Engine::new (channel.rs:327): uploads the grids once."""

    def __init__(self, grid: dict):
        lib = load()
        f, keep = make_fields(grid)
        h = C.c_void_p()
        _check(lib.mq_engine_create(C.byref(f), C.byref(h)))
        del keep
        self._h = h
        self.grid_meta = {k: grid[k] for k in ("nx", "ny", "nz", "h")}

    def close(self):
        if getattr(self, "_h", None):
            load().mq_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run_batch(self, jobs, dc, ph: Physics, path: int = PATH_BATCHED) -> list[Outcome]:
        ja = jobs_array(jobs)
        out = (Outcome * len(jobs))()
        dcv = (C.c_double * 4)(*dc)
        _check(load().mq_run_batch(self._h, ja, len(jobs), dcv, C.byref(ph), path, out))
        return list(out)

    def run_batch_ions(self, jobs, dc, ph: Physics, path: int = PATH_BATCHED, noalloc: bool = False) -> dict:
        """Engine::run_batch_ions (channel.rs:458-475); noalloc=True calls the entry point that allocates no
        device memory (mq_run_batch_ions_noalloc)."""
        ja = jobs_array(jobs)
        n = int(sum(j.n for j in jobs))
        o = {k: np.zeros(n, np.float64) for k in ("x", "y", "z", "vx", "vy", "vz")}
        o["steps"] = np.zeros(n, np.uint32)
        o["ncoll"] = np.zeros(n, np.uint32)
        o["status"] = np.zeros(n, np.uint8)
        o["job"] = np.zeros(n, np.uint32)
        o["phi_ac0"] = np.zeros(n, np.float64)
        o["cd"] = np.zeros(n, np.uint32)
        soa = soa_of(o)
        dcv = (C.c_double * 4)(*dc)
        fn = load().mq_run_batch_ions_noalloc if noalloc else load().mq_run_batch_ions
        _check(fn(self._h, ja, len(jobs), dcv, C.byref(ph), path, C.byref(soa),
                  o["job"].ctypes.data_as(C.POINTER(C.c_uint32))))
        return o

    def stats(self) -> dict:
        s = RunStats()
        _check(load().mq_last_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def field_probe(self, dc, x, y, z, a_rf=1.0, s_dc=1.0, a_ac=1.0, nearest=0):
        """The batched path's FP32 field at positions [m] (mq_field_probe): (E (n, 3) float32, crash bits)."""
        x, y, z = (np.ascontiguousarray(v, np.float64) for v in (x, y, z))
        n = len(x)
        e = np.zeros((3, n), np.float32)
        cr = np.zeros(n, np.uint8)
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        _check(load().mq_field_probe(self._h, (C.c_double * 4)(*dc), int(nearest), a_rf, s_dc, a_ac, n, _dptr(x),
                                     _dptr(y), _dptr(z), fp(e[0]), fp(e[1]), fp(e[2]),
                                     cr.ctypes.data_as(C.POINTER(C.c_uint8))))
        return e.T.copy(), cr

    def gather_probe(self, window_cells: int = 0, iters: int = 64) -> float:
        """Achievable random-gather bandwidth [B/s] of the cell table (whole table, or its first window_cells)."""
        b = C.c_double()
        _check(load().mq_gather_probe(self._h, window_cells, iters, C.byref(b)))
        return b.value

    def set_tune(self, **kw) -> None:
        t = Tune()
        _check(load().mq_get_tune(self._h, C.byref(t)))
        for k, v in kw.items():
            setattr(t, k, v)
        _check(load().mq_set_tune(self._h, C.byref(t)))

    def forward(self, j: Job, dc, ph: Physics, obj: Objective) -> float:
        """Forward op of the differentiable form: mean smoothed objective J."""
        J = C.c_double()
        _check(load().mq_forward(self._h, C.byref(j), (C.c_double * 4)(*dc), C.byref(ph), C.byref(obj), C.byref(J)))
        return J.value

    def backward(self, dJ: float = 1.0, state: bool = False):
        """Backward op: dJ * dJ/d(V_rf, f_rf, V_ac, f_ac, dc_scale) [and dJ/d(initial state), shape (6, n)]."""
        g = (C.c_double * 5)()
        n = self.stats()["n_ions"]
        gs = np.zeros(6 * n) if state else None
        _check(load().mq_backward(self._h, dJ, g, _dptr(gs) if state else None))
        return (np.array(list(g)), gs.reshape(6, n)) if state else np.array(list(g))

    def forward_per_ion(self) -> np.ndarray:
        out = np.zeros(self.stats()["n_ions"])
        _check(load().mq_forward_per_ion(self._h, _dptr(out)))
        return out

    def forward_final(self) -> np.ndarray:
        """Final positions (3, n) [m] of the last forward."""
        n = self.stats()["n_ions"]
        out = np.zeros((3, n))
        _check(load().mq_forward_final(self._h, _dptr(out[0]), _dptr(out[1]), _dptr(out[2])))
        return out

    def backward_per_ion(self, n: int) -> np.ndarray:
        out = np.zeros(5 * n)
        _check(load().mq_backward_per_ion(self._h, _dptr(out)))
        return out.reshape(5, n)

    def autotune(self, jobs, dc, ph: Physics, sample_ions: int = 0, cache_path: str | None = None,
                 force: bool = False) -> tuple[dict, float]:
        """Time candidate tunings on a subset of `jobs`; keep and return the fastest."""
        t = Tune()
        r = C.c_double()
        _check(load().mq_autotune(self._h, jobs_array(jobs), len(jobs), (C.c_double * 4)(*dc), C.byref(ph), sample_ions,
                                  cache_path.encode() if cache_path else None, int(force), C.byref(t), C.byref(r)))
        return {k: getattr(t, k) for k, _ in t._fields_}, r.value

    def get_tune(self) -> dict:
        t = Tune()
        _check(load().mq_get_tune(self._h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}


class Live:
    """A live ion population on an engine (mq_live_*; mqviz Sim::block, engine.rs:448-538)."""

    def __init__(self, eng: "Engine", species, weights, rf: RfAc, dc, ph: Physics, block_steps=128, capacity=65536,
                 inj_rate=0.0, recirculate=False, max_tags=1, residence_s=0.0, seed=1):
        sp = (Species * len(species))(*species)
        w = None if weights is None else (C.c_double * len(species))(*weights)
        o = LiveOpts(block_steps, capacity, inj_rate, int(recirculate), max_tags, residence_s, seed)
        h = C.c_void_p()
        _check(load().mq_live_create(eng._h, sp, w, len(species), C.byref(rf), (C.c_double * 4)(*dc), C.byref(ph),
                                     C.byref(o), C.byref(h)))
        self._h, self.nsp, self.capacity, self._eng = h, len(species), capacity, eng

    def close(self):
        if getattr(self, "_h", None):
            load().mq_live_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def inject(self, species: int, n: int) -> None:
        _check(load().mq_live_inject(self._h, species, n))

    def run(self, nblocks: int) -> dict:
        st = LiveState()
        _check(load().mq_live_run(self._h, nblocks, C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def set_ac(self, f_ac: float, tag: int) -> None:
        _check(load().mq_live_set_ac(self._h, f_ac, tag))

    def tally(self, tag: int = 0) -> list:
        out = (Outcome * self.nsp)()
        _check(load().mq_live_tally(self._h, tag, out))
        return list(out)

    def sweep_start(self, f_lo_khz: float, f_hi_khz: float, step_khz: float, ions_per_point: int,
                    sweeps: int = 1) -> None:
        c = LiveSweepCfg(f_lo_khz, f_hi_khz, step_khz, ions_per_point, sweeps)
        _check(load().mq_live_sweep_start(self._h, C.byref(c)))

    def sweep_stop(self) -> None:
        _check(load().mq_live_sweep_stop(self._h))

    def sweep(self) -> dict:
        """The last sweep: its state, frequencies [kHz] and counts acc[point, species, (trans, crash, return,
        other)]."""
        st = LiveSweepState()
        f = np.zeros(LIVE_MAX_POINTS, np.float64)
        acc = np.zeros((LIVE_MAX_POINTS, self.nsp, 4), np.uint64)
        _check(load().mq_live_sweep(self._h, C.byref(st), _dptr(f), acc.ctypes.data_as(C.POINTER(C.c_uint64))))
        d = {k: getattr(st, k) for k, _ in st._fields_ if k != "pad_"}
        d["f_khz"], d["acc"] = f[:st.npoints], acc[:st.npoints]
        return d

    def slots(self) -> dict:
        n = self.capacity
        o = {k: np.zeros(n, np.float64) for k in ("x", "y", "z", "vx", "vy", "vz", "phi_ac0")}
        o["steps"] = np.zeros(n, np.uint32)
        o["ncoll"] = np.zeros(n, np.uint32)
        o["status"] = np.zeros(n, np.uint8)
        o["cd"] = np.zeros(n, np.uint32)
        o["sid"] = np.zeros(n, np.uint32)
        o["species"] = np.zeros(n, np.uint32)
        o["birth"] = np.zeros(n, np.uint64)
        soa = soa_of(o)
        u32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint32))
        _check(load().mq_live_slots(self._h, C.byref(soa), u32(o["sid"]), u32(o["species"]),
                                    o["birth"].ctypes.data_as(C.POINTER(C.c_uint64))))
        return o


def fit_peak(f_khz, t, n) -> dict:
    """Weighted Gaussian-plus-baseline fit of a transmission curve (mq_fit_peak)."""
    f = np.ascontiguousarray(f_khz, np.float64)
    tt = np.ascontiguousarray(t, np.float64)
    nn = np.ascontiguousarray(np.broadcast_to(np.asarray(n, np.uint64), f.shape), np.uint64)
    out = PeakFit()
    _check(load().mq_fit_peak(_dptr(f), _dptr(tt), nn.ctypes.data_as(C.POINTER(C.c_uint64)), len(f), C.byref(out)))
    return {k: getattr(out, k) for k, _ in PeakFit._fields_}


def frac_sigma(k: int, n: int) -> float:
    """Wilson-score 1-sigma error of k/n (mq_frac_sigma; mqviz stats.rs:17-22)."""
    v = C.c_double()
    _check(load().mq_frac_sigma(k, n, C.byref(v)))
    return v.value


def ratio(ka: int, na: int, kb: int, nb: int) -> tuple[float, float, float]:
    """(ka/na)/(kb/nb), its error and 95 % upper limit (mq_ratio; mqviz stats.rs:51-69)."""
    v, e, u = C.c_double(), C.c_double(), C.c_double()
    _check(load().mq_ratio(ka, na, kb, nb, C.byref(v), C.byref(e), C.byref(u)))
    return v.value, e.value, u.value


def optimize_grid(eng: "Engine", sp: Species, rf: RfAc, dc, ph: Physics, v_ac, f_lo_khz: float, f_hi_khz: float,
                  f_step_khz: float, n: int, min_r: float, seed: int, dlnf_dlnmz: float = 0.0,
                  path: int = PATH_BATCHED, leak: Species | None = None, leak_n: int = 0) -> dict:
    """The Optimize grid search (research.rs:133-163) on the engine (mq_optimize_grid)."""
    va = np.ascontiguousarray(v_ac, np.float64)
    o = OptimizeOpts(f_lo_khz, f_hi_khz, f_step_khz, n, min_r, dlnf_dlnmz, seed, leak_n)
    rows = (OptimizeRow * len(va))()
    conf = OptimizeRow()
    best, confirmed = C.c_int32(), C.c_int32()
    _check(load().mq_optimize_grid(eng._h, C.byref(sp), C.byref(leak) if leak is not None else None, C.byref(rf),
                                   (C.c_double * 4)(*dc), C.byref(ph), path,
                                   _dptr(va), len(va), C.byref(o), rows, C.byref(best), C.byref(conf),
                                   C.byref(confirmed)))
    return {"rows": [r.as_dict() for r in rows], "best": best.value,
            "confirmation": conf.as_dict() if best.value >= 0 else None, "confirmed": bool(confirmed.value)}


def optimize_grad(eng: "Engine", j: Job, dc, ph: Physics, obj: Objective, mask: int, lo, hi, step0,
                  max_iter: int = 40, tol: float = 1e-3) -> dict:
    """Projected gradient ascent of the differentiable objective (mq_optimize_grad)."""
    jj = Job()
    C.pointer(jj)[0] = j
    o = GradOpts(mask, max_iter, (C.c_double * 5)(*lo), (C.c_double * 5)(*hi), (C.c_double * 5)(*step0), tol)
    J = C.c_double()
    ne = C.c_uint32()
    hist = np.zeros(6 * max_iter)
    _check(load().mq_optimize_grad(eng._h, C.byref(jj), (C.c_double * 4)(*dc), C.byref(ph), C.byref(obj),
                                   C.byref(o), C.byref(J), C.byref(ne), _dptr(hist)))
    theta = [jj.drive.v_rf, jj.drive.f_rf, jj.drive.v_ac, jj.drive.f_ac, jj.dc_scale]
    return {"job": jj, "theta": theta, "J": J.value, "n_eval": ne.value, "history": hist[: 6 * ne.value].reshape(-1, 6)}


def inject_reference(grid: dict, j: Job, ph: Physics) -> dict:
    """Engine::inject (channel.rs:342-374) with the reference's xorshift64 stream, on the host:
    initial states of every ion of one job, plus their xoshiro128** seed words (shape (n, 4))."""
    n = int(j.n)
    o = {k: np.zeros(n, np.float64) for k in ("x", "y", "z", "vx", "vy", "vz", "phi_ac0")}
    o["cd"] = np.zeros(n, np.uint32)
    o["steps"] = np.zeros(n, np.uint32)
    o["ncoll"] = np.zeros(n, np.uint32)
    o["status"] = np.zeros(n, np.uint8)
    rng = np.zeros((n, 4), np.uint32)
    f, keep = make_fields(grid)
    _check(load().mq_inject_reference(C.byref(f), C.byref(j), C.byref(ph), C.byref(soa_of(o)),
                                      rng.ctypes.data_as(C.POINTER(C.c_uint32))))
    del keep
    o["rng"] = rng
    return o


def summarize(status, ncoll, steps, job, njobs: int, dt: float) -> list[Outcome]:
    """summarize (channel.rs:496-506) through the C ABI."""
    status = np.ascontiguousarray(status, np.uint8)
    ncoll = np.ascontiguousarray(ncoll, np.uint32)
    steps = np.ascontiguousarray(steps, np.uint32)
    job = np.ascontiguousarray(job, np.uint32)
    out = (Outcome * njobs)()
    u32 = C.POINTER(C.c_uint32)
    _check(load().mq_summarize(status.ctypes.data_as(C.POINTER(C.c_uint8)), ncoll.ctypes.data_as(u32),
                               steps.ctypes.data_as(u32), job.ctypes.data_as(u32), len(status), njobs, dt, out))
    return list(out)


def soa_of(o: dict) -> IonSoA:
    """mq_ion_soa over the arrays of a dict of final states (optional keys may be absent)."""
    soa = IonSoA()
    for k in ("x", "y", "z", "vx", "vy", "vz", "phi_ac0"):
        if k in o:
            setattr(soa, k, _dptr(o[k]))
    for k in ("steps", "ncoll", "cd"):
        if k in o:
            setattr(soa, k, o[k].ctypes.data_as(C.POINTER(C.c_uint32)))
    soa.status = o["status"].ctypes.data_as(C.POINTER(C.c_uint8))
    return soa


def export_chion(o: dict, jobs, ph: Physics) -> np.ndarray:
    """64-byte ChIon records (channel.rs:292-294) of final states in job order."""
    n = len(o["status"])
    for k in ("x", "y", "z", "vx", "vy", "vz", "phi_ac0"):
        if k in o:
            o[k] = np.ascontiguousarray(o[k], np.float64)
    for k in ("steps", "ncoll", "cd", "job"):
        if k in o:
            o[k] = np.ascontiguousarray(o[k], np.uint32)
    o["status"] = np.ascontiguousarray(o["status"], np.uint8)
    out = np.zeros(n, CHION_DTYPE)
    ja = jobs_array(jobs)
    _check(load().mq_export_chion(C.byref(soa_of(o)), o["job"].ctypes.data_as(C.POINTER(C.c_uint32)), n, ja,
                                  len(jobs), C.byref(ph), out.ctypes.data_as(C.c_void_p)))
    return out


def summarize_chion(recs: np.ndarray, job, njobs: int) -> list[Outcome]:
    """summarize (channel.rs:496-506) over ChIon records."""
    recs = np.ascontiguousarray(recs, CHION_DTYPE)
    job = np.ascontiguousarray(job, np.uint32)
    out = (Outcome * max(njobs, 1))()
    _check(load().mq_summarize_chion(recs.ctypes.data_as(C.c_void_p), job.ctypes.data_as(C.POINTER(C.c_uint32)),
                                     len(recs), njobs, out))
    return list(out)[:njobs]


def fnv1a64(data) -> int:
    """FNV-1a 64 of raw bytes (check.rs:100-104 hashes ChIon bytes this way)."""
    b = memoryview(np.ascontiguousarray(data)).cast("B") if not isinstance(data, (bytes, bytearray)) else data
    buf = (C.c_char * len(b)).from_buffer_copy(bytes(b))
    return int(load().mq_fnv1a64(C.cast(buf, C.c_void_p), len(b)))


# ------------------------------------------------------------------ field build

BASES = ("Rf", "Ac", "DcQ1", "DcPlate", "DcDown", "DcBox")


def geom_classify(g: Geom) -> np.ndarray:
    """MQ_ELEC_* code per node, shape (nz, ny, nx) (Geom::classify on the GPU)."""
    nx, ny, nz, _ = g.dims()
    out = np.zeros(nx * ny * nz, np.uint8)
    _check(load().mq_geom_classify(C.byref(g), out.ctypes.data_as(C.POINTER(C.c_uint8))))
    return out.reshape(nz, ny, nx)


def laplace_solve(mask: np.ndarray, fixed: np.ndarray, outer_bc: float, tol: float, max_iter: int = 0):
    """fieldsolve::solve on the GPU; mask/fixed of shape (nz, ny, nx). Returns (phi, stats dict)."""
    nz, ny, nx = mask.shape
    m = np.ascontiguousarray(mask, np.uint8)
    f = np.ascontiguousarray(fixed, np.float64)
    phi = np.zeros(m.shape, np.float64)
    st = SolveStats()
    rc = load().mq_laplace_solve(nx, ny, nz, m.ctypes.data_as(C.POINTER(C.c_uint8)), _dptr(f), outer_bc, tol,
                                 max_iter, _dptr(phi), C.byref(st))
    stats = {k: getattr(st, k) for k, _ in st._fields_}
    if rc not in (0, -5):
        _check(rc)
    return phi, stats


def efield(phi: np.ndarray, h: float) -> np.ndarray:
    """E = -grad(phi) (channel.rs:168-184) on the GPU; returns shape (nz, ny, nx, 3)."""
    nz, ny, nx = phi.shape
    p = np.ascontiguousarray(phi, np.float64)
    e = np.zeros(p.shape + (3,), np.float64)
    _check(load().mq_efield(_dptr(p), nx, ny, nz, h, _dptr(e)))
    return e


def field_cache_name(g: Geom, basis: int) -> str:
    buf = C.create_string_buffer(64)
    _check(load().mq_field_cache_name(C.byref(g), basis, buf, len(buf)))
    return buf.value.decode()


def channel_grid(g: Geom, tol: float = 1e-7, cache_dir: str | None = None, potentials: bool = False) -> dict:
    """Fields::build (channel.rs:204-217) on the GPU: a grid dict for Engine(), as synth.py makes."""
    nx, ny, nz, lay = g.dims()
    n = nx * ny * nz
    grid = {"rf": np.zeros(3 * n), "ac": np.zeros(3 * n), "dc": [np.zeros(3 * n) for _ in range(4)],
            "mask": np.zeros(n, np.uint8)}
    out = ChannelFields()
    out.rf, out.ac = _dptr(grid["rf"]), _dptr(grid["ac"])
    for b in range(4):
        out.dc[b] = _dptr(grid["dc"][b])
    phis = [np.zeros(n) for _ in range(6)] if potentials else []
    for b, p in enumerate(phis):
        out.phi[b] = _dptr(p)
    out.mask = grid["mask"].ctypes.data_as(C.POINTER(C.c_uint8))
    _check(load().mq_fields_build(C.byref(g), tol, cache_dir.encode() if cache_dir else None, C.byref(out)))
    f = Fields()
    _check(load().mq_geom_fields_params(C.byref(g), C.byref(f)))
    for k in ("nx", "ny", "nz", "h", "axis_x", "axis_y", "z_inject", "z_return", "z_exit", "z_return_legacy",
              "z_exit_legacy", "crash_z0_legacy", "crash_z1_legacy"):
        grid[k] = getattr(f, k)
    grid["stats"] = [{k: getattr(out.stats[b], k) for k, _ in SolveStats._fields_} for b in range(6)]
    grid["layers_mm"] = lay
    if potentials:
        grid["phi"] = phis
    return grid


# ------------------------------------------------------------------ legacy Java potentials


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _u8ptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


def legacy_dims(opts: LegacyOpts | None = None):
    """(nx, ny, nz, check) of the Java sketch's grid."""
    o = opts or LegacyOpts.default()
    nx, ny, nz = C.c_int32(), C.c_int32(), C.c_int32()
    ch = (C.c_int32 * 3)()
    _check(load().mq_legacy_dims(C.byref(o), C.byref(nx), C.byref(ny), C.byref(nz), ch))
    return nx.value, ny.value, nz.value, tuple(ch)


def legacy_rasterize(basis: int, opts: LegacyOpts | None = None):
    """Start state of one basis on the GPU: (phi float32, electrode uint8), shape (nz, ny, nx)."""
    o = opts or LegacyOpts.default()
    nx, ny, nz, _ = legacy_dims(o)
    phi = np.zeros((nz, ny, nx), np.float32)
    el = np.zeros((nz, ny, nx), np.uint8)
    _check(load().mq_legacy_rasterize(C.byref(o), basis, _fptr(phi), _u8ptr(el)))
    return phi, el


def legacy_sweeps(phi: np.ndarray, electrode: np.ndarray, check, precision_pct: float, max_iter: int):
    """The sketch's sweeps on the GPU on any grid (nz, ny, nx); returns (phi, stats dict)."""
    nz, ny, nx = phi.shape
    p = np.ascontiguousarray(phi, np.float32).copy()
    e = np.ascontiguousarray(electrode, np.uint8)
    ch = (C.c_int32 * 3)(*check)
    st = LegacyStats()
    _check(load().mq_legacy_sweeps(nx, ny, nz, _fptr(p), _u8ptr(e), ch, precision_pct, max_iter, C.byref(st)))
    return p, st.as_dict()


def legacy_potential(basis: int, opts: LegacyOpts | None = None):
    """Rasterize + sweeps of one basis: (phi, electrode, stats dict), shape (nz, ny, nx)."""
    o = opts or LegacyOpts.default()
    nx, ny, nz, _ = legacy_dims(o)
    phi = np.zeros((nz, ny, nx), np.float32)
    el = np.zeros((nz, ny, nx), np.uint8)
    st = LegacyStats()
    _check(load().mq_legacy_potential(C.byref(o), basis, _fptr(phi), _u8ptr(el), C.byref(st)))
    return phi, el, st.as_dict()


def legacy_export(directory: str, opts: LegacyOpts | None = None) -> list:
    """Writes eject_{RF,DC,AC}.f32 and eject_{RF,AC}_electrode.u8 to directory; returns the stats per basis."""
    o = opts or LegacyOpts.default()
    st = (LegacyStats * 3)()
    _check(load().mq_legacy_export(C.byref(o), directory.encode(), st))
    return [s.as_dict() for s in st]


def fields_from_theirs(directory: str, potentials: bool = False) -> dict:
    """Fields::from_theirs (channel.rs:187-202): a grid dict for Engine() from the five exported files."""
    g = Geom.default()
    nx, ny, nz, lay = g.dims()
    n = nx * ny * nz
    grid = {"rf": np.zeros(3 * n), "ac": np.zeros(3 * n), "dc": [np.zeros(3 * n) for _ in range(4)],
            "mask": np.zeros(n, np.uint8)}
    out = ChannelFields()
    out.rf, out.ac = _dptr(grid["rf"]), _dptr(grid["ac"])
    for b in range(4):
        out.dc[b] = _dptr(grid["dc"][b])
    phis = [np.zeros(n) for _ in range(6)] if potentials else []
    for b, p in enumerate(phis):
        out.phi[b] = _dptr(p)
    out.mask = _u8ptr(grid["mask"])
    _check(load().mq_fields_from_theirs(directory.encode(), C.byref(out)))
    f = Fields()
    _check(load().mq_geom_fields_params(C.byref(g), C.byref(f)))
    for k in ("nx", "ny", "nz", "h", "axis_x", "axis_y", "z_inject", "z_return", "z_exit", "z_return_legacy",
              "z_exit_legacy", "crash_z0_legacy", "crash_z1_legacy"):
        grid[k] = getattr(f, k)
    grid["layers_mm"] = lay
    if potentials:
        grid["phi"] = phis
    return grid


# ------------------------------------------------------------------ tandem batch

TANDEM_REGION_SPLIT, TANDEM_REGION_COMBINE, TANDEM_NO_PORT = 0, 10, 255


def tandem_geom(seam_z: float, mouths, ap_r: float, entrance, l_ext: float) -> TandemGeom:
    """mq_tandem_geom: the seam, the nine mouth centres (x, y), the aperture radius, the entrance aperture's
    centre (x, z) on the cube's y = (ny - 1) h face and the entrance tube length."""
    g = TandemGeom()
    g.seam_z, g.ap_r, g.l_ext = seam_z, ap_r, l_ext
    g.entrance[0], g.entrance[1] = entrance
    for k, (x, y) in enumerate(mouths):
        g.mouth[k][0], g.mouth[k][1] = x, y
    return g


def tandem_drive(v_rf, f_rf, v_ac, tones, dc_split, dc_combine, dc_port, bias_idle) -> TandemDrive:
    d = TandemDrive()
    d.v_rf, d.f_rf, d.v_ac = v_rf, f_rf, v_ac
    for k in range(9):
        d.tones[k] = tones[k]
    for b in range(4):
        d.dc_split[b], d.dc_combine[b] = dc_split[b], dc_combine[b]
        d.dc_port[b], d.bias_idle[b] = dc_port[b], bias_idle[b]
    return d


class Tandem:
    """The tandem multi-domain batch (include/mq_tandem.h) over the C ABI."""

    def __init__(self, cube: dict, port: dict, geom: TandemGeom):
        fc, self._kc = make_fields(cube)
        fp, self._kp = make_fields(port)
        h = C.c_void_p()
        _check(load().mq_tandem_create(C.byref(fc), C.byref(fp), C.byref(geom), C.byref(h)))
        self._h = h

    def close(self) -> None:
        if self._h:
            load().mq_tandem_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, jobs, drive: TandemDrive, ph: Physics, path: int = PATH_BATCHED):
        """(per-port counts as dicts, totals dict)."""
        ja = jobs_array(jobs)
        pc = (PortCounts * 9)()
        tt = TandemTotals()
        _check(load().mq_tandem_run(self._h, ja, len(jobs), C.byref(drive), C.byref(ph), path, pc, C.byref(tt)))
        return [p.as_dict() for p in pc], tt.as_dict()

    def run_ions(self, jobs, drive: TandemDrive, ph: Physics, path: int = PATH_BATCHED):
        """(counts, totals, final states dict with region / last / job)."""
        ja = jobs_array(jobs)
        n = int(sum(j.n for j in jobs))
        o = {k: np.zeros(n, np.float64) for k in ("x", "y", "z", "vx", "vy", "vz")}
        o["steps"] = np.zeros(n, np.uint32)
        o["ncoll"] = np.zeros(n, np.uint32)
        o["status"] = np.zeros(n, np.uint8)
        o["region"] = np.zeros(n, np.uint8)
        o["last"] = np.zeros(n, np.uint8)
        o["job"] = np.zeros(n, np.uint32)
        soa = soa_of(o)
        pc = (PortCounts * 9)()
        tt = TandemTotals()
        u8 = C.POINTER(C.c_uint8)
        _check(load().mq_tandem_run_ions(self._h, ja, len(jobs), C.byref(drive), C.byref(ph), path, pc, C.byref(tt),
                                         C.byref(soa), o["region"].ctypes.data_as(u8), o["last"].ctypes.data_as(u8),
                                         o["job"].ctypes.data_as(C.POINTER(C.c_uint32))))
        return [p.as_dict() for p in pc], tt.as_dict(), o

    def stats(self) -> dict:
        s = RunStats()
        _check(load().mq_tandem_last_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def set_plain_recs(self, on: bool) -> None:
        """mq_tandem_set_plain_recs: the FP64 path reads per-cell records (True) or the node tables."""
        _check(load().mq_tandem_set_plain_recs(self._h, int(bool(on))))


# ------------------------------------------------------------------ scan API

def format_f64(x: float, display: bool = False) -> str:
    """Rust {:?} (display=False) or {} (display=True) notation of an f64."""
    buf = C.create_string_buffer(64)
    _check(load().mq_format_f64(x, int(display), buf, len(buf)))
    return buf.value.decode()


def scan_key(geom_key: str, j: Job, dc, ph: Physics, path: int = PATH_BATCHED) -> str:
    """Cache key of one job (scan.rs:9-11) for a run on `path`."""
    buf = C.create_string_buffer(4096)
    _check(load().mq_scan_key(geom_key.encode(), C.byref(j), (C.c_double * 4)(*dc), C.byref(ph), int(path), buf,
                              len(buf)))
    return buf.value.decode()


def adaptive_fine_points(f_khz, t, n_lo: int, base: float, f_step: float) -> list[float]:
    f = np.ascontiguousarray(f_khz, np.float64)
    tt = np.ascontiguousarray(t, np.float64)
    n = C.c_size_t()
    out = np.zeros(4 * len(f) * 16 + 8)
    _check(load().mq_adaptive_fine_points(_dptr(f), _dptr(tt), len(f), n_lo, base, f_step, _dptr(out), len(out),
                                          C.byref(n)))
    return list(out[: n.value])


def cached_batch(eng: "Engine", geom_key: str, jobs, dc, ph: Physics, cache_path: str,
                 path: int = PATH_BATCHED) -> tuple[list[Outcome], int]:
    """scan::cached_batch (scan.rs:31-47); returns (outcomes, jobs computed)."""
    ja = jobs_array(jobs)
    out = (Outcome * len(jobs))()
    nc = C.c_size_t()
    _check(load().mq_cached_batch(eng._h, geom_key.encode(), ja, len(jobs), (C.c_double * 4)(*dc), C.byref(ph), path,
                                  cache_path.encode(), out, C.byref(nc)))
    return list(out), nc.value


def adaptive_scan(eng: "Engine", geom_key: str, sp: Species, rf: RfAc, dc, ph: Physics, f_lo: float, f_hi: float,
                  f_step: float, n_lo: int, n_hi: int, base: float, seed0: int, cache_path: str,
                  path: int = PATH_BATCHED) -> list[tuple]:
    """scan::adaptive_scan (scan.rs:52-78): sorted (f_kHz, T, C, n) points."""
    cap = 8 * int((f_hi - f_lo) / f_step + 2) + 16
    out = (ScanPoint * cap)()
    n = C.c_size_t()
    _check(load().mq_adaptive_scan(eng._h, geom_key.encode(), C.byref(sp), C.byref(rf), (C.c_double * 4)(*dc),
                                   C.byref(ph), f_lo, f_hi, f_step, n_lo, n_hi, base, seed0, path, cache_path.encode(),
                                   out, cap, C.byref(n)))
    return [(p.f_khz, p.transmitted, p.crashed, p.n) for p in out[: n.value]]
