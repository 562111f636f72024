"""GPU: the hand-written adjoint of the differentiable form (mq_forward / mq_backward)
against the oracle's FP64 dual-number forward-mode gradient and against central
finite differences of the GPU forward (h = 1e-6 relative)."""
import numpy as np
import pytest

import mqhip
import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1():
    g = synth.c1_grid()
    e = mqhip.Engine(g)
    yield g, e
    e.close()


def setting(n=256, nsteps=2000, v_dc=5.0):
    ph = mqhip.Physics.correct(0.0).copy(dt=1e-7, injection=mqhip.INJ_GAUSSIAN, inj_sigma_mm=0.1, legacy_force=1)
    job = mqhip.job(499.0, 1.0, 0.0, 100.0, 500e3, 0.5, 100e3, n, 1499, dc_scale=v_dc / 5.0)
    obj = mqhip.Objective(nsteps, 0, 2.0e-3, 2e-5, 1.0e-3, 0.3)
    return job, ph, obj


def test_adjoint_matches_dual_numbers(c1, oracle_mod):
    g, e = c1
    job, ph, obj = setting()
    print("[gpu] adjoint vs dual numbers", flush=True)
    J = e.forward(job, [0, 5, 0, 0], ph, obj)
    grad, gstate = e.backward(1.0, state=True)
    Jr, gr, gsr = oracle_mod.grad_reference(g, job, [0, 5, 0, 0], ph, obj, 8)
    print(f"[gpu]   J {J:.15g} vs {Jr:.15g}; grad {grad} vs {gr}", flush=True)
    assert J == pytest.approx(Jr, rel=1e-12)
    scale = np.abs(gr).max()
    assert np.allclose(grad, gr, rtol=1e-9, atol=1e-12 * scale)
    sscale = np.abs(gsr).max()
    assert np.allclose(gstate, gsr, rtol=1e-8, atol=1e-10 * sscale)


def test_float_tables_match_dual_numbers(oracle_mod):
    """A grid whose basis values are all exactly floats runs the differentiable form on float
    copies of the node tables (half the bytes per gather); the arithmetic is the same FP64, so
    the bars of the double-table test above hold unchanged."""
    g = synth.c1_grid()
    g["rf"] = g["rf"].astype(np.float32).astype(np.float64)
    g["ac"] = g["ac"].astype(np.float32).astype(np.float64)
    g["dc"] = [d.astype(np.float32).astype(np.float64) if d is not None else None for d in g["dc"]]
    e = mqhip.Engine(g)
    try:
        job, ph, obj = setting()
        job.dc_scale *= 5.0   # the gate voltage on the job; the combined DC (1 x basis) stays exactly float
        print("[gpu] adjoint on float node tables vs dual numbers", flush=True)
        J = e.forward(job, [0, 1, 0, 0], ph, obj)
        grad, gstate = e.backward(1.0, state=True)
    finally:
        e.close()
    Jr, gr, gsr = oracle_mod.grad_reference(g, job, [0, 1, 0, 0], ph, obj, 8)
    assert J == pytest.approx(Jr, rel=1e-12)
    assert np.allclose(grad, gr, rtol=1e-9, atol=1e-12 * np.abs(gr).max())
    assert np.allclose(gstate, gsr, rtol=1e-8, atol=1e-10 * np.abs(gsr).max())


@pytest.fixture(scope="module")
def c1_float():
    """The C1 grid rounded to floats: the differentiable form then runs on per-cell float records."""
    g = synth.c1_grid()
    g["rf"] = g["rf"].astype(np.float32).astype(np.float64)
    g["ac"] = g["ac"].astype(np.float32).astype(np.float64)
    g["dc"] = [d.astype(np.float32).astype(np.float64) if d is not None else None for d in g["dc"]]
    e = mqhip.Engine(g)
    yield g, e
    e.close()


@pytest.mark.parametrize("n,K", [(300, 0), (300, 7), (64, 1), (1, 0), (777, 700)])
def test_float_tables_ragged_blocks_match_dual_numbers(c1_float, oracle_mod, n, K):
    """The float-record kernels with ion counts that leave partial wavefronts and blocks,
    and with checkpoint intervals that give many, odd, even and single segments (700 steps:
    K = 7 gives 100 segments, K = 0 picks 27 -> 26, K = 1 gives 700, K = 700 gives one)."""
    g, e = c1_float
    job, ph, obj = setting(n=n, nsteps=700)
    job.dc_scale *= 5.0
    obj = mqhip.Objective(700, K, obj.z_exit, obj.scale, obj.r_leak, obj.w_leak)
    print(f"[gpu] float-record adjoint, n={n}, ckpt={K}", flush=True)
    J = e.forward(job, [0, 1, 0, 0], ph, obj)
    grad, gstate = e.backward(1.0, state=True)
    Jr, gr, gsr = oracle_mod.grad_reference(g, job, [0, 1, 0, 0], ph, obj, 8)
    assert J == pytest.approx(Jr, rel=1e-12)
    assert np.allclose(grad, gr, rtol=1e-9, atol=1e-12 * np.abs(gr).max())
    assert np.allclose(gstate, gsr, rtol=1e-8, atol=1e-10 * np.abs(gsr).max())


def test_adjoint_matches_central_differences(c1):
    """Central differences with h = 1e-6 relative and their Richardson extrapolation (h, h/2).
    Operating point with a 0.5 V gate: no ion reaches the barrier top, so the objective is
    smooth on the scale of h (at 5 V the ions reflected near the barrier top make the
    pathwise derivative of a 0.02 mm sigmoid ill-conditioned; the dual-number test above
    covers that case exactly)."""
    g, e = c1
    job, ph, obj = setting(n=512, nsteps=1500, v_dc=0.5)
    print("[gpu] adjoint vs central differences", flush=True)
    e.forward(job, [0, 5, 0, 0], ph, obj)
    grad = e.backward(1.0)
    names = [("drive", "v_rf"), ("drive", "f_rf"), ("drive", "v_ac"), ("drive", "f_ac"), (None, "dc_scale")]

    def central(grp, name, h):
        base = getattr(job.drive if grp else job, name)
        vals = []
        for sgn in (+1, -1):
            j2 = mqhip.Job()
            j2.sp, j2.drive, j2.n, j2.seed, j2.dc_scale = job.sp, job.drive, job.n, job.seed, job.dc_scale
            setattr(j2.drive if grp else j2, name, base + sgn * h)
            vals.append(e.forward(j2, [0, 5, 0, 0], ph, obj))
        return (vals[0] - vals[1]) / (2 * h)

    for k, (grp, name) in enumerate(names):
        h = 1e-6 * abs(getattr(job.drive if grp else job, name))
        d1, d2 = central(grp, name, h), central(grp, name, h / 2)
        rich = (4 * d2 - d1) / 3
        print(f"[gpu]   {name}: adjoint {grad[k]:.10g} central {d1:.10g} richardson {rich:.10g}", flush=True)
        assert rich == pytest.approx(grad[k], rel=1e-4, abs=1e-9 * np.abs(grad).max())
        assert d1 == pytest.approx(grad[k], rel=1e-3, abs=1e-9 * np.abs(grad).max())


def test_per_ion_outputs_sum_to_totals(c1):
    g, e = c1
    job, ph, obj = setting(n=300, nsteps=500)
    J = e.forward(job, [0, 5, 0, 0], ph, obj)
    Ji = e.forward_per_ion()
    grad = e.backward(1.0)
    gi = e.backward_per_ion(300)
    assert Ji.mean() == pytest.approx(J, rel=1e-13)
    assert np.allclose(gi.sum(axis=1), grad, rtol=1e-12, atol=1e-15 * np.abs(grad).max())


def test_checkpoint_interval_does_not_change_gradient(c1):
    g, e = c1
    job, ph, obj = setting(n=128, nsteps=700)
    out = []
    for K in (1, 7, 0, 700):
        o = mqhip.Objective(obj.nsteps, K, obj.z_exit, obj.scale, obj.r_leak, obj.w_leak)
        J = e.forward(job, [0, 5, 0, 0], ph, o)
        out.append((J, e.backward(2.0)))
    for J, gr in out[1:]:
        assert J == out[0][0] and np.array_equal(gr, out[0][1])


def test_adjoint_errors(c1):
    g, e = c1
    job, ph, obj = setting(n=8, nsteps=10)
    e.run_batch([job], [0, 5, 0, 0], ph)
    with pytest.raises(mqhip.MqError, match="preceding mq_forward"):
        e.backward()
    with pytest.raises(mqhip.MqError, match="thermal collision model only"):
        e.forward(job, [0, 5, 0, 0], ph.copy(pressure_pa=0.1, legacy_collisions=1), obj)


@pytest.mark.parametrize("tables", ["double", "float"])
def test_frozen_collisions_match_dual_numbers(c1, c1_float, oracle_mod, tables):
    """Thermal N2 at 1 mTorr (a few collisions per ion over 2,000 steps): the adjoint with the Philox draws frozen
    against the oracle's dual numbers through the same collisions (tests/test_oracle_grad.py pins those against
    central differences), on the double node tables and on the float cell records; without gas the same job gives
    another J and gradient."""
    g, e = c1 if tables == "double" else c1_float
    _, ph, obj = setting()
    # the float records need an exactly-float combined DC: the gate voltage goes on the job there
    dc, s_dc = ([0, 5, 0, 0], 1.0) if tables == "double" else ([0, 1, 0, 0], 5.0)
    job = mqhip.job(499.0, 1.0, 230.0, 100.0, 500e3, 0.5, 100e3, 256, 1499, dc_scale=s_dc)
    gas = ph.copy(pressure_pa=0.133322)
    print(f"[gpu] adjoint with frozen collisions vs dual numbers ({tables} tables)", flush=True)
    J = e.forward(job, dc, gas, obj)
    grad, gstate = e.backward(1.0, state=True)
    Jr, gr, gsr = oracle_mod.grad_reference(g, job, dc, gas, obj, 8)
    J0 = e.forward(job, dc, ph, obj)
    g0 = e.backward(1.0)
    print(f"[gpu]   J {J:.15g} vs {Jr:.15g} (no gas {J0:.6g}); grad {grad} vs {gr}", flush=True)
    assert J == pytest.approx(Jr, rel=1e-12)
    assert np.allclose(grad, gr, rtol=1e-9, atol=1e-12 * np.abs(gr).max())
    assert np.allclose(gstate, gsr, rtol=1e-8, atol=1e-10 * np.abs(gsr).max())
    assert J != J0 and not np.allclose(grad, g0, rtol=1e-3)


def test_adjoint_per_ion_at_full_size_matches_dual_numbers(oracle_mod):
    """The differentiable form at its full size (65,536 collision-free ions, 20,000 steps,
    128x128x512 grid, FP32-rounded bases): per-ion adjoint contributions of the 16 ions with
    the largest |dJ_i/dV_rf| (orbits whose sensitivity grows to ~1e5 per volt, where central
    differences at 1e-6 relative are far outside the linear range) and of 48 random ions
    against the oracle's FP64 dual-number derivative of the same ions."""
    grid = synth.c2_grid()
    for k in ("rf", "ac"):
        grid[k] = grid[k].astype(np.float32).astype(np.float64)
    grid["dc"] = [d.astype(np.float32).astype(np.float64) if d is not None else None for d in grid["dc"]]
    ph = mqhip.Physics.correct(0.0).copy(dt=1e-7, injection=mqhip.INJ_GAUSSIAN, inj_sigma_mm=0.1, legacy_force=1)
    n = 65536
    job = mqhip.job(499.0, 1.0, 0.0, 100.0, 500e3, 0.5, 100e3, n, 4499, dc_scale=0.5)
    obj = mqhip.Objective(20000, 0, 18.6e-3, 2e-5, 1.0e-3, 0.3)
    dc = [0.0, 1.0, 0.0, 0.0]
    e = mqhip.Engine(grid)
    try:
        print("[gpu] full-size adjoint: forward + backward of 65536 ions x 20000 steps", flush=True)
        e.forward(job, dc, ph, obj)
        Ji = e.forward_per_ion()
        e.backward(1.0)
        gi = e.backward_per_ion(n) * n
    finally:
        e.close()
    rng = np.random.default_rng(7)
    top = np.argsort(np.abs(gi[0]))[::-1][:16]
    pick = np.unique(np.concatenate([top, rng.choice(n, 48, replace=False)])).astype(np.uint32)
    Jr, gr = oracle_mod.grad_ions(grid, job, dc, ph, obj, pick, 8)
    print(f"[gpu]   largest |dJ_i/dV_rf| {np.abs(gi[0][top[0]]):.6g} (ion {top[0]})", flush=True)
    assert np.abs(gi[0][top[0]]) > 1e3   # the chaotic ions are in the check
    assert np.allclose(Ji[pick], Jr, rtol=1e-11, atol=1e-14)
    for k in range(5):
        sc = np.abs(gr[k]).max()
        assert np.allclose(gi[k][pick], gr[k], rtol=1e-8, atol=1e-12 * sc), k
