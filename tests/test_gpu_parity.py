"""GPU parity of the HIP path against the CPU oracle, through the C ABI.

FP64 operator applies are compared bit for bit: the kernels evaluate every
reference sum in the reference's order with contraction disabled.  Results
that involve transcendental functions (manufactured sources, exact boundary
values) are compared with the tolerances stated in each test.
"""
import numpy as np
import pytest

import helpers as H
import oracle_ffi as O

pytestmark = pytest.mark.gpu

WAVE_CFGS = {
    "wave48_E4_1": H.sphere_wave_cfg(48),
    "wave22_E4u_short": H.sphere_wave_cfg(22, scheme=O.E4U_1, alpha=(0.1, 0.2)),
    "wave30_radial_E4u": H.sphere_wave_cfg(30, linear=False, scheme=O.E4U_1),
}
HEAT_CFGS = {
    "heat2d_21": H.heat2d_cfg(21),
    "heat3d_E4_2": H.heat3d_cfg(),
    "heat3d_E2_2": H.heat3d_cfg(scheme=O.E2_2),
    # odd z extent: one point per lane in the x-marching tiles
    "heat3d_odd": H.heat3d_cfg(n=(21, 23, 27)),
}
ALL = {**WAVE_CFGS, **HEAT_CFGS}
# large enough for the derivative's several-planes-per-lane launch along x and
# y (at least 2048 workgroups of four planes); extents that are not multiples
# of the tile or of the four planes
MULTI_PLANE = {"heat3d_multi_plane": H.heat3d_cfg(n=(130, 66, 258), center=(0.5, 0.45, 0.55), radius=0.3)}
# z extent a multiple of 4 (four points per lane in the vector z kernel), with
# floating and with Dirichlet grid faces: every grid-wall closure end is
# evaluated by its own lanes (four per end along z), the closures at the
# sphere (B terms) come from the slow-row list
for _faces, _bc in (("floating", [O.FLOATING] * 6), ("dirichlet", [O.DIRICHLET, O.FLOATING] * 3)):
    _cfg = H.heat3d_cfg(n=(130, 66, 256), center=(0.5, 0.45, 0.55), radius=0.3)
    _cfg["grid_bc"] = _bc
    MULTI_PLANE[f"heat3d_multi_plane_z4_{_faces}"] = _cfg
# odd z extent: the z rows take the point-wise kernel with four planes per
# lane, and its grid-wall ends the four-lane form
_cfg = H.heat3d_cfg(n=(130, 66, 257), center=(0.5, 0.45, 0.55), radius=0.3)
_cfg["grid_bc"] = [O.FLOATING] * 6
MULTI_PLANE["heat3d_multi_plane_zodd_floating"] = _cfg


def _rand_field(osys, seed):
    rng = np.random.default_rng(seed)
    f = osys.field()
    f.D[:] = rng.standard_normal(f.D.size)
    for d in range(3):
        f.R[d][:] = rng.standard_normal(f.R[d].size)
    return f


def _pair(ctx, cfg):
    osys = H.oracle_system(cfg)
    prob = H.product_problem(ctx, cfg)
    return osys, prob


@pytest.mark.parametrize("name", list(ALL) + list(MULTI_PLANE))
@pytest.mark.parametrize("plus", [False, True])
def test_derivative_apply_bitexact(ctx, name, plus):
    cfg = {**ALL, **MULTI_PLANE}[name]
    osys, prob = _pair(ctx, cfg)
    u = _rand_field(osys, 1)
    for d in range(3):
        if prob.n[d] < 2:
            continue
        du = _rand_field(osys, 2)  # eq must overwrite rows only; plus_eq accumulates
        gu, gdu = prob.field(), prob.field()
        gu.upload(u.flat())
        gdu.upload(du.flat())
        osys.deriv_apply(d, u, du, plus_eq=plus)
        prob.derivative(d)(gu, gdu, 1 if plus else 0)
        got = gdu.download()
        want = du.flat()
        assert np.array_equal(got, want), f"{name} dir {d}: max diff {np.max(np.abs(got - want))}"


@pytest.mark.parametrize("name", list(WAVE_CFGS))
def test_gradient_bitexact(ctx, name):
    osys, prob = _pair(ctx, WAVE_CFGS[name])
    u = _rand_field(osys, 3)
    want = osys.gradient(u)
    gu = prob.field()
    gu.upload(u.flat())
    outs = [prob.field() for _ in range(3)]
    prob.gradient(gu)(*outs)
    for c in range(3):
        assert np.array_equal(outs[c].download(), want[c].flat())


@pytest.mark.parametrize("name", list(HEAT_CFGS))
def test_laplacian_bitexact(ctx, name):
    osys, prob = _pair(ctx, HEAT_CFGS[name])
    u = _rand_field(osys, 4)
    want = osys.laplacian(u)
    gu, gl = prob.field(), prob.field()
    gu.upload(u.flat())
    prob.laplacian(gu)(gl)
    assert np.array_equal(gl.download(), want.flat())


@pytest.mark.parametrize("name", list(WAVE_CFGS))
@pytest.mark.parametrize("fused", [True, False])
def test_wave_rhs_bitexact(ctx, name, fused):
    """Scalar-wave RHS has no transcendental in the linear form; the radial
    form computes gG with sqrt and division, both correctly rounded."""
    osys, prob = _pair(ctx, WAVE_CFGS[name])
    u = _rand_field(osys, 5)
    want = osys.rhs(u, 0.3)
    gu, gr = prob.field(), prob.field()
    gu.upload(u.flat())
    prob.rhs(gu, gr, 0.3, fused=fused)
    assert np.array_equal(gr.download(), want.flat())


@pytest.mark.parametrize("name", list(HEAT_CFGS))
@pytest.mark.parametrize("fused", [True, False])
def test_heat_rhs(ctx, name, fused):
    """Heat RHS adds the manufactured source (sin, exp on the device vs glibc
    on the host): tolerance 1e-13 relative to the max-norm of the result."""
    osys, prob = _pair(ctx, HEAT_CFGS[name])
    u = osys.field()
    osys.initialize(u, 0.0)
    want = osys.rhs(u, 0.25).flat()
    gu, gr = prob.field(), prob.field()
    gu.upload(u.flat())
    prob.rhs(gu, gr, 0.25, fused=fused)
    got = gr.download()
    scale = np.max(np.abs(want))
    assert np.max(np.abs(got - want)) <= 1e-13 * scale


@pytest.mark.parametrize("name", ["wave48_E4_1", "heat2d_21", "heat3d_E4_2"])
@pytest.mark.parametrize("fused", [False, True])
def test_rk4_steps(ctx, name, fused):
    """Several RK4 steps from the exact initial condition against the
    oracle's step; tolerance 1e-11 relative max-norm (full solves)."""
    cfg = ALL[name]
    osys, prob = _pair(ctx, cfg)
    u0 = osys.field()
    osys.initialize(u0, 0.0)
    osys.update_boundary(u0, 0.0)
    dt = osys.dt(0.5, 0.8)
    assert abs(dt - prob.dt(0.5, 0.8)) == 0.0
    nsteps = 3
    gu = prob.field()
    gu.upload(u0.flat())
    out, rk, srhs = osys.field(), osys.field(), osys.field()
    t = 0.0
    for _ in range(nsteps):
        osys.rk4_step(u0, out, rk, srhs, t, dt)
        u0.D[:] = out.D
        for d in range(3):
            u0.R[d][:] = out.R[d]
        t += dt
    if fused:
        prob.rk4_fused(gu, 0.0, dt, nsteps)
    else:
        g_out, g_rk, g_srhs = prob.field(), prob.field(), prob.field()
        tt = 0.0
        for _ in range(nsteps):
            prob.rk4(gu, g_out, g_rk, g_srhs, tt, dt)
            gu.upload(g_out.download())
            tt += dt
    got = gu.download()
    want = u0.flat()
    err = np.max(np.abs(got - want)) / np.max(np.abs(want))
    assert err <= 1e-11, f"{name}: relative max-norm error {err:.3e}"


@pytest.mark.parametrize("name", ["wave48_E4_1", "wave30_radial_E4u", "wave22_E4u_short",
                                  "heat3d_E4_2", "heat3d_E2_2", "heat2d_21", "heat3d_odd"])
@pytest.mark.parametrize("variant,ty,xchunk", [(1, 4, 16), (2, 8, 7), (3, 8, 64), (4, 4, 1000)])
def test_march_kernel_bitexact(ctx, name, variant, ty, xchunk):
    """The x-marching form (interior and solid points in registers and LDS,
    special points from a list) gives the same doubles as the point-wise
    tiles (same operation order), over several RK4 steps, D and cut points."""
    cfg = ALL[name]
    osys, prob = _pair(ctx, cfg)
    u0 = osys.field()
    osys.initialize(u0, 0.0)
    osys.update_boundary(u0, 0.0)
    dt = osys.dt(0.5, 1.0)
    a, b = prob.field(), prob.field()
    a.upload(u0.flat())
    b.upload(u0.flat())
    prob.set_tuning(0)
    prob.rk4_fused(a, 0.0, dt, 3)
    prob.set_tuning(variant, ty, xchunk)
    prob.rk4_fused(b, 0.0, dt, 3)
    ga, gb = a.download(), b.download()
    assert np.array_equal(ga, gb), f"max diff {np.max(np.abs(ga - gb))}"


@pytest.mark.parametrize("name", ["wave48_E4_1", "heat3d_E4_2"])
def test_march_point_classes(ctx, name):
    """The x-marching launch updates every point of D; the special records
    hold the faces and the remaining closure rows."""
    osys, prob = _pair(ctx, ALL[name])
    prob.set_tuning(1)
    main, special = prob.stage_points()
    assert main == prob.N
    assert 0 < special < prob.N
    prob.set_tuning(0)
    assert prob.stage_points() == (prob.N, 0)


def test_autotune_keeps_state(ctx):
    cfg = WAVE_CFGS["wave48_E4_1"]
    osys, prob = _pair(ctx, cfg)
    u0 = osys.field()
    osys.initialize(u0, 0.0)
    g = prob.field()
    g.upload(u0.flat())
    v, ty, xc, ms = prob.autotune(g, 0.0, osys.dt(0.5, 1.0))
    assert 0 <= v <= 4 and ms > 0
    assert np.array_equal(g.download(), u0.flat())


def test_stats_ignore_nan_like_the_reference(ctx):
    """compute_scalar_stats keeps the first strict maximum (MaxLoc with `e >
    val`, scalar_system_utils.hpp:102-113): a NaN point never wins.  An all-NaN
    field keeps the MaxLoc identity for D (lowest double, index INT_MAX), and
    err = max(err_d, R errors = 0) reads 0 (:169)."""
    cfg = HEAT_CFGS["heat3d_E4_2"]
    osys, prob = _pair(ctx, cfg)
    u = osys.field()
    osys.initialize(u, 0.0)
    fl = osys.mesh.fluid()
    u.D[fl[40]] += 0.125
    u.D[fl[123]] = np.nan
    gu = prob.field()
    gu.upload(u.flat())
    got, want = prob.stats(gu, 0.0), osys.stats(u, 0.0)
    assert got[0] == want[0] and got[4] == want[4] == fl[40]
    u.D[fl] = np.nan
    gu.upload(u.flat())
    got, want = prob.stats(gu, 0.0), osys.stats(u, 0.0)
    assert got[0] == want[0] == 0.0
    assert got[3] == want[3] == -np.finfo(np.float64).max
    assert got[4] == want[4] == 2147483647


def test_stats_match_oracle(ctx):
    cfg = HEAT_CFGS["heat3d_E4_2"]
    osys, prob = _pair(ctx, cfg)
    u = osys.field()
    osys.initialize(u, 0.0)
    u.D[:] += 1e-3 * np.random.default_rng(6).standard_normal(u.D.size)
    gu = prob.field()
    gu.upload(u.flat())
    want = osys.stats(u, 0.1)
    got = prob.stats(gu, 0.1)
    assert got[4] == want[4]  # location of the max error
    np.testing.assert_allclose(got[:4], want[:4], rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("n", [1, 2, 1001, 1 << 16])
def test_copy(ctx, n):
    """shx_copy moves every double, including the odd tail."""
    import ctypes as C
    from shoccs_amd._lib import lib

    L = lib()
    a, b = C.c_void_p(), C.c_void_p()
    assert L.shx_alloc(ctx.p, 8 * n, C.byref(a)) == 0
    assert L.shx_alloc(ctx.p, 8 * n, C.byref(b)) == 0
    src = np.random.default_rng(n).standard_normal(n)
    assert L.shx_h2d(ctx.p, a, src.ctypes.data_as(C.c_void_p), 8 * n) == 0
    assert L.shx_fill(ctx.p, b, -1.0, n) == 0
    assert L.shx_copy(ctx.p, b, a, n) == 0
    out = np.empty(n)
    assert L.shx_d2h(ctx.p, out.ctypes.data_as(C.c_void_p), b, 8 * n) == 0
    L.shx_free(ctx.p, a)
    L.shx_free(ctx.p, b)
    assert np.array_equal(out, src)


def test_context_close_releases_problems():
    """Closing a context first releases the problems and fields made on it,
    so that objects outliving the context (a failing test's traceback, for
    one) never reach a finalized context."""
    import shoccs_amd as S

    c = S.Context(0)
    prob = H.product_problem(c, HEAT_CFGS["heat3d_E4_2"])
    f = prob.field()
    c.close()
    assert prob.p is None and f.c is None
    del f, prob


# ---------------------------------------------------------------- mixed form
# No reference counterpart: the mixed form (FP32 stage register, compensated
# FP32 interior sums, FP64 u0 / accumulator / closures / B terms) is checked
# against the FP64 form and the exact solution.
MIXED_CFGS = {
    "heat3d_E4_2": HEAT_CFGS["heat3d_E4_2"],
    "wave_box_E4u": dict(H.sphere_wave_cfg(40, scheme=O.E4U_1, alpha=(-0.7733323791884821, 0.1623961700641681)),
                         spheres=[], obj_bc=[]),
    "wave30_radial_E4u": WAVE_CFGS["wave30_radial_E4u"],
}


def _run(prob, u0flat, dt, nsteps, precision, variant=3, ty=4, xc=16):
    import shoccs_amd as S

    prob.set_tuning(variant, ty, xc)
    prob.set_precision(precision)
    g = prob.field()
    g.upload(u0flat)
    prob.rk4_fused(g, 0.0, dt, nsteps)
    out = g.download()
    prob.set_precision(S.F64)
    return out


@pytest.mark.parametrize("name", list(MIXED_CFGS))
def test_mixed_form_close_to_fp64(ctx, name):
    """After 5 steps the mixed form differs from the FP64 form by at most
    1e-5 of the solution's max-norm, and its error against the exact
    solution is within 1.5x of the FP64 form's (plus 1e-7 of the max-norm)."""
    import shoccs_amd as S

    cfg = MIXED_CFGS[name]
    osys, prob = _pair(ctx, cfg)
    u0 = osys.field()
    osys.initialize(u0, 0.0)
    osys.update_boundary(u0, 0.0)
    dt = osys.dt(0.5, 0.8)
    n = 5
    f64 = _run(prob, u0.flat(), dt, n, S.F64)
    mix = _run(prob, u0.flat(), dt, n, S.F32_COMP)
    scale = np.max(np.abs(f64))
    assert np.max(np.abs(mix - f64)) <= 1e-5 * scale
    exact = osys.exact(n * dt).flat()
    fl = osys.mesh.fluid()
    e64 = np.max(np.abs(f64[fl] - exact[fl]))
    emx = np.max(np.abs(mix[fl] - exact[fl]))
    assert emx <= 1.5 * e64 + 1e-7 * scale, (emx, e64)


@pytest.mark.parametrize("name", ["heat3d_E4_2", "wave_box_E4u"])
def test_mixed_form_same_doubles_for_every_tiling(ctx, name):
    """Tile shape and placement of the special work do not change the mixed
    form's results (the same arithmetic per point).  Bounded solutions only:
    a solution that outgrows FP32's range (E4_1's singular grid-wall
    closures) overflows the FP32 stage register."""
    import shoccs_amd as S

    osys, prob = _pair(ctx, MIXED_CFGS[name])
    u0 = osys.field()
    osys.initialize(u0, 0.0)
    osys.update_boundary(u0, 0.0)
    dt = osys.dt(0.5, 0.8)
    ref = _run(prob, u0.flat(), dt, 3, S.F32_COMP, variant=3, ty=4, xc=16)
    assert np.all(np.isfinite(ref))
    for v, ty, xc in [(1, 8, 7), (2, 8, 64), (4, 4, 1000)]:
        assert np.array_equal(_run(prob, u0.flat(), dt, 3, S.F32_COMP, v, ty, xc), ref)


def test_mixed_form_needs_the_marching_tiles(ctx):
    import shoccs_amd as S

    osys, prob = _pair(ctx, HEAT_CFGS["heat3d_E4_2"])
    prob.set_tuning(0)
    prob.set_precision(S.F32_COMP)
    assert prob.tuning()[0] == 3 and prob.precision() == S.F32_COMP
    with pytest.raises(S.ShoccsError):
        prob.set_tuning(0)
    prob.set_precision(S.F64)
    prob.set_tuning(0)


# ------------------------------------------------- euler and the driver
@pytest.mark.parametrize("name", ["wave48_E4_1", "heat3d_E4_2", "heat2d_21"])
def test_euler_steps(ctx, name):
    """integrators::euler on the device against the oracle's step (1e-11
    relative max-norm over 3 steps, as the RK4 steps)."""
    cfg = ALL[name]
    osys, prob = _pair(ctx, cfg)
    u0 = osys.field()
    osys.initialize(u0, 0.0)
    osys.update_boundary(u0, 0.0)
    dt = 0.25 * osys.dt(0.5, 0.8)
    g = prob.field()
    g.upload(u0.flat())
    go, gs = prob.field(), prob.field()
    out, srhs = osys.field(), osys.field()
    t = 0.0
    for _ in range(3):
        osys.euler_step(u0, out, srhs, t, dt)
        u0.D[:] = out.D
        for d in range(3):
            u0.R[d][:] = out.R[d]
        prob.euler(g, go, gs, t, dt)
        g.upload(go.download())
        t += dt
    got, want = g.download(), u0.flat()
    assert np.max(np.abs(got - want)) <= 1e-11 * np.max(np.abs(want))


def _oracle_cycle(osys, integrator, max_step, max_time, h_cfl, p_cfl, max_error):
    """simulation_cycle::run restated over the oracle (stats on u1, the
    controller bounds, valid())."""
    import shoccs_amd as S

    u0 = osys.field()
    osys.initialize(u0, 0.0)
    u1 = osys.field()
    u1.D[:] = u0.D
    for d in range(3):
        u1.R[d][:] = u0.R[d]
    osys.update_boundary(u0, 0.0)
    st = osys.stats(u1, 0.0)
    rows = [st]
    t, step = 0.0, 0
    rk, srhs = osys.field(), osys.field()
    valid = lambda v: np.isfinite(v) and abs(v) <= max_error
    while t < max_time and step < max_step and valid(st[0]):
        dt = osys.dt(h_cfl, p_cfl)
        if integrator == S.EULER:
            osys.euler_step(u0, u1, srhs, t, dt)
        else:
            osys.rk4_step(u0, u1, rk, srhs, t, dt)
        t += dt
        step += 1
        st = osys.stats(u1, t)
        rows.append(st)
        u0.D[:] = u1.D
        for d in range(3):
            u0.R[d][:] = u1.R[d]
    completed = not (t < max_time and step < max_step)
    return dict(time=t, step=step, completed=completed, u=u0.flat(), stats=np.array(rows))


@pytest.mark.parametrize("integrator", ["RK4", "RK4_FUSED", "EULER"])
def test_driver_matches_the_cycle(ctx, integrator):
    """shoccs_run follows simulation_cycle::run: same number of steps, end
    time (max_time is not a multiple of dt: the last step overshoots), final
    state and per-step statistics as the loop over the oracle."""
    import shoccs_amd as S

    cfg = HEAT_CFGS["heat2d_21"]
    osys, prob = _pair(ctx, cfg)
    integ = getattr(S, integrator)
    dt = osys.dt(1.0, 0.8)
    max_time = 6.5 * dt
    want = _oracle_cycle(osys, S.RK4 if integ == S.RK4_FUSED else integ, 10 ** 6, max_time, 1.0,
                         0.8, 1e6)
    g = prob.field()
    got = prob.run(g, integ, max_time=max_time, h_cfl=1.0, p_cfl=0.8)
    assert got["step"] == want["step"] == 7 and got["time"] == want["time"]
    assert got["completed"] and want["completed"]
    scale = np.max(np.abs(want["u"]))
    assert np.max(np.abs(g.download() - want["u"])) <= 1e-11 * scale
    assert got["stats"].shape == want["stats"].shape
    np.testing.assert_allclose(got["stats"][:, :4], want["stats"][:, :4], rtol=1e-9, atol=1e-15)
    assert got["linf"] == got["stats"][-1, 0]


def test_driver_stops_on_invalid_stats(ctx):
    """valid() false ends the run prematurely: the L-inf error is reported
    as NaN, and the step that broke the bound is the last one."""
    import shoccs_amd as S

    cfg = dict(WAVE_CFGS["wave30_radial_E4u"])
    cfg["params"] = list(cfg["params"][:4]) + [1e-3]  # max_error
    osys, prob = _pair(ctx, cfg)
    g = prob.field()
    got = prob.run(g, S.RK4_FUSED, max_step=50, h_cfl=0.5)
    assert not got["completed"] and np.isnan(got["linf"])
    assert 0 < got["step"] < 50
    assert got["stats"][-1, 0] > 1e-3 and np.all(got["stats"][:-1, 0] <= 1e-3)


def test_driver_zero_steps_without_bounds(ctx):
    """step_controller with neither max_step nor max_time runs no step."""
    import shoccs_amd as S

    osys, prob = _pair(ctx, HEAT_CFGS["heat2d_21"])
    g = prob.field()
    got = prob.run(g, S.RK4)
    assert got["step"] == 0 and got["completed"] and got["stats"].shape[0] == 1


# ---------------------------------------------------------------- batches
BATCH_CFGS = [H.heat3d_cfg(n=(20, 24, 28)),
              H.heat3d_cfg(n=(26, 22, 30), center=(0.45, 0.5, 0.52), radius=0.25),
              H.heat3d_cfg(n=(18, 34, 24), center=(0.55, 0.52, 0.47), radius=0.18)]
# one odd z extent: the whole batch takes one point per lane, the problems
# stepped alone two (where even)
BATCH_CFGS_ODD = BATCH_CFGS[:2] + [H.heat3d_cfg(n=(19, 22, 25), center=(0.52, 0.5, 0.49), radius=0.2)]


def _batch_setup(ctx, precision, cfgs=BATCH_CFGS):
    import shoccs_amd as S

    pairs = [_pair(ctx, c) for c in cfgs]
    init, dts = [], []
    for osys, prob in pairs:
        u0 = osys.field()
        osys.initialize(u0, 0.0)
        osys.update_boundary(u0, 0.0)
        init.append(u0.flat())
        dts.append(osys.dt(1.0, 0.8))
        prob.set_tuning(3, 4, 16)
        prob.set_precision(precision)
    return pairs, init, dts


@pytest.mark.parametrize("precision", ["F64", "F32_COMP"])
@pytest.mark.parametrize("odd", [False, True])
def test_batch_matches_single_problems(ctx, precision, odd):
    """A batch of heat problems of different shapes steps every problem to
    the same doubles as stepping it alone (same kernels, same stage scalars),
    eagerly and by graph replay, and keeps their clocks."""
    import shoccs_amd as S

    prec = getattr(S, precision)
    pairs, init, dts = _batch_setup(ctx, prec, BATCH_CFGS_ODD if odd else BATCH_CFGS)
    want = []
    for (osys, prob), u, dt in zip(pairs, init, dts):
        g = prob.field()
        g.upload(u)
        prob.rk4_fused(g, 0.0, dt, 4)
        want.append(g.download())
        g.free()
    fields = []
    for (osys, prob), u in zip(pairs, init):
        f = prob.field()
        f.upload(u)
        fields.append(f)
    b = S.Batch(ctx, [p for _, p in pairs], fields, [0.0] * len(pairs), dts)
    b.steps(2)
    b.capture(1)
    b.launch()
    b.launch()
    for f, w in zip(fields, want):
        assert np.array_equal(f.download(), w)
    np.testing.assert_array_equal(b.times(), [4 * dt for dt in dts])
    b.close()
    for _, prob in pairs:
        prob.set_precision(S.F64)


@pytest.mark.parametrize("odd", [False, True])
def test_batch_matches_oracle(ctx, odd):
    """Every problem of a batch (different shapes and spheres; with an odd z
    extent the whole batch takes one point per lane) against the oracle's
    RK4 schedule stepping that problem alone: four steps, eagerly and by
    graph replay, to 1e-11 of the max-norm."""
    import shoccs_amd as S

    cfgs = (BATCH_CFGS_ODD if odd else BATCH_CFGS) + [
        H.heat3d_cfg(n=(30, 20, 22), center=(0.5, 0.47, 0.5), radius=0.22),
        H.heat3d_cfg(n=(24, 28, 18), center=(0.48, 0.53, 0.5), radius=0.2)]
    pairs, init, dts = _batch_setup(ctx, S.F64, cfgs)
    want = []
    for (osys, _), u, dt in zip(pairs, init, dts):
        u0 = osys.field()
        osys.initialize(u0, 0.0)
        osys.update_boundary(u0, 0.0)
        out, rk, srhs = osys.field(), osys.field(), osys.field()
        t = 0.0
        for _ in range(4):
            osys.rk4_step(u0, out, rk, srhs, t, dt)
            u0.D[:] = out.D
            for d in range(3):
                u0.R[d][:] = out.R[d]
            t += dt
        want.append(u0.flat())
    fields = []
    for (_, prob), u in zip(pairs, init):
        f = prob.field()
        f.upload(u)
        fields.append(f)
    b = S.Batch(ctx, [p for _, p in pairs], fields, [0.0] * len(pairs), dts)
    b.steps(2)
    b.capture(1)
    b.launch()
    b.launch()
    for q, (f, w) in enumerate(zip(fields, want)):
        err = np.max(np.abs(f.download() - w)) / np.max(np.abs(w))
        assert err <= 1e-11, f"problem {q}: {err:.2e}"
    b.close()


def test_batch64_full_size(ctx):
    """The 64-problem heat configuration at full size (per-axis sizes 32..128
    from seed 0, a Dirichlet sphere each, sin-product solution): 50 steps as
    one captured graph; every problem stays finite and within its
    discretisation error of the exact solution: below 5e-4 and below h_min^2
    for every problem at t = 50 dt (the second-order bound of the E4
    closures)."""
    import shoccs_amd as S

    rng = np.random.default_rng(0)
    probs, fields, dts, hs = [], [], [], []
    for _ in range(64):
        n = tuple(int(v) for v in rng.integers(32, 129, size=3))
        while True:
            c = rng.uniform(0.3, 0.7, size=3)
            r = float(rng.uniform(0.15, 0.35))
            try:
                p = S.Problem(ctx, S.SYS_HEAT, n, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0),
                              spheres=[[c[0], c[1], c[2], r]], shape_bc=[S.BC_DIRICHLET],
                              grid_bc=[S.BC_DIRICHLET] * 6, scheme=S.E4_2, params=[1.0, S.MMS_SINPROD])
                break
            except S.ShoccsError:
                pass
        f = p.field()
        p.initialize(f, 0.0)
        p.update_boundary(f, 0.0)
        probs.append(p)
        fields.append(f)
        dts.append(p.dt(1.0, 0.5))
        hs.append(1.0 / (min(n) - 1))
    b = S.Batch(ctx, probs, fields, [0.0] * 64, dts)
    b.capture(50)
    b.launch()
    times = b.times()
    errs = np.array([p.stats(f, t)[0] for p, f, t in zip(probs, fields, times)])
    np.testing.assert_allclose(times, [50 * dt for dt in dts], rtol=1e-12)
    assert np.all(np.isfinite(errs))
    # measured: max error 3.3e-4, max error / h_min^2 0.42
    assert errs.max() < 5e-4
    assert np.all(errs <= 1.0 * np.array(hs) ** 2)
    b.close()
    for p in probs:
        p.close()


def test_batch_rejects_mixed_families(ctx):
    import shoccs_amd as S

    (o1, p1), (o2, p2) = _pair(ctx, BATCH_CFGS[0]), _pair(ctx, WAVE_CFGS["wave30_radial_E4u"])
    with pytest.raises(S.ShoccsError):
        S.Batch(ctx, [p1, p2], [p1.field(), p2.field()], [0.0, 0.0], [1e-3, 1e-3])
