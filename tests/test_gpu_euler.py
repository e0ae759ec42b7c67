"""The compressible Euler system on the device (shx_euler_*) against the CPU
oracle (oracle/euler.c): wall update, rhs and RK4 steps, value for value.
Every step is plain FP64 arithmetic and the derivative applies are
bit-identical.  The device applies the derivatives to negated fluxes where the
oracle negates the summed rhs; negation commutes with rounding to nearest, so
only an exact zero may carry the other sign (np.array_equal counts -0.0 equal
to 0.0)."""
import numpy as np
import pytest

import helpers as H
import oracle_ffi as O
from test_euler_oracle import E4U_ALPHA, euler_cfg, set_state, vortex

pytestmark = pytest.mark.gpu

F, D = O.FLOATING, O.DIRICHLET

CASES = {
    "sphere3d": euler_cfg((24, 22, 20), (0, 0, 0), (1, 1, 1), spheres=[(0.5, 0.48, 0.52, 0.26)],
                          grid_bc=[D, F, F, F, F, F]),
    "sphere3d_noncubic": euler_cfg((30, 18, 26), (0, 0, 0), (1.2, 0.8, 1.0), spheres=[(0.6, 0.4, 0.5, 0.2)],
                                   grid_bc=[D, F, D, F, F, F]),
    "vortex2d": euler_cfg((41, 41), (-5, -5, 0), (5, 5, 0)),
}
# large enough for the four-planes-per-lane pass along x (at least 2048
# workgroups of four planes), extents off the tile and plane multiples
LARGE = {"sphere3d_multi_plane": euler_cfg((130, 66, 258), (0, 0, 0), (1.0, 0.5, 2.0),
                                           spheres=[(0.5, 0.26, 1.0, 0.2)], grid_bc=[D, F, F, D, F, F])}


def _state(e, name, rng):
    """A smooth flow (or the vortex) with a small random part, positive
    density and pressure everywhere (solid points included)."""
    if name == "vortex2d":
        f = vortex(0.2)
    else:
        def f(x, y, z):
            rho = 1.0 + 0.1 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) * np.cos(np.pi * z)
            u = 0.3 + 0.05 * np.cos(2 * np.pi * z)
            v = 0.05 * np.sin(2 * np.pi * x)
            w = -0.04 * np.cos(2 * np.pi * y)
            p = 1 / 1.4 + 0.05 * np.cos(2 * np.pi * (x + y))
            return [rho, rho * u, rho * v, rho * w, p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)]
    U = e.fields()
    set_state(e, U, f)
    for k in range(5):
        U[k].D[:] *= 1.0 + 1e-3 * rng.standard_normal(U[k].D.size)
    return U


def _pair(ctx, name):
    import shoccs_amd as S

    cfg = {**CASES, **LARGE}[name]
    e = O.OracleEuler(cfg)
    pcfg = dict(kind=O.SYS_HEAT, n=cfg["n"], lo=cfg["lo"], hi=cfg["hi"], spheres=cfg["spheres"],
                obj_bc=[D] * len(cfg["spheres"]), grid_bc=cfg["grid_bc"], scheme=cfg["scheme"], alpha=cfg["alpha"],
                params=[1.0, O.MMS_SINPROD])
    prob = H.product_problem(ctx, pcfg)
    return e, prob, S.Euler(prob)


def _upload(ge, U):
    G = ge.fields()
    ge.upload(G, [u.flat() for u in U])
    return G


def _flat(U):
    return [u.flat() for u in U]


@pytest.mark.parametrize("name", list(CASES) + list(LARGE))
def test_wall_and_rhs_bitexact(ctx, name):
    e, prob, ge = _pair(ctx, name)
    U = _state(e, name, np.random.default_rng(3))
    G = _upload(ge, U)
    e.wall(U)
    ge.wall(G)
    for a, b in zip(ge.download(G), _flat(U)):
        assert np.array_equal(a, b)
    want = e.rhs(U)
    out = ge.fields()
    ge.rhs(G, out)
    for k, (a, b) in enumerate(zip(ge.download(out), _flat(want))):
        assert np.array_equal(a, b), f"{name} component {k}: {np.max(np.abs(a - b))}"
    assert max(np.max(np.abs(w.D)) for w in want) > 0
    ge.close()
    prob.close()


def _dt(e, U, cfl=0.4):
    """cfl h / (|u| + c)_max"""
    rho, m, E = U[0].D, np.stack([U[k].D for k in (1, 2, 3)]), U[4].D
    p = 0.4 * (E - 0.5 * np.sum(m * m, axis=0) / rho)
    smax = np.max(np.sqrt(np.sum(m * m, axis=0)) / rho + np.sqrt(1.4 * p / rho))
    return cfl * min(h for h in e.mesh.h() if h > 0) / smax


@pytest.mark.parametrize("name", list(CASES))
def test_rk4_steps_bitexact(ctx, name):
    """Four RK4 steps (rk4.cpp:16-47 over the five fields), u0 <- out between
    steps, at dt = 0.4 h / (|u| + c)_max."""
    e, prob, ge = _pair(ctx, name)
    U = _state(e, name, np.random.default_rng(4))
    e.wall(U)
    dt = _dt(e, U)
    G = _upload(ge, U)
    out, rk, srhs = e.fields(), e.fields(), e.fields()
    Go, Gr, Gs = ge.fields(), ge.fields(), ge.fields()
    for _ in range(4):
        e.rk4_step(U, out, rk, srhs, dt)
        ge.rk4_step(G, Go, Gr, Gs, dt)
        for k in range(5):
            U[k].D[:] = out[k].D
            for d in range(3):
                U[k].R[d][:] = out[k].R[d]
            G[k].copy_from(Go[k])
    for k, (a, b) in enumerate(zip(ge.download(G), _flat(U))):
        assert np.array_equal(a, b), f"{name} component {k}: {np.max(np.abs(a - b))}"
    ms = ge.time_steps(G, Go, Gr, Gs, dt, 2)
    assert ms > 0
    ge.close()
    prob.close()


def _device_steps(ge, U, dt, nsteps):
    """nsteps device steps through rk4_step, u0 <- out between them"""
    G = _upload(ge, U)
    Go, Gr, Gs = ge.fields(), ge.fields(), ge.fields()
    for _ in range(nsteps):
        ge.rk4_step(G, Go, Gr, Gs, dt)
        for k in range(5):
            G[k].copy_from(Go[k])
    return ge.download(G)


@pytest.mark.parametrize("precision", ["F64", "F32_COMP"])
@pytest.mark.parametrize("nsteps", [1, 2, 3])
def test_time_steps_match_steps(ctx, precision, nsteps):
    """The timing loop (u0 and out swapping roles from step to step) leaves
    the result of nsteps single steps in both u0 and out, value for value."""
    import shoccs_amd as S

    e, prob, ge = _pair(ctx, "sphere3d_noncubic")
    ge.set_precision(getattr(S, precision))
    U = _state(e, "sphere3d_noncubic", np.random.default_rng(5))
    e.wall(U)
    dt = _dt(e, U)
    want = _device_steps(ge, U, dt, nsteps)
    G = _upload(ge, U)
    Go, Gr, Gs = ge.fields(), ge.fields(), ge.fields()
    ge.time_steps(G, Go, Gr, Gs, dt, nsteps)
    for got in (ge.download(G), ge.download(Go)):
        for k in range(5):
            assert np.array_equal(got[k], want[k]), f"{precision} {nsteps} component {k}"
    ge.close()
    prob.close()


@pytest.mark.parametrize("name", list(CASES) + list(LARGE))
def test_mixed_form_against_fp64(ctx, name):
    """The mixed form (FP32 stage inputs and interior flux-derivative rows with
    Kahan-compensated sums; FP64 closures, B terms, cut points and RK
    registers) after four steps: within 1e-5 of the FP64 steps relative to
    the field's largest magnitude (for a momentum component, the largest
    momentum component's), and not identical to them.  Expected size: each
    FP32 stage input is rounded by about 1e-7, the derivative scales that by
    about 1.5 / h and the stage by dt = 0.4 h / 1.3, some 5e-8 per stage and
    1e-6 over the sixteen stages."""
    import shoccs_amd as S

    e, prob, ge = _pair(ctx, name)
    U = _state(e, name, np.random.default_rng(6))
    e.wall(U)
    dt = _dt(e, U)
    want = _device_steps(ge, U, dt, 4)
    ge.set_precision(S.F32_COMP)
    got = _device_steps(ge, U, dt, 4)
    worst = 0.0
    mom = max(np.max(np.abs(want[k])) for k in (1, 2, 3))
    for k in range(5):
        assert np.all(np.isfinite(got[k]))
        scale = mom if k in (1, 2, 3) else np.max(np.abs(want[k]))
        rel = np.max(np.abs(got[k] - want[k])) / scale
        worst = max(worst, rel)
        assert rel <= 1e-5, f"{name} component {k}: {rel:.3e}"
    assert worst > 0.0
    ge.close()
    prob.close()


def test_euler_refuses_cut_point_rows(ctx):
    import shoccs_amd as S

    cfg = H.cut_cfgs()["floating_E4u_1"]
    prob = H.product_problem(ctx, cfg)
    with pytest.raises(S.ShoccsError):
        S.Euler(prob)
    prob.close()


def test_euler_refuses_one_direction(ctx):
    """The last direction's pass completes each RK stage, so a mesh with one
    direction is refused."""
    import shoccs_amd as S

    cfg = dict(kind=O.SYS_HEAT, n=(40, 1, 1), lo=(0, 0, 0), hi=(1, 0, 0), spheres=[], obj_bc=[],
               grid_bc=[D, F, F, F, F, F], scheme=O.E4U_1, alpha=E4U_ALPHA, params=[1.0, O.MMS_SINPROD])
    prob = H.product_problem(ctx, cfg)
    with pytest.raises(S.ShoccsError):
        S.Euler(prob)
    prob.close()
