"""Heat problems whose manufactured solution is given by functions of (time,
location), evaluated on the host as the reference's lua_mms does
(src/mms/lua_mms.hpp:12-79), on the GPU: the reference's own known answers
(heat.t.cpp:43-364, rk4_v2.t.cpp:28-143, simulation_cycle.t.cpp:22-99) and the
oracle's values on the same problems."""
import math

import numpy as np
import pytest

import oracle_ffi as O
import test_lua_functions as LF
import test_lua_mms as T

pytestmark = pytest.mark.gpu


def _pair(ctx, mesh, sphere, body_bc, grid_bc, k, fns):
    """The product's problem (functions called through ctypes) and the
    oracle's, on the same configuration."""
    import shoccs_amd as S

    prob = S.Problem(ctx, S.SYS_HEAT, mesh["n"], mesh["lo"], mesh["hi"], spheres=[sphere], shape_bc=[body_bc],
                     grid_bc=grid_bc, scheme=S.E2_2, alpha=(0.0, 0.0), params=[k, S.MMS_HOST],
                     mms=S.mms_functions(*fns))
    return prob, T.host_system(mesh, sphere, body_bc, grid_bc, k, fns)


@pytest.mark.parametrize("case", ["heat - E2", "heat - E2 - floating", "2D heat - E2 - floating"])
def test_heat_rhs_known_answer_and_oracle(ctx, case):
    if case == "heat - E2":
        prob, osys = _pair(ctx, T.MESH3, T.SPHERE3, T.D, T.BC3, 0.1, T.cubic(math.sin, math.cos))
    elif case == "heat - E2 - floating":
        prob, osys = _pair(ctx, T.MESH3, T.SPHERE3, T.F, T.BC3, 0.1, T.bilinear3())
    else:
        prob, osys = _pair(ctx, T.MESH2, T.SPHERE2, T.F, T.BC2, 0.1, T.bilinear())
    u, r = prob.field(), prob.field()
    prob.initialize(u, 0.0)
    ou = osys.field()
    osys.initialize(ou, 0.0)
    np.testing.assert_array_equal(u.download(), ou.flat())
    assert abs(prob.stats(u, 0.0)[0]) <= 1e-13
    prob.update_boundary(u, 0.0)
    osys.update_boundary(ou, 0.0)
    np.testing.assert_array_equal(u.download(), ou.flat())
    prob.rhs(u, r, 0.0)
    got = r.download()
    np.testing.assert_array_equal(got, osys.rhs(ou, 0.0).flat())
    # heat.t.cpp: rhs 1 at fluid points and free cut points, zero at solid
    # points and on the Dirichlet faces
    n = prob.n
    N = n[0] * n[1] * n[2]
    D = got[:N]
    solid = int(np.count_nonzero(u.download()[:N] == 0.0))
    faces = n[1] * n[2] + n[0] * n[1] - n[1] if n[2] > 1 else n[1]
    assert solid == int(np.count_nonzero(D == 0.0)) - faces
    assert D.sum() == pytest.approx(N - int(np.count_nonzero(D == 0.0)), rel=1.2e-5)
    if case != "heat - E2":
        off = N
        for d in range(3 if n[2] > 1 else 2):
            nr = prob.nR[d]
            assert nr > 0 and got[off:off + nr].sum() == pytest.approx(float(nr), rel=1.2e-5)
            off += nr
    prob.close()


def test_rk4_step_known_answer(ctx):
    """rk4_v2.t.cpp: one RK4 step of a solution linear in time, exact."""
    prob, osys = _pair(ctx, T.MESH3, T.SPHERE3, T.D, T.BC3, 1.0, T.cubic(lambda t: t, lambda t: 1.0))
    u0, out, rk, srhs = prob.field(), prob.field(), prob.field(), prob.field()
    prob.initialize(u0, 0.0)
    prob.update_boundary(u0, 0.0)
    dt = prob.dt(1.0, 1.0)
    prob.rk4(u0, out, rk, srhs, 0.0, dt)
    assert prob.stats(out, dt)[0] <= 1e-13
    ou, oo, ork, osr = osys.field(), osys.field(), osys.field(), osys.field()
    osys.initialize(ou, 0.0)
    osys.update_boundary(ou, 0.0)
    osys.rk4_step(ou, oo, ork, osr, 0.0, dt)
    np.testing.assert_allclose(out.download(), oo.flat(), rtol=0, atol=1e-11 * np.max(np.abs(oo.flat())))
    prob.close()


CYCLE_2D = """
simulation = {
    mesh = { index_extents = {21, 22}, domain_bounds = { min = {1, 1.1}, max = {3, 3.3} } },
    domain_boundaries = { xmin = "dirichlet", ymin = "neumann", ymax = "neumann" },
    shapes = { { type = "sphere", center = {2.0001, 2.5656565}, radius = 0.25, boundary_condition = "floating" } },
    scheme = { order = 2, type = "E2" },
    system = { type = "heat", diffusivity = 1.0 },
    integrator = { type = "rk4" },
    step_controller = { max_step = 5 },
    manufactured_solution = {
        type = "lua",
        call = function(time, loc)
            local x, y, z = loc[1], loc[2], loc[3]
            return (time +
                x * x * y + y * y * x + 3 * x * y + x + y)
        end,
        ddt = function(time, loc)
            return 1.0
        end,
        grad = function(time, loc)
            local x, y, z = loc[1], loc[2], loc[3]
            return 2. * x * y + y * y + 3. * y + 1,
                    x * x + 2. * y * x + 3. * x + 1,
                    0
        end,
        lap = function(time, loc)
            local x, y, z = loc[1], loc[2], loc[3]
            return 2. * y + 2. * x
        end,
        div = function(time, loc)
            return 0.0
        end
    }
}
"""


def test_cycle_2d_from_lua(ctx):
    """simulation_cycle.t.cpp "cycle - 2D": the configuration's Lua functions
    drive shoccs_run; t = 0.0125 and L-inf < 0.05 after 5 steps, and the run
    equals the oracle's cycle with the same solution."""
    import shoccs_amd as S

    cfg = S.Config(CYCLE_2D)
    prob = cfg.problem(ctx)
    u = prob.field()
    res = prob.run(u, **cfg.run_kwargs())
    assert res["step"] == 5 and res["completed"]
    assert res["time"] == pytest.approx(0.0125, abs=1e-10)
    assert res["linf"] < 0.05
    assert cfg.mms_error() == ""
    osys = T.host_system(T.MESH2, T.SPHERE2, T.F, T.BC2, 1.0, T.cycle2d())
    ou = osys.field()
    osys.initialize(ou, 0.0)
    osys.update_boundary(ou, 0.0)
    out, rk, srhs = osys.field(), osys.field(), osys.field()
    dt = osys.dt(1.0, 1.0)
    t = 0.0
    for _ in range(5):
        osys.rk4_step(ou, out, rk, srhs, t, dt)
        ou.D[:] = out.D
        for d in range(3):
            ou.R[d][:] = out.R[d]
        t += dt
    want = ou.flat()
    err = np.max(np.abs(u.download() - want)) / np.max(np.abs(want))
    assert err <= 1e-11
    assert res["linf"] == pytest.approx(osys.stats(ou, t)[0], rel=1e-9)
    prob.close()


def test_heat_e2_from_lua(ctx):
    """The "heat - E2" configuration text of heat.t.cpp read by the Lua
    subset: its functions give the same rhs as the Python ones."""
    import shoccs_amd as S

    cfg = S.Config(LF.HEAT_E2 % ("dirichlet", LF.CUBIC))
    prob = cfg.problem(ctx)
    u, r = prob.field(), prob.field()
    prob.initialize(u, 0.0)
    prob.update_boundary(u, 0.0)
    prob.rhs(u, r, 0.0)
    osys = T.host_system(T.MESH3, T.SPHERE3, T.D, T.BC3, 0.1, T.cubic(math.sin, math.cos))
    ou = osys.field()
    osys.initialize(ou, 0.0)
    osys.update_boundary(ou, 0.0)
    np.testing.assert_array_equal(r.download(), osys.rhs(ou, 0.0).flat())
    prob.close()


def test_host_solution_refuses_fused_forms(ctx):
    import shoccs_amd as S

    prob, _ = _pair(ctx, T.MESH3, T.SPHERE3, T.D, T.BC3, 1.0, T.cubic(lambda t: t, lambda t: 1.0))
    u = prob.field()
    prob.initialize(u, 0.0)
    with pytest.raises(S.ShoccsError, match="host-evaluated"):
        prob.rk4_fused(u, 0.0, prob.dt(1.0, 1.0), 1)
    prob.close()
