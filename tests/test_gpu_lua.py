"""Runs driven by a Lua `simulation` configuration (shoccs_config_parse +
shoccs_run) against the oracle's cycle of the same problem."""
import numpy as np
import pytest

import oracle_ffi as O

pytestmark = pytest.mark.gpu

RADIAL_2D = """
-- the reference's 2-D varying-coefficient wave (lua-configs/brady_livescu_4_3_n61.lua), short
simulation = {
    mesh = { index_extents = {31, 31}, domain_bounds = {math.sqrt(2), math.sqrt(2)} },
    domain_boundaries = { xmin = "dirichlet", ymin = "dirichlet" },
    shapes = {},
    scheme = { order = 1, type = "E4u", alpha = {-0.7733323791884821, 0.1623961700641681} },
    system = { type = "scalar wave", center = {-0.25, -0.25, 0}, radius = 0, max_error = 10.0 },
    integrator = { type = "rk4"%s },
    step_controller = { max_step = 25, cfl = { hyperbolic = 0.8 } },
}
"""

HEAT_2D = """
-- config 1: 2-D heat with a Dirichlet circle, sin-product solution, 101^2 (20 steps)
simulation = {
    mesh = { index_extents = {101, 101}, domain_bounds = {1, 1} },
    domain_boundaries = { xmin = "dirichlet", xmax = "dirichlet", ymin = "dirichlet", ymax = "dirichlet" },
    shapes = { { type = "sphere", center = {0.5, 0.5}, radius = 0.23, boundary_condition = "dirichlet" } },
    scheme = { order = 2, type = "E4" },
    system = { type = "heat", diffusivity = 1 },
    manufactured_solution = { type = "sin_product" },
    integrator = { type = "rk4" },
    step_controller = { max_step = 20, cfl = { parabolic = 0.8 } },
}
"""


def _oracle_cycle(cfg, nsteps, h_cfl, p_cfl):
    st = O.stencil(cfg.scheme, cfg.alpha)
    obj_bc = cfg.shape_bc
    osys = O.OracleSystem(cfg.kind, cfg.n, cfg.lo, cfg.hi, cfg.spheres if cfg.spheres else np.zeros((0, 4)),
                          obj_bc, list(cfg.grid_bc), st, cfg.params)
    u = osys.field()
    osys.initialize(u, 0.0)
    osys.update_boundary(u, 0.0)
    out, rk, srhs = osys.field(), osys.field(), osys.field()
    dt = osys.dt(h_cfl, p_cfl)
    t = 0.0
    for _ in range(nsteps):
        osys.rk4_step(u, out, rk, srhs, t, dt)
        u.D[:] = out.D
        for d in range(3):
            u.R[d][:] = out.R[d]
        t += dt
    return u.flat(), t


@pytest.mark.parametrize("text,steps", [(RADIAL_2D % "", 25), (RADIAL_2D % ', form = "fused"', 25),
                                        (HEAT_2D, 20)])
def test_lua_configured_run_matches_oracle(ctx, text, steps):
    import shoccs_amd as S

    cfg = S.Config(text)
    prob = cfg.problem(ctx)
    u = prob.field()
    res = prob.run(u, **cfg.run_kwargs())
    assert res["step"] == steps and res["completed"]
    want, t = _oracle_cycle(cfg, steps, cfg.h_cfl, cfg.p_cfl)
    assert res["time"] == pytest.approx(t, rel=1e-15)
    got = u.download()
    err = np.max(np.abs(got - want)) / np.max(np.abs(want))
    assert err <= 1e-11, f"relative max-norm error {err:.3e}"
    prob.close()
