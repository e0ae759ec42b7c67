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


HEAT_FLOATING_2D = """
-- 2-D heat with a floating circle, two gaussians, a Neumann face (the kind of
-- problem the reference's heat.lua sets up)
local dx = 1 / 40
simulation = {
    mesh = { index_extents = {41, 41}, domain_bounds = {1, 1} },
    domain_boundaries = { xmin = "neumann", xmax = "dirichlet", ymin = "dirichlet", ymax = "dirichlet" },
    shapes = { { type = "sphere", center = {0.52, 0.5}, radius = 0.2, boundary_condition = "floating" } },
    scheme = { order = 2, type = "E4" },
    system = { type = "heat", diffusivity = 1 / 30 },
    manufactured_solution = { type = "gaussian",
        { center = {0.4, 0.6}, variance = {0.25, 0.2}, amplitude = 1.5, frequency = 2 },
        { center = {0.7, 0.3}, variance = {0.3, 0.35}, amplitude = -10 * dx, frequency = 0.5 } },
    integrator = { type = "rk4" },
    step_controller = { max_step = 12, cfl = { parabolic = 0.5 } },
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
                                        (HEAT_2D, 20), (HEAT_FLOATING_2D, 12)])
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


@pytest.mark.parametrize("fused", [False, True])
def test_configured_run_writes_fields(ctx, tmp_path, fused):
    """The io table of a configuration (field_io, src/io/field_io.cpp:32-65):
    dumps at step 0 and every second step, U as the run left it, Error =
    |U - exact| by write_scalar_error, geometry and XDMF alongside."""
    import xml.etree.ElementTree as ET

    import shoccs_amd as S

    text = RADIAL_2D % (', form = "fused"' if fused else "")
    text = text.replace("step_controller = { max_step = 25,", "step_controller = { max_step = 5,")
    text = text.replace("\n}\n", f"\n    io = {{ write_every_step = 2, dir = '{tmp_path / 'io'}' }},\n}}\n")
    cfg = S.Config(text)
    assert cfg.io["write_every_step"] == 2
    prob = cfg.problem(ctx)
    u = prob.field()
    io = S.FieldIO(prob, **cfg.io)
    res = prob.run(u, io=io, **cfg.run_kwargs())
    io.close()
    assert res["step"] == 5
    d = tmp_path / "io"
    assert sorted(p.name for p in d.iterdir() if p.name.startswith("U.")) == ["U.000000", "U.000001", "U.000002"]
    # dump 0: the initialized field with its boundary values
    u0 = prob.field()
    prob.initialize(u0, 0.0)
    prob.update_boundary(u0, 0.0)
    np.testing.assert_array_equal(np.fromfile(d / "U.000000"), u0.download())
    # dump 2 is step 4: rerun 4 steps and compare U and its Error
    u4 = prob.field()
    res4 = prob.run(u4, **dict(cfg.run_kwargs(), max_step=4))
    U = np.fromfile(d / "U.000002")
    np.testing.assert_array_equal(U, u4.download())
    ex = prob.field()
    prob.exact(ex, res4["time"])
    np.testing.assert_array_equal(np.fromfile(d / "Error.000002"), prob.error_field(U, ex.download()))
    root = ET.fromstring((d / "view.xmf").read_text().split("\n", 2)[2])
    grids = root.find("Domain").find("Grid").findall("Grid")
    assert [g.get("Name") for g in grids] == ["0", "1", "2"]
    assert float(grids[2].find("Time").get("Value")) == res4["time"]
    prob.close()
