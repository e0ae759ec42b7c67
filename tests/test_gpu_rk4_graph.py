"""One problem's fused RK4 steps captured as a hipGraph (shx_rk4_graph_*):
every replay equals the same number of eager fused steps, value for value,
and advances the device clock; the inflow sphere (free cut points in the
special work) and the mixed form included; the oracle's RK4 schedule within
1e-11 after the replays."""
import numpy as np
import pytest

import helpers as H
import oracle_ffi as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,precision", [("inflow_e2p", 0), ("inflow_e2p", 1), ("dirichlet_e4u", 0),
                                            ("heat3d", 0)])
def test_graph_replays_equal_eager_steps(ctx, name, precision):
    import shoccs_amd as S

    if name == "inflow_e2p":
        cfg = H.wave_e2p_cfg(34)
    elif name == "dirichlet_e4u":
        cfg = H.sphere_wave_cfg(30, scheme=O.E4U_1, alpha=(0.1, 0.2))
    else:
        cfg = H.heat3d_cfg()
    prob = H.product_problem(ctx, cfg)
    prob.set_precision(S.F32_COMP if precision else S.F64)
    dt = prob.dt(0.5, 0.5)
    ua, ub = prob.field(), prob.field()
    for u in (ua, ub):
        prob.initialize(u, 0.0)
        prob.update_boundary(u, 0.0)
    nsteps = 3
    g = S.Rk4Graph(prob, ua, 0.0, dt, nsteps)
    for rep in range(2):
        g.launch()
        prob.rk4_fused(ub, rep * nsteps * dt, dt, nsteps)
        np.testing.assert_array_equal(ua.download(), ub.download())
        assert g.time() == pytest.approx((rep + 1) * nsteps * dt, rel=1e-15)
    g.close()
    if precision == 0:
        osys = H.oracle_system(cfg)
        u = osys.field()
        osys.initialize(u, 0.0)
        osys.update_boundary(u, 0.0)
        out, rk, srhs = osys.field(), osys.field(), osys.field()
        t = 0.0
        for _ in range(2 * nsteps):
            osys.rk4_step(u, out, rk, srhs, t, dt)
            u.D[:] = out.D
            for d in range(3):
                u.R[d][:] = out.R[d]
            t += dt
        want = u.flat()
        err = np.max(np.abs(ua.download() - want)) / np.max(np.abs(want))
        assert err <= 1e-11
    prob.close()


def test_graph_refuses_neumann_faces(ctx):
    import shoccs_amd as S

    cfg = H.heat3d_cfg()
    cfg["grid_bc"] = [O.NEUMANN] + [O.DIRICHLET] * 5
    prob = H.product_problem(ctx, cfg)
    u = prob.field()
    prob.initialize(u, 0.0)
    with pytest.raises(S.ShoccsError):
        S.Rk4Graph(prob, u, 0.0, prob.dt(0.5, 0.5), 2)
    prob.close()
