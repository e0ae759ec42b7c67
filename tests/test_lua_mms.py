"""Manufactured solutions given by functions of (time, location), as the
reference's lua_mms (src/mms/lua_mms.hpp:12-79) calls its Lua functions, on the
oracle (CPU).  The known answers are the reference's own tests:

* heat.t.cpp:43-144 "heat - E2" (Dirichlet sphere, a cubic solution),
  :146-256 "heat - E2 - floating" and :258-364 "2D heat - E2 - floating"
  (floating spheres, solutions linear in each variable): solutions whose
  laplacian the E2 operators reproduce exactly, so after initialize and
  update_boundary at t = 0 the rhs is cos(0) = 1 at every fluid point and every
  free cut point, zero at solid points and on the Dirichlet grid faces;
* rk4_v2.t.cpp:28-143: one RK4 step of a solution linear in time is exact
  (L-inf error below 1e-13);
* simulation_cycle.t.cpp:22-99 "cycle - 2D": 5 steps of the 21 x 22 heat
  problem with a floating circle reach t = 0.0125 with L-inf error < 0.05.

The functions below are the Lua functions of those tests, written in Python
with the same operations in the same order.
"""
import math

import numpy as np
import pytest

import oracle_ffi as O

MESH3 = dict(n=(21, 22, 23), lo=(1.0, 1.1, 0.3), hi=(3.0, 3.3, 2.2))
MESH2 = dict(n=(21, 22, 1), lo=(1.0, 1.1, 0.0), hi=(3.0, 3.3, 0.0))
SPHERE3 = [2.0001, 2.5656565, 1.313131311, 0.25]
SPHERE2 = [2.0001, 2.5656565, 0.0, 0.25]
D, F, N = O.DIRICHLET, O.FLOATING, O.NEUMANN
# xmin dirichlet, ymin/ymax neumann, zmax dirichlet (heat.t.cpp:55-60)
BC3 = [D, F, N, N, F, D]
BC2 = [D, F, N, N, F, F]


def cubic(timefn, dtfn):
    """heat.t.cpp:79-101 (timefn = math.sin, dtfn = math.cos) and
    rk4_v2.t.cpp:69-93 (time itself, ddt 1)."""
    def call(t, loc):
        x, y, z = loc
        return (timefn(t) + x * x * (y + z) + y * y * (x + z) + z * z * (x + y) + 3 * x * y * z + x + y + z)

    def ddt(t, loc):
        return dtfn(t)

    def grad(t, loc):
        x, y, z = loc
        return (2. * x * (y + z) + y * y + z * z + 3. * y * z + 1,
                x * x + 2. * y * (x + z) + z * z + 3. * x * z + 1,
                x * x + y * y + 2. * z * (x + y) + 3. * x * y + 1)

    def lap(t, loc):
        x, y, z = loc
        return 2. * (y + z) + 2. * (x + z) + 2. * (x + y)

    return call, ddt, grad, lap


def bilinear3():
    """heat.t.cpp:182-204 (3-D floating case)."""
    def call(t, loc):
        x, y, z = loc
        return (math.sin(t) + x * (y + z) + y * (x + z) + z * (x + y) + x + y + z)

    def grad(t, loc):
        x, y, z = loc
        return 2. * (y + z) + 1, 2. * (x + z) + 1, 2. * (x + y) + 1

    return call, (lambda t, loc: math.cos(t)), grad, (lambda t, loc: 0.0)


def bilinear():
    """heat.t.cpp:294-312 (2-D floating case)."""
    def call(t, loc):
        x, y, z = loc
        return (math.sin(t) + x * y + x + y)

    return call, (lambda t, loc: math.cos(t)), (lambda t, loc: (loc[1] + 1, loc[0] + 1, 0.0)), \
        (lambda t, loc: 0.0)


def cycle2d():
    """simulation_cycle.t.cpp:62-83."""
    def call(t, loc):
        x, y, z = loc
        return (t + x * x * y + y * y * x + 3 * x * y + x + y)

    def grad(t, loc):
        x, y, z = loc
        return 2. * x * y + y * y + 3. * y + 1, x * x + 2. * y * x + 3. * x + 1, 0

    def lap(t, loc):
        x, y, z = loc
        return 2. * y + 2. * x

    return call, (lambda t, loc: 1.0), grad, lap


def host_system(mesh, sphere, body_bc, grid_bc, k, fns):
    st = O.stencil(O.E2_2, (0.0, 0.0))
    osys = O.OracleSystem(O.SYS_HEAT, mesh["n"], mesh["lo"], mesh["hi"], [sphere], [body_bc], grid_bc, st,
                          [k, O.MMS_HOST])
    osys.set_mms(O.mms_fns(*fns))
    return osys


@pytest.mark.parametrize("case", ["heat - E2", "heat - E2 - floating", "2D heat - E2 - floating"])
def test_heat_rhs_known_answer(case):
    if case == "heat - E2":
        osys = host_system(MESH3, SPHERE3, D, BC3, 0.1, cubic(math.sin, math.cos))
    elif case == "heat - E2 - floating":
        osys = host_system(MESH3, SPHERE3, F, BC3, 0.1, bilinear3())
    else:
        osys = host_system(MESH2, SPHERE2, F, BC2, 0.1, bilinear())
    n = osys.mesh.m.n
    u = osys.field()
    osys.initialize(u, 0.0)
    solid = int(np.count_nonzero(u.D == 0.0))
    assert abs(osys.stats(u, 0.0)[0]) <= 1e-13
    osys.update_boundary(u, 0.0)
    r = osys.rhs(u, 0.0)
    rhs_zero = int(np.count_nonzero(r.D == 0.0))
    if n[2] > 1:  # the Dirichlet faces xmin and zmax, sharing one edge of n[1] points
        faces = n[1] * n[2] + n[0] * n[1] - n[1]
    else:
        faces = n[1]
    assert solid == rhs_zero - faces
    npts = n[0] * n[1] * n[2]
    assert r.D.sum() == pytest.approx(npts - rhs_zero, rel=1.2e-5)
    fluid = np.zeros(npts, bool)
    m = osys.mesh.m
    fluid[np.ctypeslib.as_array(m.fluid, (m.nfluid,))] = True
    onface = r.D == 0.0
    assert np.allclose(r.D[fluid & ~onface], 1.0, rtol=0, atol=1e-9)
    if case != "heat - E2":
        for d in range(3 if n[2] > 1 else 2):
            nr = len(r.R[d])
            assert nr > 0
            assert r.R[d].sum() == pytest.approx(float(nr), rel=1.2e-5)


def test_rk4_step_known_answer():
    osys = host_system(MESH3, SPHERE3, D, BC3, 1.0, cubic(lambda t: t, lambda t: 1.0))
    u = osys.field()
    osys.initialize(u, 0.0)
    osys.update_boundary(u, 0.0)
    out, rk, srhs = osys.field(), osys.field(), osys.field()
    dt = osys.dt(1.0, 1.0)
    osys.rk4_step(u, out, rk, srhs, 0.0, dt)
    assert osys.stats(out, dt)[0] <= 1e-13


def test_cycle_2d_known_answer():
    osys = host_system(MESH2, SPHERE2, F, BC2, 1.0, cycle2d())
    u = osys.field()
    osys.initialize(u, 0.0)
    osys.update_boundary(u, 0.0)
    out, rk, srhs = osys.field(), osys.field(), osys.field()
    dt = osys.dt(1.0, 1.0)
    t = 0.0
    for _ in range(5):
        osys.rk4_step(u, out, rk, srhs, t, dt)
        u.D[:] = out.D
        for d in range(3):
            u.R[d][:] = out.R[d]
        t += dt
    assert t == pytest.approx(0.0125, abs=1e-10)
    assert osys.stats(u, t)[0] < 0.05
