"""Pins of the oracle's E2-poly scheme (src/stencils/polyE2_1.cpp) and of the
per-cut-point inflow condition of the constant-velocity wave (CPU).

E2-poly's golden closure vector is checked in test_oracle.py with the other
schemes; here: query shapes, the reference's exactness cases
(polyE2_1.t.cpp: floating and Dirichlet closures differentiate a linear
function exactly, left and right), whole operators with a floating sphere
exact on linear fields at fluid and cut points (closures, interpolated
cross-direction rows), and the inflow classification (Dirichlet exactly where
a . n > 0)."""
import numpy as np
import pytest

import helpers as H
import oracle_ffi as O

FA = [0.1, 0.2, 0.3, 0.4, 0.5, 0.6]  # polyE2_1.t.cpp "floating", "dirichlet"
DA = [-0.1, -0.2, -0.3]
# scalar_wave.lua:29-40
LUA_FA = [0.7039278390946743, 0.5390086175376538, -0.647109821986589, 0.2051508287133347,
          0.6062051039572746, 0.8148425279273044]
LUA_DA = [-0.10739761225713096, 0.8736492896991024, 0.40413606410467495]
LUA_IA = [0.01, -0.01, 0.01, 0.04]


def test_query():
    """polyE2_1.cpp:25-41: P1 R3 T4, Dirichlet drops a row, no Neumann."""
    L = O.lib()
    st = O.stencil(O.E2P_1, FA + DA + [0] * 4)
    q = L.orc_query_max(st)
    assert (q.p, q.r, q.t, q.x) == (1, 3, 4, 0)
    d = L.orc_query(st, O.DIRICHLET)
    assert (d.r, d.t) == (2, 4)
    n = L.orc_query(st, O.NEUMANN)
    assert (n.r, n.t) == (0, 0)
    c = O.interior(st, 0.5)
    assert np.array_equal(c[:3], [-1.0, 0.0, 1.0])


@pytest.mark.parametrize("bc,psi", [(O.FLOATING, 0.2), (O.DIRICHLET, 0.0), (O.FLOATING, 1.0),
                                    (O.DIRICHLET, 0.73)])
@pytest.mark.parametrize("right", [False, True])
def test_closures_exact_on_linears(bc, psi, right):
    """polyE2_1.t.cpp "floating" / "dirichlet": mesh of t - 1 points on
    [-1, 5], the wall psi h outside it; every row differentiates 2x + 1 to 2.
    A floating row i sits at column i, a Dirichlet row at column i + 1 (its
    wall is data)."""
    st = O.stencil(O.E2P_1, FA + DA + [0] * 4)
    mesh = np.linspace(-1.0, 5.0, 3)
    h = mesh[1] - mesh[0]
    c, _, q = O.nbs(st, h, bc, psi, right)
    m = np.concatenate([mesh, [5.0 + psi * h]]) if right else np.concatenate([[-1.0 - psi * h], mesh])
    for i in range(q.r):
        assert abs(c[i * q.t:(i + 1) * q.t] @ (2 * m + 1) - 2.0) < 1e-13


def _linear_field(osys, f):
    m = osys.mesh
    u = osys.field()
    X, Y, Z = np.meshgrid(*[np.asarray(m.m.xyz[d][:m.m.n[d]]) for d in range(3)], indexing="ij")
    u.D[:] = f(X.ravel(), Y.ravel(), Z.ravel())
    for d in range(3):
        if m.nR(d):
            p = m.r_arrays(d)["pos"]
            u.R[d][:] = f(p[:, 0], p[:, 1], p[:, 2])
    return u, (X.ravel(), Y.ravel(), Z.ravel())


def test_gradient_exact_on_linears_floating_sphere():
    """E2-poly with a floating sphere: the gradient of a linear field is exact
    at the fluid points off the Dirichlet faces and at every cut point, whose
    rows interpolate across the other directions (derivative.cpp:159-257)."""
    cfg = H.sphere_wave_cfg(31, scheme=O.E2P_1, alpha=LUA_FA + LUA_DA + LUA_IA)
    cfg.update(spheres=[[0.52, 0.47, 0.49, 0.27]], obj_bc=[O.FLOATING])
    osys = H.oracle_system(cfg)
    m = osys.mesh
    g = (0.7, -1.3, 0.4)
    u, (X, Y, Z) = _linear_field(osys, lambda x, y, z: g[0] * x + g[1] * y + g[2] * z + 0.25)
    du = osys.gradient(u)
    fl = m.fluid()
    fl = fl[(X[fl] > 0) & (Y[fl] > 0) & (Z[fl] > 0)]  # Dirichlet inflow faces carry no rows
    for c in range(3):
        assert np.max(np.abs(du[c].D[fl] - g[c])) < 1e-12
        for d in range(3):
            assert m.nR(d) > 0
            assert np.max(np.abs(du[c].R[d] - g[c])) < 1e-12


def test_inflow_classification():
    """Shape condition INFLOW: a cut point is Dirichlet where the velocity
    leaves the body into the fluid (a . n > 0, n the outward normal) and
    floating elsewhere.  On a linear field the rhs -a . grad u is exactly 0 at
    the Dirichlet cut points (scalar_wave.cpp:70-95 zeroes them) and equals
    -a . g, computed from the floating rows, at the others."""
    a = 1 / np.sqrt(3.0)
    cfg = H.sphere_wave_cfg(31, scheme=O.E2P_1, alpha=LUA_FA + LUA_DA + LUA_IA)
    cfg["obj_bc"] = [O.INFLOW]
    osys = H.oracle_system(cfg)
    m = osys.mesh
    g = (0.7, -1.3, 0.4)
    u, _ = _linear_field(osys, lambda x, y, z: g[0] * x + g[1] * y + g[2] * z)
    rhs = osys.rhs(u, 0.0)
    want = -a * sum(g)
    nd = nf = 0
    for d in range(3):
        n = (m.r_arrays(d)["pos"] - 0.5) / 0.3  # the sphere's outward normal, closed form
        an = a * n.sum(axis=1)
        dirichlet = an > 0
        nd += dirichlet.sum()
        nf += (~dirichlet).sum()
        assert np.all(rhs.R[d][dirichlet] == 0.0)
        assert np.max(np.abs(rhs.R[d][~dirichlet] - want)) < 1e-11
    assert nd > 0 and nf > 0


def test_inflow_needs_a_velocity():
    """INFLOW is resolved from the constant velocity: other systems refuse it."""
    cfg = H.heat3d_cfg(n=(12, 13, 14))
    cfg["obj_bc"] = [O.INFLOW]
    with pytest.raises(ValueError):
        H.oracle_system(cfg)


def _linf_history(cfg, steps, every):
    s = H.oracle_system(cfg)
    u = s.field()
    s.initialize(u, 0.0)
    s.update_boundary(u, 0.0)
    out, rk, srhs = s.field(), s.field(), s.field()
    dt, t, hist = s.dt(0.5, 1.0), 0.0, []
    for k in range(steps):
        s.rk4_step(u, out, rk, srhs, t, dt)
        t += dt
        u, out = out, u
        if (k + 1) % every == 0:
            hist.append(s.stats(u, t)[0])
    return hist


def test_reference_schemes_grow_on_the_inflow_sphere():
    """Config 2's wave with the inflow sphere at 48^3, CFL 0.5: the
    reference's 4th-order E4u (its scalar-wave parameters) and its E2-poly
    (scalar_wave.lua's parameters) both grow without bound within 100 steps,
    while the same wave without the sphere stays at its discretisation
    error."""
    e4u = H.sphere_wave_cfg(48, scheme=O.E4U_1, alpha=(-0.7733323791884821, 0.1623961700641681))
    e4u["obj_bc"] = [O.INFLOW]
    h = _linf_history(e4u, 50, 25)
    assert h[-1] > 1e3 * h[0] and h[-1] > 1e3, h
    h = _linf_history(H.wave_e2p_cfg(48), 100, 25)
    assert h[-1] > 10.0 and h[-1] > 10 * min(h), h
    free = dict(e4u, spheres=[], obj_bc=[])
    h = _linf_history(free, 100, 25)
    assert max(h) < 0.2, h
