"""Pins of the oracle's scalar-wave exact solutions and boundary updates
(CPU).  The values are compared with closed forms evaluated here in numpy,
so a sign, a 2 pi or a shift slip in the oracle's formulas fails.

* radial wave: the reference's own check (src/systems/scalar_wave.t.cpp:43-145):
  21 x 22 x 23 mesh on [1, 3] x [1.1, 3.3] x [0.3, 2.2], Dirichlet sphere
  centre (2.0001, 2.5656565, 1.313131311) radius 0.25 (the system takes the
  first shape's centre and radius, scalar_wave.cpp:140-160); after
  initialize the error is 0; after update_boundary at t = 0.25 the Dirichlet
  faces xmin and zmax hold sin(2 pi (|x - c| - r - t)) to 1e-10 relative;
  the rhs is finite.
* constant-velocity wave (the extension for the translated Gaussian pulse):
  faces and Dirichlet cut points hold A exp(-|x - x0 - a t|^2 / (2 w^2)).
* the same wave without a body converges at order about 4 (E4u interior
  rows, closures one order lower) over three grids.
"""
import numpy as np

import helpers as H
import oracle_ffi as O

D, F, N = O.DIRICHLET, O.FLOATING, O.NEUMANN
BL = (-0.7733323791884821, 0.1623961700641681)


def _coords(osys):
    m = osys.mesh
    xyz = [np.asarray(m.m.xyz[d][:m.m.n[d]]) for d in range(3)]
    X, Y, Z = np.meshgrid(*xyz, indexing="ij")
    return X.ravel(), Y.ravel(), Z.ravel()


def test_radial_update_boundary_reference_case():
    c = (2.0001, 2.5656565, 1.313131311)
    rad, t = 0.25, 0.25
    n = (21, 22, 23)
    cfg = dict(kind=O.SYS_WAVE_RADIAL, n=n, lo=(1, 1.1, 0.3), hi=(3, 3.3, 2.2), spheres=[list(c) + [rad]],
               obj_bc=[D], grid_bc=[D, F, N, N, F, D], scheme=O.E2_2, alpha=(0.0, 0.0),
               params=[c[0], c[1], c[2], rad])
    osys = H.oracle_system(cfg)
    u = osys.field()
    osys.initialize(u, 0.0)
    assert abs(osys.stats(u, 0.0)[0]) <= 1e-13
    osys.update_boundary(u, t)

    def sol(x, y, z):
        r = np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2)
        return np.sin(2 * np.pi * (r - rad - t))

    # the reference test's own coordinates: min + index * (max - min) / (n - 1)
    i, j, k = np.meshgrid(np.arange(n[0]), np.arange(n[1]), np.arange(n[2]), indexing="ij")
    x = 1.0 + i * (2.0 / 20)
    y = 1.1 + j * (2.2 / 21)
    z = 0.3 + k * (1.9 / 22)
    Dv = u.D.reshape(n)
    for face in (np.s_[0, :, :], np.s_[:, :, -1]):
        want = sol(x[face], y[face], z[face])
        assert np.allclose(Dv[face], want, rtol=1e-10, atol=1e-12)
    # Dirichlet cut points take the same solution
    for d in range(3):
        p = osys.mesh.r_arrays(d)["pos"]
        if len(p):
            assert np.allclose(u.R[d], sol(p[:, 0], p[:, 1], p[:, 2]), rtol=1e-12, atol=1e-13)
    rhs = osys.rhs(u, t)
    assert np.all(np.isfinite(rhs.D))


def test_gaussian_pulse_exact_and_boundary():
    a = 1 / np.sqrt(3.0)
    x0, w, amp, t = (0.25, 0.3, 0.2), 0.07, 1.3, 0.25
    cfg = H.sphere_wave_cfg(24, scheme=O.E4U_1, alpha=BL)
    cfg["params"] = [a, a, a, x0[0], x0[1], x0[2], w, amp]
    osys = H.oracle_system(cfg)

    def sol(x, y, z, tt):
        r2 = (x - x0[0] - a * tt) ** 2 + (y - x0[1] - a * tt) ** 2 + (z - x0[2] - a * tt) ** 2
        return amp * np.exp(-r2 / (2 * w * w))

    X, Y, Z = _coords(osys)
    ex = osys.exact(t)
    assert np.allclose(ex.D, sol(X, Y, Z, t), rtol=1e-13, atol=1e-300)
    u = osys.field()
    osys.initialize(u, 0.0)
    fl = osys.mesh.fluid()
    assert np.allclose(u.D[fl], sol(X[fl], Y[fl], Z[fl], 0.0), rtol=1e-13, atol=1e-300)
    osys.update_boundary(u, t)
    n = osys.mesh.m.n
    for d in range(3):  # inflow faces xmin, ymin, zmin are Dirichlet
        face = [slice(None)] * 3
        face[d] = 0
        idx = np.arange(osys.mesh.N).reshape(tuple(n))[tuple(face)].ravel()
        assert np.allclose(u.D[idx], sol(X[idx], Y[idx], Z[idx], t), rtol=1e-13, atol=1e-300)
        p = osys.mesh.r_arrays(d)["pos"]
        assert np.allclose(u.R[d], sol(p[:, 0], p[:, 1], p[:, 2], t), rtol=1e-13, atol=1e-300)


def _wave_error(n, T=0.2, cfl=0.5):
    a = 1 / np.sqrt(3.0)
    cfg = H.sphere_wave_cfg(n, scheme=O.E4U_1, alpha=BL)
    cfg.update(spheres=[], obj_bc=[], params=[a, a, a, 0.4, 0.4, 0.4, 0.1, 1.0])
    s = H.oracle_system(cfg)
    u = s.field()
    s.initialize(u, 0.0)
    s.update_boundary(u, 0.0)
    out, rk, srhs = s.field(), s.field(), s.field()
    nsteps = int(np.ceil(T / (cfl / (n - 1))))
    dt, t = T / nsteps, 0.0
    for _ in range(nsteps):
        s.rk4_step(u, out, rk, srhs, t, dt)
        t += dt
        u, out = out, u
    return s.stats(u, t)[0]


def test_linear_wave_converges_at_fourth_order():
    """Sphere-free constant-velocity wave (E4u with the reference's
    scalar-wave closure parameters, Dirichlet inflow faces, floating outflow
    faces), pulse width 0.1 from (0.4, 0.4, 0.4) to t = 0.2: measured orders
    3.7, 3.9, 3.9 on 25 -> 33 -> 49 -> 65 points per side."""
    errs = [_wave_error(n) for n in (33, 49, 65)]
    h = [1 / 32, 1 / 48, 1 / 64]
    orders = [np.log(errs[q] / errs[q + 1]) / np.log(h[q] / h[q + 1]) for q in range(2)]
    assert errs[-1] < 5e-4
    assert min(orders) > 3.6, orders
