"""The Euler oracle (oracle/euler.c), pinned without a reference
implementation: free-stream preservation, gas at rest around a slip-wall
sphere, the wall-state invariants, and convergence on the isentropic vortex
whose formulas the reference's stub carries (src/systems/inviscid_vortex.cpp:13-116)."""
import numpy as np
import pytest

import oracle_ffi as O

G = 1.4


E4U_ALPHA = (-0.7733323791884821, 0.1623961700641681)


def euler_cfg(n, lo, hi, spheres=(), grid_bc=None, scheme=O.E4U_1, alpha=E4U_ALPHA):
    """E4u by default.  E4_1's rows sum to zero, but at grid walls it clamps
    psi = 1 to 1 - 1e-4 (E4_1.cpp:100-101), where its closure coefficients
    reach about 4.5e5 / h: no explicit RK4 step of usable size is stable.
    Short segments between a body and a wall are refused with either scheme
    when their two closure blocks would overlap."""
    n = tuple(n) + (1,) * (3 - len(n))
    return dict(n=n, lo=lo, hi=hi, spheres=[list(s) for s in spheres],
                grid_bc=list(grid_bc) if grid_bc else [O.FLOATING] * 6, scheme=scheme, alpha=alpha)


def set_state(e, U, f):
    """U[k] = f(x, y, z)[k] at the mesh points and the cut points."""
    m = e.mesh
    X, Y, Z = np.meshgrid(*[np.asarray(m.m.xyz[d][:m.m.n[d]]) for d in range(3)], indexing="ij")
    vals = f(X.ravel(), Y.ravel(), Z.ravel())
    for k in range(5):
        U[k].D[:] = vals[k]
    for d in range(3):
        if m.nR(d):
            p = m.r_arrays(d)["pos"]
            vals = f(p[:, 0], p[:, 1], p[:, 2])
            for k in range(5):
                U[k].R[d][:] = vals[k]


def uniform(rho, u, p):
    def f(x, y, z):
        one = np.ones_like(x)
        m = [rho * u[a] * one for a in range(3)]
        E = p / (G - 1) + 0.5 * rho * (u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
        return [rho * one] + m + [E * one]

    return f


def vortex(t, eps=5.0, M0=0.5):
    """inviscid_vortex.cpp:13-116: the isentropic vortex centred at the
    origin, carried by a unit stream along x."""
    def f(x, y, z):
        xh, yh = x - t, y
        fr = 1.0 - (xh * xh + yh * yh)
        rho = (1.0 - 0.5 * (G - 1) * (eps * M0 / (2 * np.pi)) ** 2 * np.exp(fr)) ** (1.0 / (G - 1))
        ru = rho * (1.0 - eps * yh * np.exp(fr / 2) / (2 * np.pi))
        rv = rho * eps * xh / (2 * np.pi) * np.exp(0.5 * fr)
        p = rho ** G / (G * M0 * M0)
        E = p / (G - 1) + 0.5 * (ru * ru + rv * rv) / rho
        return [rho, ru, rv, 0 * rho, E]

    return f


def test_free_stream_is_preserved():
    """A uniform stream through a box without bodies: every rhs row of every
    variable vanishes (closure rows sum to zero)."""
    e = O.OracleEuler(euler_cfg((16, 18, 20), (0, 0, 0), (1, 1.1, 1.2), scheme=O.E4_1, alpha=(0.1, 0.7)))
    U = e.fields()
    set_state(e, U, uniform(1.0, (0.3, -0.2, 0.1), 1 / G))
    r = e.rhs(U)
    scale = 10.0 / min(e.mesh.h())  # fluxes of order 1, rows of order 1/h
    for k in range(5):
        assert np.max(np.abs(r[k].D)) < 1e-12 * scale * 100, k  # rounding through the closures


def test_gas_at_rest_around_a_slip_wall_sphere():
    """Uniform density and pressure at rest: the wall state equals the fluid
    state, every flux is constant, the rhs vanishes.  (p = 0.8 and E = p / 0.4
    = 2 convert exactly, so the wall's pressure is the fluid's to the bit;
    the sphere keeps every psi above 0.09, where the cut-cell closure weights,
    which grow like 1 / psi, stay moderate.)"""
    e = O.OracleEuler(euler_cfg((24, 24, 24), (0, 0, 0), (1, 1, 1), spheres=[(0.5, 0.5, 0.5, 0.26)]))
    U = e.fields()
    set_state(e, U, uniform(1.25, (0, 0, 0), 0.8))
    e.wall(U)
    r = e.rhs(U)
    scale = 10.0 / min(e.mesh.h())
    for k in range(5):
        assert np.max(np.abs(r[k].D)) < 1e-10 * scale, k


def test_wall_state_invariants():
    """After the wall update every cut point holds its fluid neighbour's
    density and pressure with zero normal momentum."""
    cfg = euler_cfg((20, 22, 24), (0, 0, 0), (1, 1, 1), spheres=[(0.5, 0.5, 0.5, 0.25)])
    e = O.OracleEuler(cfg)
    U = e.fields()
    rng = np.random.default_rng(1)
    N = e.mesh.N
    U[0].D[:] = 1.0 + 0.2 * rng.random(N)
    for k in (1, 2, 3):
        U[k].D[:] = rng.standard_normal(N) * 0.3
    U[4].D[:] = 2.5 + 0.5 * rng.random(N) + 0.5 * (U[1].D ** 2 + U[2].D ** 2 + U[3].D ** 2) / U[0].D
    e.wall(U)
    n = e.mesh.m.n
    strides = np.array([n[1] * n[2], n[2], 1])
    for d in range(3):
        r = e.mesh.r_arrays(d)
        nrm = np.array([e.mesh.m.R[d][i].normal[:3] for i in range(e.mesh.nR(d))])
        fl = r["solid"] @ strides + np.where(r["ray_outside"] != 0, -strides[d], strides[d])
        rho = U[0].R[d][: e.mesh.nR(d)]
        m = np.stack([U[k].R[d][: e.mesh.nR(d)] for k in (1, 2, 3)], axis=1)
        E = U[4].R[d][: e.mesh.nR(d)]
        np.testing.assert_array_equal(rho, U[0].D[fl])
        assert np.max(np.abs(np.sum(m * nrm, axis=1))) < 1e-15
        p_w = (G - 1) * (E - 0.5 * np.sum(m * m, axis=1) / rho)
        mf = np.stack([U[k].D[fl] for k in (1, 2, 3)], axis=1)
        p_f = (G - 1) * (U[4].D[fl] - 0.5 * np.sum(mf * mf, axis=1) / U[0].D[fl])
        np.testing.assert_allclose(p_w, p_f, rtol=1e-13)
        assert np.all(np.isin(fl, e.mesh.fluid()))  # the neighbour is a fluid point


def test_isentropic_vortex_converges():
    """rhs(U_vortex(t)) - dU/dt -> 0 under refinement on [-5, 5]^2 (E4u,
    floating faces, where the vortex has decayed to the free stream): at
    least third order between 41 and 81 points per side (measured 3.7).  A
    wrong flux term or sign leaves an error of the size of dU/dt."""
    scheme, alpha, order = O.E4U_1, E4U_ALPHA, 3.0
    errs = []
    t = 0.3
    for n in (41, 81):
        e = O.OracleEuler(euler_cfg((n, n), (-5, -5, 0), (5, 5, 0), scheme=scheme, alpha=alpha))
        U = e.fields()
        set_state(e, U, vortex(t))
        r = e.rhs(U)
        h = 1e-6
        Up, Um = e.fields(), e.fields()
        set_state(e, Up, vortex(t + h))
        set_state(e, Um, vortex(t - h))
        err = max(np.max(np.abs(r[k].D - (Up[k].D - Um[k].D) / (2 * h))) for k in range(5))
        errs.append(err)
    rate = np.log2(errs[0] / errs[1])
    assert rate >= order, (errs, rate)
