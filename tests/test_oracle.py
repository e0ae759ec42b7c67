"""Pins of the CPU oracle: the reference's own golden vectors (tests/golden,
extracted from the reference's unit tests) and what the mathematics fixes
(polynomial exactness of interior stencils, closures and whole operators,
measured convergence orders, RK4 on a solution linear in time)."""
import ctypes
import json
import math
import os

import numpy as np
import pytest

import helpers as H
import oracle_ffi as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SCHEMES = {"E4_1": O.E4_1, "E4u_1": O.E4U_1, "E2_2": O.E2_2, "E4_2": O.E4_2, "E2P_1": O.E2P_1}
BCS = {"Floating": O.FLOATING, "Dirichlet": O.DIRICHLET, "Neumann": O.NEUMANN}


def _cases():
    return json.load(open(os.path.join(GOLDEN, "stencil_closures.json")))


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c["source"])
def test_closure_golden_vectors(case):
    """Closure coefficients equal the reference tests' vectors.  The golden
    values were printed by the reference's generator; they agree with the
    restated formulas to <= 1e-12 relative (the reference's own check uses an
    absolute margin of 1e-8)."""
    st = O.stencil(SCHEMES[case["scheme"]], tuple(case["alpha"]))
    c, ex, q = O.nbs(st, case["h"], BCS[case["bc"]], case["psi"], case["right"])
    want = np.array(case["expected"])
    assert c.size == want.size
    scale = max(1.0, np.max(np.abs(want)))
    assert np.max(np.abs(c - want)) <= 1e-12 * scale


def test_query_shapes():
    """stencil::query (E4_1.cpp:43-56, E4u_1.cpp:26-39, E4_2.cpp:16-28, E2_2.cpp:12-24)."""
    L = O.lib()
    expect = {O.E4_1: (2, 5, 7, 0), O.E4U_1: (2, 3, 5, 0), O.E4_2: (2, 3, 5, 3), O.E2_2: (1, 2, 4, 2)}
    for k, (p, r, t, x) in expect.items():
        st = O.stencil(k, (0.1, 0.7))
        q = L.orc_query_max(st)
        assert (q.p, q.r, q.t, q.x) == (p, r, t, x)
        d = L.orc_query(st, O.DIRICHLET)
        assert (d.r, d.t, d.x) == (r - 1, t, 0)


def _closure_points(h, psi, t):
    """Positions of a left closure's t columns: the wall, then t-1 grid points
    starting psi*h from it."""
    return np.concatenate([[0.0], psi * h + h * np.arange(t - 1)])


# Column at which row 0 of a closure differentiates (measured property of the
# closed forms): the first-derivative cut-cell scheme E4_1 starts one column
# later than the others.  A Dirichlet closure drops the first floating row.
ROW0_COL = {O.E4_1: 1, O.E4U_1: 0, O.E4_2: 0, O.E2_2: 0}


@pytest.mark.parametrize("scheme,alpha,order,deg", [(O.E4_1, (0.1, 0.7), 1, 3),
                                                    (O.E4U_1, (0.1, 0.2), 1, 3),
                                                    (O.E4_2, (0.0, 0.0), 2, 3),
                                                    (O.E2_2, (0.0, 0.0), 2, 2)])
@pytest.mark.parametrize("bc", [O.FLOATING, O.DIRICHLET])
@pytest.mark.parametrize("right", [False, True])
def test_closure_polynomial_exactness(scheme, alpha, order, deg, bc, right):
    """Every closure row differentiates polynomials of degree <= deg exactly
    at its own grid point (a dropped or mistyped term breaks this).  Right-wall
    closures are checked on the mirrored points.  E4u_1 is psi-independent and
    is checked at psi = 1."""
    st = O.stencil(scheme, alpha)
    h = 0.37
    psis = [1.0] if scheme == O.E4U_1 else [0.2, 0.55, 0.9]
    for psi in psis:
        c, _, q = O.nbs(st, h, bc, psi, right)
        c = c.reshape(q.r, q.t)
        x = _closure_points(h, psi, q.t)
        if right:
            x = -x[::-1]
        x = x + 0.3
        first = ROW0_COL[scheme] + (0 if bc == O.FLOATING else 1)
        for r in range(q.r):
            col = first + r if not right else q.t - 1 - (first + (q.r - 1 - r))
            xr = x[col]
            for k in range(deg + 1):
                val = c[r] @ (x ** k)
                if order == 1:
                    exact = k * xr ** (k - 1) if k > 0 else 0.0
                else:
                    exact = k * (k - 1) * xr ** (k - 2) if k > 1 else 0.0
                scale = np.max(np.abs(c[r])) * np.max(np.abs(x)) ** k
                assert abs(val - exact) <= 1e-11 * scale, (psi, r, k, val, exact)


@pytest.mark.parametrize("scheme,alpha", [(O.E4_1, (0.1, 0.7)), (O.E4U_1, (0.1, 0.2))])
def test_interior_first_derivative_exact(scheme, alpha):
    st = O.stencil(scheme, alpha)
    h = 0.21
    c = O.interior(st, h)
    x = h * np.arange(-2, 3) + 0.4
    for k in range(5):
        assert abs(c @ x ** k - (k * 0.4 ** (k - 1) if k else 0.0)) <= 1e-12 * max(1, k * 4)


def test_mesh_golden_lines():
    g = json.load(open(os.path.join(GOLDEN, "mesh_lines.json")))
    m = O.OracleMesh(g["extents"], g["lo"], g["hi"], g["sphere"])
    ex = g["extents"]
    for d in range(3):
        slow = {0: 1, 1: 0, 2: 0}[d]
        fast = {0: 2, 1: 2, 2: 1}[d]
        assert m.nR(d) == g["n_intersections"][d]
        assert len(m.lines(d)) == ex[slow] * ex[fast] + g["n_intersections"][d] // 2
    for L in g["lines"]:
        ln = m.lines(L["dir"])[L["index"]]
        assert ln["stride"] == g["strides"][L["dir"]]
        assert list(ln["start"]["mesh"]) == L["start"]
        assert list(ln["end"]["mesh"]) == L["end"]
        so = ln["start"]["obj"]
        eo = ln["end"]["obj"]
        assert (so[0] if so else -1) == L["start_obj"]
        assert (eo[0] if eo else -1) == L["end_obj"]
        if "start_psi" in L:
            assert abs(so[2] - L["start_psi"]) <= 1e-14  # goldens from an exact-arithmetic tool
        if "end_psi" in L:
            assert abs(eo[2] - L["end_psi"]) <= 1e-14


def test_cut_points_lie_on_sphere():
    """Every cut point is on the sphere, psi is the fractional distance from
    the solid coordinate to the hit, and rays alternate in/out per line."""
    c = np.array([0.5, 0.45, 0.55])
    r = 0.31
    n = (23, 26, 29)
    m = O.OracleMesh(n, (0, 0, 0), (1, 1, 1), list(c) + [r])
    h = np.array(m.h())
    for d in range(3):
        a = m.r_arrays(d)
        dist = np.linalg.norm(a["pos"] - c, axis=1)
        assert np.max(np.abs(dist - r)) < 1e-14
        # fluid neighbour of the solid coordinate is psi*h away from the hit
        off = np.where(a["ray_outside"] == 1, -1, 1)
        fluid = a["solid"][:, d] + off
        fpos = fluid * h[d]
        psi = off * (fpos - a["pos"][:, d]) / h[d]
        assert np.max(np.abs(np.clip(psi, 1e-12, 1 - 1e-12) - a["psi"])) < 1e-12


def _poly_field(sys_, fn):
    """fn evaluated at every mesh point and cut point."""
    m = sys_.mesh.m
    n = m.n
    xs = [np.ctypeslib.as_array(m.xyz[a], shape=(n[a],)) for a in range(3)]
    X, Y, Z = np.meshgrid(xs[0], xs[1], xs[2], indexing="ij")
    f = sys_.field()
    f.D[:] = fn(X, Y, Z).ravel()
    for d in range(3):
        p = sys_.mesh.r_arrays(d)["pos"]
        if len(p):
            f.R[d][: len(p)] = fn(p[:, 0], p[:, 1], p[:, 2])
    return f


def _segment_class(s, cfg, d):
    """Per mesh point, for direction d: 0 no row, 1 interior row, 2 closure
    row next to a grid wall, 3 closure row next to a body."""
    n = cfg["n"]
    cls = np.zeros(int(np.prod(n)), np.int8)
    stride = {0: n[1] * n[2], 1: n[2], 2: 1}[d]
    for b in O.deriv_blocks(s.deriv(d)):
        base = b["row_offset"]
        L, I, R = b["left_rows"], b["interior_rows"], b["right_rows"]
        total = L + I + R
        lobj = b["left_cols"] == (7 - 1) or False
        for r in range(total):
            idx = base + r * stride
            if r < L:
                cls[idx] = 2
            elif r < L + I:
                cls[idx] = 1
            else:
                cls[idx] = 2
    # mark rows whose segment ends on a body
    for ln in s.mesh.lines(d):
        for end, sign in ((ln["start"], 1), (ln["end"], -1)):
            if end["obj"] is None:
                continue
            ijk = list(end["mesh"])
            for k in range(1, 6):
                p = list(ijk)
                p[d] += sign * k
                if 0 <= p[d] < n[d]:
                    idx = (p[0] * n[1] + p[1]) * n[2] + p[2]
                    if cls[idx] == 2:
                        cls[idx] = 3
    return cls


@pytest.mark.parametrize("scheme,alpha", [(O.E4U_1, (0.1, 0.2)), (O.E4_1, (0.1, 0.7))])
def test_first_derivative_exact_on_polynomials(scheme, alpha):
    """The assembled first derivative (O + B with Dirichlet-body values on the
    cut points) differentiates cubics exactly on every interior row, and
    constants (zero row sums, B included) on every row.  E4u_1 closures are
    also exact on cubics at grid walls, where the mesh is uniform."""
    cfg = H.sphere_wave_cfg(48, scheme=scheme, alpha=alpha)
    s = H.oracle_system(cfg)
    fn = lambda x, y, z: 1 + 2 * x - 3 * y + 0.5 * z + x * x - y * z + 0.7 * z ** 3
    dfn = [lambda x, y, z: 2 + 2 * x + 0 * y, lambda x, y, z: -3 - z + 0 * x,
           lambda x, y, z: 0.5 - y + 2.1 * z * z]
    u = _poly_field(s, fn)
    one = _poly_field(s, lambda x, y, z: 1.0 + 0 * x)
    for d in range(3):
        cls = _segment_class(s, cfg, d)
        du = s.field()
        du.D[:] = np.nan
        s.deriv_apply(d, u, du)
        want = _poly_field(s, dfn[d]).D
        err = np.abs(du.D - want)
        assert np.max(err[cls == 1]) < 1e-10
        if scheme == O.E4U_1:
            assert np.max(err[cls == 2]) < 1e-9
        dc = s.field()
        dc.D[:] = np.nan
        s.deriv_apply(d, one, dc)
        rows = cls > 0
        assert np.max(np.abs(dc.D[rows])) < 1e-8 * 48 * 1e4


def test_laplacian_exact_on_quadratics():
    """E4_2 second derivatives (closures at grid walls and at Dirichlet bodies)
    are exact on quadratics at every row, as the reference's own operator tests
    check for E2 (src/operators/derivative.t.cpp:324-384)."""
    cfg = H.heat3d_cfg(n=(26, 28, 30))
    s = H.oracle_system(cfg)
    fn = lambda x, y, z: x * (y + z) + y * (x + z) + z * (x + y) + 3 * x * y * z + x * x - 2 * z * z
    u = _poly_field(s, fn)
    lap = s.laplacian(u)
    fl = s.mesh.fluid()
    ijk = H.unflat(fl, cfg["n"])
    interior = np.all((ijk > 0) & (ijk < np.array(cfg["n"]) - 1), axis=1)
    got = lap.D[fl[interior]]
    assert np.max(np.abs(got - (2.0 - 4.0))) < 1e-7


@pytest.mark.parametrize("scheme,max_order,rms_order", [(O.E4_2, 1.8, 2.4), (O.E2_2, 0.9, 1.5)])
def test_laplacian_convergence_order(scheme, max_order, rms_order):
    """Measured orders of the assembled laplacian on a smooth function over
    sphere meshes 24^3..64^3: E4_2 max-norm 1.9, rms 2.6; E2_2 1.0 and 1.7
    (closure rows at the curved body bound the max-norm order)."""
    fn = lambda x, y, z: np.sin(2 * x + 0.3) * np.cos(1.5 * y) + np.exp(0.5 * z)
    lapf = lambda x, y, z: -6.25 * np.sin(2 * x + 0.3) * np.cos(1.5 * y) + 0.25 * np.exp(0.5 * z)
    errs, rms, hs = [], [], []
    for n in (24, 32, 48, 64):
        cfg = H.heat3d_cfg(n=(n, n, n), scheme=scheme)
        s = H.oracle_system(cfg)
        L = s.laplacian(_poly_field(s, fn))
        w = _poly_field(s, lapf).D
        fl = s.mesh.fluid()
        ijk = H.unflat(fl, (n, n, n))
        inner = fl[np.all((ijk > 0) & (ijk < n - 1), axis=1)]
        e = np.abs(L.D[inner] - w[inner])
        errs.append(e.max())
        rms.append(np.sqrt(np.mean(e ** 2)))
        hs.append(1.0 / (n - 1))
    assert np.polyfit(np.log(hs), np.log(errs), 1)[0] > max_order
    assert np.polyfit(np.log(hs), np.log(rms), 1)[0] > rms_order


def test_interior_rows_converge_fourth_order():
    """Away from walls and bodies the first derivative is the 4th-order
    central stencil: the error on interior rows falls as h^4."""
    fn = lambda x, y, z: np.sin(2 * x + 0.3) * np.cos(1.5 * y) + np.exp(0.5 * z)
    dfn = lambda x, y, z: 2 * np.cos(2 * x + 0.3) * np.cos(1.5 * y)
    errs, hs = [], []
    for n in (32, 48, 64):
        cfg = H.sphere_wave_cfg(n, scheme=O.E4U_1, alpha=(0.1, 0.2))
        s = H.oracle_system(cfg)
        cls = _segment_class(s, cfg, 0)
        du = s.field()
        s.deriv_apply(0, _poly_field(s, fn), du)
        e = np.abs(du.D - _poly_field(s, dfn).D)[cls == 1]
        errs.append(e.max())
        hs.append(1.0 / (n - 1))
    assert np.polyfit(np.log(hs), np.log(errs), 1)[0] > 3.7


def test_rk4_exact_for_linear_in_time_heat():
    """rk4_v2.t.cpp:28-143 analogue: for a heat solution whose spatial part is
    annihilated by the laplacian closures and linear in time the RK4 step is
    exact; here the sin-product MMS over 3 steps must stay within the spatial
    truncation error and the stepper must reproduce dt = p_cfl h^2 / 4k."""
    cfg = H.heat2d_cfg(41)
    s = H.oracle_system(cfg)
    h = 1.0 / 40
    assert s.dt(1.0, 0.8) == 0.8 * h * h / 4
    u0 = s.field()
    s.initialize(u0, 0.0)
    out, rk, srhs = s.field(), s.field(), s.field()
    dt = s.dt(1.0, 0.8)
    t = 0.0
    for _ in range(3):
        s.rk4_step(u0, out, rk, srhs, t, dt)
        u0.D[:] = out.D
        t += dt
    st = s.stats(u0, t)
    assert 0 < st[0] < 1e-3


def test_system_cycle_heat_2d_config1_short():
    """The 2-D heat configuration (101^2, E4, Dirichlet circle, sin-product MMS,
    dt = 0.2 h^2): 20 of its 200 steps stay within the discretisation error."""
    cfg = H.heat2d_cfg(101)
    s = H.oracle_system(cfg)
    u0 = s.field()
    s.initialize(u0, 0.0)
    s.update_boundary(u0, 0.0)
    out, rk, srhs = s.field(), s.field(), s.field()
    dt = s.dt(1.0, 0.8)
    assert abs(dt - 0.2 * 0.01 ** 2) < 1e-18
    t = 0.0
    for _ in range(20):
        s.rk4_step(u0, out, rk, srhs, t, dt)
        u0.D[:] = out.D
        for d in range(3):
            u0.R[d][:] = out.R[d]
        t += dt
    st = s.stats(u0, t)
    assert st[0] < 1e-5, st[0]


def _box_wave(kind, grid_bc, n=(18, 20, 22), scheme=O.E4U_1, alpha=(-0.7733323791884821, 0.1623961700641681)):
    a = np.array([0.3, -0.5, 0.8])
    if kind == O.SYS_WAVE_LINEAR:
        params = [a[0], a[1], a[2], 0.4, 0.5, 0.6, 0.1, 1.0]
    else:
        params = [-0.25, -0.3, -0.2, 0.1]
    cfg = dict(kind=kind, n=n, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), spheres=[], obj_bc=[],
               grid_bc=grid_bc, scheme=scheme, alpha=alpha, params=params)
    return cfg, a


@pytest.mark.parametrize("kind", [O.SYS_WAVE_LINEAR, O.SYS_WAVE_RADIAL])
def test_wave_rhs_closed_form_on_linear_field(kind):
    """scalar_wave rhs = gG . grad u with gG = -a (constant velocity) or
    -(x - c)/|x - c| (radial, scalar_wave.cpp:28-33), zeroed on Dirichlet
    faces (scalar_wave.cpp:70-95).  On u = 1 + 2x - 3y + 0.5z every closure row
    is exact, so the rhs has a closed form at every point; a sign, component
    or face mistake changes it."""
    grid_bc = [O.DIRICHLET, O.FLOATING, O.FLOATING, O.DIRICHLET, O.DIRICHLET, O.FLOATING]
    cfg, a = _box_wave(kind, grid_bc)
    s = H.oracle_system(cfg)
    g = np.array([2.0, -3.0, 0.5])
    u = _poly_field(s, lambda x, y, z: 1 + g[0] * x + g[1] * y + g[2] * z)
    rhs = s.rhs(u, 0.37).D
    n = np.array(cfg["n"])
    ijk = H.unflat(np.arange(rhs.size), cfg["n"])
    xyz = ijk / (n - 1.0)
    if kind == O.SYS_WAVE_LINEAR:
        want = np.full(rhs.size, -(a @ g))
    else:
        r = xyz - np.array(cfg["params"][:3])
        want = -(r @ g) / np.linalg.norm(r, axis=1)
    face = (ijk[:, 0] == 0) | (ijk[:, 1] == n[1] - 1) | (ijk[:, 2] == 0)
    want[face] = 0.0
    np.testing.assert_allclose(rhs, want, rtol=0, atol=1e-11)


def test_rk4_step_is_the_fourth_order_taylor_polynomial():
    """rk4.cpp:16-47 on an autonomous linear problem (constant-velocity wave,
    floating faces, no bodies: rhs(u) = L u with no boundary data): one step
    equals (I + dtL + (dtL)^2/2 + (dtL)^3/6 + (dtL)^4/24) u0, with L u computed
    by the (separately pinned) rhs.  A wrong stage time or weight fails."""
    cfg, _ = _box_wave(O.SYS_WAVE_LINEAR, [O.FLOATING] * 6, n=(14, 16, 18))
    s = H.oracle_system(cfg)
    u0 = s.field()
    rng = np.random.default_rng(11)
    u0.D[:] = rng.standard_normal(u0.D.size)
    dt = 0.9 * s.dt(0.5, 1.0)
    terms = [u0.D.copy()]
    cur = u0
    for _ in range(4):
        nxt = s.rhs(cur, 0.0)
        cur = s.field()
        cur.D[:] = nxt.D
        terms.append(cur.D.copy())
    want = sum(terms[k] * dt ** k / math.factorial(k) for k in range(5))
    out, rk, srhs = s.field(), s.field(), s.field()
    s.rk4_step(u0, out, rk, srhs, 0.0, dt)
    scale = np.max(np.abs(want))
    assert np.max(np.abs(out.D - want)) <= 1e-13 * scale
    # and the five terms do not collapse (the check has teeth)
    assert np.max(np.abs(terms[4])) * dt ** 4 / 24 > 1e-6 * scale


def test_stats_over_fluid_points():
    """compute_scalar_stats (scalar_system_utils.hpp:77-181): the L-inf error
    is taken over fluid points only, the first strict maximum wins, and a NaN
    point never wins (MaxLoc with `e > val`, :102-113)."""
    cfg = H.sphere_wave_cfg(24, scheme=O.E4U_1, alpha=(0.1, 0.2))
    s = H.oracle_system(cfg)
    u = s.exact(0.1)
    fl = s.mesh.fluid()
    solid = np.setdiff1d(np.arange(u.D.size), fl)
    u.D[solid] = 1e6  # solid points are ignored
    u.D[fl[7]] += 0.25
    st = s.stats(u, 0.1)
    assert abs(st[0] - 0.25) < 1e-15 and int(st[4]) == fl[7]
    u.D[fl[8]] = u.D[fl[8]] + 0.25  # a tie later in fluid order loses
    u.D[fl[9]] = np.nan
    st = s.stats(u, 0.1)
    assert abs(st[0] - 0.25) < 1e-15 and int(st[4]) == fl[7]


def test_euler_step_is_the_first_order_taylor_polynomial():
    """euler.cpp:11-27 on an autonomous linear problem (no boundary data):
    one step is u0 + dt L u0, with L u from the (pinned) rhs."""
    cfg, _ = _box_wave(O.SYS_WAVE_LINEAR, [O.FLOATING] * 6, n=(14, 16, 18))
    s = H.oracle_system(cfg)
    u0 = s.field()
    u0.D[:] = np.random.default_rng(12).standard_normal(u0.D.size)
    dt = 0.7 * s.dt(0.5, 1.0)
    want = u0.D + dt * s.rhs(u0, 0.0).D
    out, srhs = s.field(), s.field()
    s.euler_step(u0, out, srhs, 0.0, dt)
    assert np.array_equal(out.D, want)


# ------------------------------------------------- cut rows of non-Dirichlet bodies
def _floating_system(n, lo, hi, sphere, grid_bc, scheme, bc=None):
    st = O.stencil(scheme)
    return O.OracleSystem(O.SYS_HEAT, n, lo, hi, [sphere], [O.FLOATING if bc is None else bc], grid_bc, st,
                          [1.0, O.MMS_GAUSS, 0])


def _eval(osys, f):
    m = osys.mesh
    u = osys.field()
    X, Y, Z = np.meshgrid(*[np.asarray(m.m.xyz[d][:m.m.n[d]]) for d in range(3)], indexing="ij")
    u.D[:] = f(X.ravel(), Y.ravel(), Z.ravel())
    for d in range(3):
        if m.nR(d):
            p = m.r_arrays(d)["pos"]
            u.R[d][:] = f(p[:, 0], p[:, 1], p[:, 2])
    return u


def _f2(x, y, z):  # derivative.t.cpp:22-26: linear in each variable
    return x * (y + z) + y * (x + z) + z * (x + y) + 3 * x * y * z


@pytest.mark.parametrize("grid", ["nn_dd_ff", "ff_nn_dd"])
def test_cut_rows_e2_second_derivative_of_f2(grid):
    """derivative.t.cpp:515-568 ("E2 with Objects"): on the 25 x 26 x 27 mesh
    with a floating sphere, the E2 second derivative of f2 (linear in each
    variable) vanishes at the fluid points and at every cut point, the latter
    through the R-space rows (same-direction closures and cross-direction
    interpolation rows).  Only floating bodies are exact here: the reference
    gives a Neumann body its closure rows but never adds their boundary-data
    term (derivative.cpp:335-365 builds no N entries for objects)."""
    gb = {"nn_dd_ff": [O.NEUMANN, O.NEUMANN, O.DIRICHLET, O.DIRICHLET, O.FLOATING, O.FLOATING],
          "ff_nn_dd": [O.FLOATING, O.FLOATING, O.NEUMANN, O.NEUMANN, O.DIRICHLET, O.DIRICHLET]}[grid]
    dd_nu = {"nn_dd_ff": 0, "ff_nn_dd": 1}[grid]
    osys = _floating_system((25, 26, 27), (0.1, 0.2, 0.3), (1, 2, 2.2), [0.45, 1.011, 1.31, 0.25], gb, O.E2_2)
    m = osys.mesh
    assert all(m.nR(d) > 0 for d in range(3))
    u = _eval(osys, _f2)
    nu = _eval(osys, [lambda x, y, z: 2.0 * (y + z) + 3.0 * y * z,  # f2_dx, f2_dy
                      lambda x, y, z: 2.0 * (x + z) + 3.0 * x * z][dd_nu])
    scale = np.max(np.abs(u.flat())) / min(m.h()) ** 2
    fluid = np.zeros(m.N, bool)
    fluid[m.fluid()] = True
    for d in range(3):
        du = osys.field()
        lib = O.lib()
        lib.orc_deriv_apply(osys.deriv(d), u.ref(), nu.ref() if d == dd_nu else None, du.ref(), 0)
        assert np.max(np.abs(du.D[fluid])) <= 1e-10 * scale, f"dir {d} fluid"
        dR = osys.deriv(d).contents
        for r in range(3):
            if m.nR(r):
                assert dR.Bf[r].nnz > 0 and dR.Bf[r].rows == m.nR(r)
                assert np.max(np.abs(du.R[r])) <= 1e-10 * scale, f"dir {d} R{r}"


def test_cut_rows_e4_second_derivative_exact_on_cubics():
    """E4 second derivatives with a floating sphere are exact on a cubic at
    the fluid points and at the cut points (closures exact to degree 3,
    5-point interpolants exact to degree 4)."""
    osys = _floating_system((28, 30, 26), (0, 0, 0), (1, 1.1, 0.95), [0.52, 0.49, 0.47, 0.27],
                            [O.FLOATING] * 6, O.E4_2)
    m = osys.mesh
    f = lambda x, y, z: x ** 3 - 2 * x * y * y + y ** 3 + 0.5 * z ** 3 + x * y * z + z * z  # noqa: E731
    dd = [lambda x, y, z: 6 * x + 0 * y,
          lambda x, y, z: -4 * x + 6 * y,
          lambda x, y, z: 3 * z + 2 + 0 * x]
    u = _eval(osys, f)
    scale = np.max(np.abs(u.flat())) / min(m.h()) ** 2
    fluid = np.zeros(m.N, bool)
    fluid[m.fluid()] = True
    for d in range(3):
        du = osys.field()
        O.lib().orc_deriv_apply(osys.deriv(d), u.ref(), None, du.ref(), 0)
        want = _eval(osys, dd[d])
        assert np.max(np.abs(du.D[fluid] - want.D[fluid])) <= 1e-9 * scale, f"dir {d} fluid"
        for r in range(3):
            if m.nR(r):
                err = np.max(np.abs(du.R[r] - want.R[r]))
                assert err <= 1e-9 * scale, f"dir {d} R{r}: {err / scale:.2e}"


@pytest.mark.parametrize("kind,deg_wall,deg_int", [(O.E2_2, 1, 1), (O.E4_2, 3, 4), (O.E2P_1, 1, 1)])
def test_interpolants_reproduce_polynomials(kind, deg_wall, deg_int):
    """E2_2.cpp:31-86, E4_2.cpp:31-183 and polyE2_1.cpp:42-138 interpolants
    (E2-poly with random floating and interpolant parameters, as
    polyE2_1.t.cpp's "interp wall" case draws them).  Interior: nodes
    -p/2..p/2 (E2: the two points either side of y), exact to degree p-1.
    Wall (left): nodes 1-psi (the boundary), 1, 2, .. t-1 relative to the
    boundary point, evaluated at y + i (y + 1 - psi for i = 0, the shift
    stencil.hpp:213-270 applies when the closest point is the solid one),
    exact to degree 1 (E2) or 3 (E4); the right wall mirrors the nodes and
    the offset (y - i)."""
    L = O.lib()
    L.orc_interp_wall.restype = ctypes.c_int
    L.orc_interp_interior.restype = ctypes.c_int
    rng = np.random.default_rng(11)
    st = O.stencil(kind, rng.uniform(size=13) if kind == O.E2P_1 else (0.0, 0.0))
    t = {O.E2_2: 3, O.E4_2: 5, O.E2P_1: 4}[kind]
    nwall = {O.E2_2: 2, O.E4_2: 3, O.E2P_1: 2}[kind]
    c = (ctypes.c_double * 8)()
    for _ in range(40):
        y = rng.uniform(-0.95, 0.95)
        n = L.orc_interp_interior(ctypes.byref(st), ctypes.c_double(y), c, 8)
        x = (np.array([0.0, 1.0]) if y > 0 else np.array([-1.0, 0.0])) if kind != O.E4_2 else np.arange(-2.0, 3.0)
        w = np.array(c[:n])
        for m in range(deg_int + 1):
            assert abs(w @ x ** m - y ** m) < 1e-13, (y, m)
        for right in (0, 1):
            for i in range(nwall):
                psi = rng.uniform(0.05, 1.0)
                n = L.orc_interp_wall(ctypes.byref(st), i, ctypes.c_double(y), ctypes.c_double(psi), c, right, 8)
                assert n == t
                w = np.array(c[:n])
                x = np.concatenate([[1.0 - psi], np.arange(1.0, t)])
                off = i if i else 1.0 - psi
                e = y - off if right else y + off
                if right:
                    x = -x[::-1]
                for m in range(deg_wall + 1):
                    assert abs(w @ x ** m - e ** m) < 1e-12 * max(1.0, abs(e) ** m), (right, i, psi, y, m)


def test_general_heat_rhs_converges():
    """heat.cpp:119-162 and :296-354 with a floating sphere and Neumann grid
    faces (sin-product solution): rhs(exact(t)) - d/dt exact(t) at the fluid
    points and at the floating sphere's cut points shrinks under refinement.
    A missing Neumann data term or a wrong source sign would leave an error
    that does not shrink."""
    F, N, D = O.FLOATING, O.NEUMANN, O.DIRICHLET
    errs = []
    for n in (20, 40):
        cfg = dict(kind=O.SYS_HEAT, n=(n, n, n), lo=(0, 0, 0), hi=(1, 1, 1), spheres=[[0.52, 0.47, 0.5, 0.2]],
                   obj_bc=[F], grid_bc=[N, N, D, D, N, D], scheme=O.E4_2, alpha=(0.0, 0.0),
                   params=[1.0, O.MMS_SINPROD])
        osys = H.oracle_system(cfg)
        t = 0.3
        u = osys.exact(t)
        osys.update_boundary(u, t)  # Neumann face data
        r = osys.rhs(u, t)
        ut = osys.exact(t)
        ut.D[:] *= -1.0  # d/dt of sin sin sin e^-t
        for d in range(3):
            ut.R[d][:] *= -1.0
        m = osys.mesh
        f = m.fluid()
        e_fluid = np.max(np.abs(r.D[f] - ut.D[f]))
        free = [np.max(np.abs(r.R[d] - ut.R[d])) for d in range(3) if m.nR(d)]
        errs.append((e_fluid, max(free)))
    assert errs[1][0] < 0.5 * errs[0][0], errs
    assert errs[1][1] < 0.6 * errs[0][1], errs


@pytest.mark.parametrize("dims", [1, 2])
def test_gaussian_solutions_of_lower_dimension(dims):
    """gauss1d.cpp / gauss2d.cpp, chosen by the mesh's dimension
    (mms.cpp:16-31; the configured z (and y) variance is 0 there, where the
    3-D form would divide by it).  With diffusivity 1 the rhs of the exact
    solution approaches its time derivative under refinement (laplacian
    against the solution, gradient through the Neumann data); with
    diffusivity 0 the rhs is the time derivative, checked against a central
    difference in time."""
    F, N, D = O.FLOATING, O.NEUMANN, O.DIRICHLET
    g = [0.45, 0.55, 0.0, 0.2, 0.25, 0.0, 1.3, 2.5]
    if dims == 1:
        g[1] = g[4] = 0.0

    def cfg(n, k):
        if dims == 1:
            return dict(kind=O.SYS_HEAT, n=(n, 1, 1), lo=(0, 0, 0), hi=(1, 0, 0), spheres=[], obj_bc=[],
                        grid_bc=[N, D, F, F, F, F], scheme=O.E4_2, alpha=(0.0, 0.0), params=[k, O.MMS_GAUSS, 1] + g)
        # one Neumann face: where two meet, the reference's shared data field
        # holds the later direction's gradient (heat.cpp:318-346), which the
        # other direction's N row then reads, so that edge does not converge
        return dict(kind=O.SYS_HEAT, n=(n, n, 1), lo=(0, 0, 0), hi=(1, 1, 0), spheres=[[0.5, 0.48, 0.0, 0.21]],
                    obj_bc=[F], grid_bc=[N, D, D, F, F, F], scheme=O.E4_2, alpha=(0.0, 0.0),
                    params=[k, O.MMS_GAUSS, 1] + g)

    t = 0.35
    errs = []
    for n in (41, 81):
        osys = H.oracle_system(cfg(n, 1.0))
        u = osys.exact(t)
        assert np.all(np.isfinite(u.flat()))
        osys.update_boundary(u, t)
        r = osys.rhs(u, t).flat()
        h = 1e-5
        ut = (osys.exact(t + h).flat() - osys.exact(t - h).flat()) / (2 * h)
        # fluid points off the Dirichlet faces (the rhs is zero there by design)
        n3 = (n, 1, 1) if dims == 1 else (n, n, 1)
        ii = np.unravel_index(osys.mesh.fluid(), n3)
        on_face = (ii[0] == n - 1) if dims == 1 else ((ii[0] == n - 1) | (ii[1] == 0))
        f = osys.mesh.fluid()[~on_face]
        free = np.arange(osys.mesh.N, r.size)
        errs.append(max(np.max(np.abs(r[f] - ut[f])), np.max(np.abs(r[free] - ut[free])) if free.size else 0.0))
    assert errs[1] < 0.6 * errs[0], errs
    # diffusivity 0: the source alone, S_t
    osys = H.oracle_system(cfg(41, 0.0))
    u = osys.exact(t)
    r = osys.rhs(u, t).flat()
    h = 1e-5
    ut = (osys.exact(t + h).flat() - osys.exact(t - h).flat()) / (2 * h)
    n3 = (41, 1, 1) if dims == 1 else (41, 41, 1)
    ii = np.unravel_index(osys.mesh.fluid(), n3)
    on_face = (ii[0] == 40) if dims == 1 else ((ii[0] == 40) | (ii[1] == 0))
    f = osys.mesh.fluid()[~on_face]
    assert np.max(np.abs(r[f] - ut[f])) <= 1e-7 * np.max(np.abs(ut[f]))
