"""Adjoint: transposed derivative applies, the RK4 step adjoint, and misfit
gradients through checkpointed solves (no reference counterpart; pinned by
the explicit transpose of the oracle-assembled operator, dot-product tests
and central finite differences)."""
import numpy as np
import pytest

import helpers as H
import oracle_ffi as O

pytestmark = pytest.mark.gpu

# small grids (the explicit-transpose test assembles dense matrices), with
# spheres small enough that no segment is shorter than its two closures
CFGS = {
    "wave16_E4u": H.sphere_wave_cfg(16, scheme=O.E4U_1, alpha=(-0.7733323791884821, 0.1623961700641681),
                                    radius=0.2),
    "wave22_E4u_short": H.sphere_wave_cfg(22, scheme=O.E4U_1, alpha=(0.1, 0.2)),
    "wave18_radial": H.sphere_wave_cfg(18, linear=False, scheme=O.E4U_1, radius=0.2),
    "heat3d": H.heat3d_cfg(n=(14, 16, 18), radius=0.18),
    "heat3d_odd": H.heat3d_cfg(n=(13, 15, 17), radius=0.15),
    "heat2d": H.heat2d_cfg(21),
}


def _rand(prob, seed, scale=1.0):
    return scale * np.random.default_rng(seed).standard_normal(prob.flat_size)


def _parts(prob, flat):
    """(D, [R0, R1, R2]) views of a flat field."""
    N = prob.N
    out, o = [], N
    for d in range(3):
        out.append(flat[o:o + prob.nR[d]])
        o += prob.nR[d]
    return flat[:N], out


@pytest.mark.parametrize("name", list(CFGS))
def test_derivative_transpose_dot_product(ctx, name):
    """<D u, y> = <u, D^T y> for every direction, with u's D and cut-point
    parts, to 1e-12 relative."""
    cfg = CFGS[name]
    prob = H.product_problem(ctx, cfg)
    for d in range(3):
        if prob.n[d] < 2:
            continue
        u = _rand(prob, 10 + d)
        y = _rand(prob, 20 + d)
        gu, gdu, gy, gt = prob.field(), prob.field(), prob.field(), prob.field()
        gu.upload(u)
        gdu.upload(np.zeros(prob.flat_size))
        gy.upload(y)
        gt.upload(np.zeros(prob.flat_size))
        prob.derivative(d)(gu, gdu, 0)
        prob.derivative_T(d)(gy, gt, 0)
        du, ut = gdu.download(), gt.download()
        duD, _ = _parts(prob, du)
        yD, _ = _parts(prob, y)
        uD, uR = _parts(prob, u)
        utD, utR = _parts(prob, ut)
        lhs = duD @ yD
        rhs = uD @ utD + uR[d] @ utR[d]
        scale = np.linalg.norm(duD) * np.linalg.norm(yD) + np.linalg.norm(u) * np.linalg.norm(ut)
        assert abs(lhs - rhs) <= 1e-12 * scale, f"{name} dir {d}: {lhs} vs {rhs}"
        # accumulate form adds
        prob.derivative_T(d)(gy, gt, 1)
        assert np.allclose(gt.download(), 2 * ut, rtol=1e-15, atol=0)
        for f in (gu, gdu, gy, gt):
            f.free()
    prob.close()


@pytest.mark.parametrize("name", ["wave16_E4u", "heat3d_odd"])
def test_derivative_transpose_explicit(ctx, name):
    """D^T y against the explicit transpose of the operator assembled by the
    oracle (one oracle apply per unit column), 1e-13 relative."""
    cfg = CFGS[name]
    osys = H.oracle_system(cfg)
    prob = H.product_problem(ctx, cfg)
    N, nR = prob.N, prob.nR
    for d in range(3):
        if prob.n[d] < 2:
            continue
        ncol = N + nR[d]
        A = np.zeros((N, ncol))
        e = osys.field()
        du = osys.field()
        for c in range(ncol):
            e.D[:] = 0.0
            for r in range(3):
                e.R[r][:] = 0.0
            if c < N:
                e.D[c] = 1.0
            else:
                e.R[d][c - N] = 1.0
            du.D[:] = 0.0
            osys.deriv_apply(d, e, du)
            A[:, c] = du.D
        y = _rand(prob, 30 + d)
        yD, _ = _parts(prob, y)
        want = A.T @ yD
        gy, gt = prob.field(), prob.field()
        gy.upload(y)
        gt.upload(np.zeros(prob.flat_size))
        prob.derivative_T(d)(gy, gt, 0)
        tD, tR = _parts(prob, gt.download())
        got = np.concatenate([tD, tR[d]])
        scale = np.max(np.abs(A)) * np.max(np.abs(yD)) * 16
        assert np.max(np.abs(got - want)) <= 1e-13 * scale, f"{name} dir {d}"
        gy.free()
        gt.free()
    prob.close()


def _step_pair(prob, dt, t):
    import shoccs_amd as S

    prob.set_tuning(0)
    prob.set_precision(S.F64)

    def step(flat):
        f = prob.field()
        f.upload(flat)
        prob.rk4_fused(f, t, dt, 1)
        out = f.download()
        f.free()
        return out

    return step


@pytest.mark.parametrize("name", ["wave16_E4u", "wave18_radial", "heat3d", "heat2d"])
def test_rk4_adjoint_dot_product(ctx, name):
    """The step is affine in its input: with M d = step(u + d) - step(u),
    <M d, ybar> = <d, M^T ybar>, M^T from shx_rk4_adjoint (1e-10 relative)."""
    cfg = CFGS[name]
    osys = H.oracle_system(cfg)
    prob = H.product_problem(ctx, cfg)
    u0 = osys.field()
    osys.initialize(u0, 0.0)
    osys.update_boundary(u0, 0.0)
    u = u0.flat()
    dt = osys.dt(0.5, 0.5)
    t = 0.1
    step = _step_pair(prob, dt, t)
    rng = np.random.default_rng(7)
    delta = rng.standard_normal(u.size)
    ybar = rng.standard_normal(u.size)
    Md = step(u + delta) - step(u)
    gu, gb = prob.field(), prob.field()
    gu.upload(u)
    gb.upload(ybar)
    prob.rk4_adjoint(gu, gb, t, dt)
    MTy = gb.download()
    lhs, rhs = Md @ ybar, delta @ MTy
    scale = np.linalg.norm(Md) * np.linalg.norm(ybar) + np.linalg.norm(delta) * np.linalg.norm(MTy)
    assert abs(lhs - rhs) <= 1e-10 * scale, f"{name}: {lhs} vs {rhs}"
    prob.close()


def _solver_setup(ctx, cfg, nsteps, every, seed=3):
    from shoccs_amd.adjoint import Solve

    osys = H.oracle_system(cfg)
    prob = H.product_problem(ctx, cfg)
    dt = osys.dt(0.3, 0.5)
    u0 = osys.field()
    osys.initialize(u0, 0.0)
    osys.update_boundary(u0, 0.0)
    base = u0.flat()
    rng = np.random.default_rng(seed)
    sol = Solve(prob, 0.0, dt, nsteps, every)
    u_obs = sol.run(base + 0.1 * rng.standard_normal(base.size))
    return prob, sol, base, u_obs, rng


@pytest.mark.parametrize("name", ["wave16_E4u", "heat3d"])
def test_misfit_gradient_initial_field_fd(ctx, name):
    """dJ/du0 from the checkpointed adjoint against central differences along
    random directions (J is quadratic in u0, so the difference is exact up to
    rounding): 1e-6 relative."""
    prob, sol, u0, u_obs, rng = _solver_setup(ctx, CFGS[name], nsteps=23, every=5)
    J0 = sol.misfit(u0, u_obs)
    g, _ = sol.gradient(sol.run(u0) - u_obs, with_coef=False)
    for k in range(4):
        v = rng.standard_normal(u0.size)
        eps = 1e-3
        fd = (sol.misfit(u0 + eps * v, u_obs) - sol.misfit(u0 - eps * v, u_obs)) / (2 * eps)
        ad = g @ v
        assert abs(fd - ad) <= 1e-6 * max(abs(fd), abs(ad), 1e-300), f"{name} dir {k}: fd {fd} ad {ad} (J {J0})"
    sol.close()
    prob.close()


@pytest.mark.parametrize("name", ["wave16_E4u", "heat3d"])
def test_misfit_gradient_coefficients_fd(ctx, name):
    """dJ/dcoef (the stored interior coefficients of the three operators)
    against central differences along random directions: 1e-6 relative."""
    prob, sol, u0, u_obs, rng = _solver_setup(ctx, CFGS[name], nsteps=12, every=4)
    c0 = sol.coef()
    sol.run(u0)
    _, gc = sol.gradient(sol.run(u0) - u_obs)
    w = c0.shape[1]
    for k in range(3):
        v = rng.standard_normal(c0.shape) * np.abs(c0).max()
        eps = 1e-6
        sol.set_coef(c0 + eps * v)
        jp = sol.misfit(u0, u_obs)
        sol.set_coef(c0 - eps * v)
        jm = sol.misfit(u0, u_obs)
        sol.set_coef(c0)
        fd = (jp - jm) / (2 * eps)
        ad = float(np.sum(gc[:, :w] * v))
        assert abs(fd - ad) <= 1e-6 * max(abs(fd), abs(ad)), f"{name} dir {k}: fd {fd} ad {ad}"
    sol.close()
    prob.close()


def test_torch_function_matches_solver(ctx):
    torch = pytest.importorskip("torch")
    from shoccs_amd.adjoint import solve_fn

    prob, sol, u0, u_obs, _ = _solver_setup(ctx, CFGS["wave16_E4u"], nsteps=9, every=4)
    c0 = sol.coef()
    g_ref, gc_ref = sol.gradient(sol.run(u0) - u_obs)
    f = solve_fn(sol)
    tu = torch.tensor(u0, dtype=torch.float64, requires_grad=True)
    tc = torch.tensor(c0, dtype=torch.float64, requires_grad=True)
    J = 0.5 * ((f(tu, tc) - torch.tensor(u_obs)) ** 2).sum()
    J.backward()
    np.testing.assert_allclose(tu.grad.numpy(), g_ref, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(tc.grad.numpy(), gc_ref[:, :c0.shape[1]], rtol=1e-12, atol=1e-300)
    sol.close()
    prob.close()


@pytest.mark.parametrize("name", ["wave16_E4u", "heat3d"])
def test_kept_stage_adjoint_matches_recomputing_adjoint(ctx, name):
    """The solver's gradient (segment replays keeping the stage inputs, step
    adjoints over them, coefficient gradient accumulated on the device)
    against the step adjoint that recomputes its stages through the
    reference schedule (shx_rk4_adjoint), step by step from the same states:
    1e-11 relative in u0 and in the coefficients."""
    prob, sol, u0, u_obs, _ = _solver_setup(ctx, CFGS[name], nsteps=7, every=3)
    uT = sol.run(u0)
    g_fast, gc_fast = sol.gradient(uT - u_obs)
    # reference: states from single fused steps, then shx_rk4_adjoint in reverse
    states = [prob.field() for _ in range(7)]
    states[0].upload(u0)
    for k in range(1, 7):
        states[k].copy_from(states[k - 1])
        prob.rk4_fused(states[k], sol._time(k - 1), sol.dt, 1)
    ubar = prob.field()
    ubar.upload(uT - u_obs)
    gc = np.zeros(27)
    for k in range(6, -1, -1):
        prob.rk4_adjoint(states[k], ubar, sol._time(k), sol.dt, gc)
    g_ref = ubar.download()
    scale = np.max(np.abs(g_ref))
    assert np.max(np.abs(g_fast - g_ref)) <= 1e-11 * scale
    gc = gc.reshape(3, 9)
    assert np.max(np.abs(gc_fast - gc)) <= 1e-11 * np.max(np.abs(gc))
    for f in states + [ubar]:
        f.free()
    sol.close()
    prob.close()
