"""Derivative, gradient and laplacian with floating and Neumann bodies: the
cut-point rows (du.R) and the D rows beside them, against the oracle, bit
for bit (the reference's derivative.cpp:475-519 apply order)."""

import numpy as np
import pytest

import helpers as H
import oracle_ffi as O

pytestmark = pytest.mark.gpu

CUT = H.cut_cfgs()


def _rand_field(osys, seed):
    rng = np.random.default_rng(seed)
    f = osys.field()
    f.D[:] = rng.standard_normal(f.D.size)
    for d in range(3):
        f.R[d][:] = rng.standard_normal(f.R[d].size)
    return f


def _upload(prob, f):
    g = prob.field()
    g.upload(f.flat())
    return g


@pytest.mark.parametrize("name", list(CUT))
@pytest.mark.parametrize("plus", [False, True])
def test_derivative_with_cut_rows_bitexact(ctx, name, plus):
    """du.R accumulates the cut-point rows whatever the op; du.D follows op."""
    cfg = CUT[name]
    osys, prob = H.oracle_system(cfg), H.product_problem(ctx, cfg)
    u = _rand_field(osys, 1)
    gu = _upload(prob, u)
    for d in range(3):
        if prob.n[d] < 2:
            continue
        du = _rand_field(osys, 2 + d)
        gdu = _upload(prob, du)
        osys.deriv_apply(d, u, du, plus_eq=plus)
        prob.derivative(d)(gu, gdu, 1 if plus else 0)
        got, want = gdu.download(), du.flat()
        assert np.array_equal(got, want), f"{name} dir {d}: max diff {np.max(np.abs(got - want))}"
        gdu.free()
    prob.close()


@pytest.mark.parametrize("name", ["objects_E2", "two_spheres_E2"])
def test_derivative_with_neumann_data_bitexact(ctx, name):
    """The (u, nu, du) form on grids with Neumann faces: N nu.D added after
    the cut-point rows and the D rows."""
    cfg = CUT[name]
    osys, prob = H.oracle_system(cfg), H.product_problem(ctx, cfg)
    u, nu = _rand_field(osys, 5), _rand_field(osys, 6)
    gu, gnu = _upload(prob, u), _upload(prob, nu)
    for d in range(3):
        du = _rand_field(osys, 7 + d)
        gdu = _upload(prob, du)
        O.lib().orc_deriv_apply(osys.deriv(d), u.ref(), nu.ref(), du.ref(), 0)
        prob.derivative(d)(gu, gdu, 0, nu=gnu)
        assert np.array_equal(gdu.download(), du.flat()), f"{name} dir {d}"
        gdu.free()
    prob.close()


@pytest.mark.parametrize("name", list(CUT))
def test_gradient_and_laplacian_with_cut_rows_bitexact(ctx, name):
    cfg = CUT[name]
    osys, prob = H.oracle_system(cfg), H.product_problem(ctx, cfg)
    u = _rand_field(osys, 11)
    gu = _upload(prob, u)
    want = osys.gradient(u)
    outs = [_upload(prob, _rand_field(osys, 12 + c)) for c in range(3)]  # zeroed first
    prob.gradient(gu)(*outs)
    for c in range(3):
        assert np.array_equal(outs[c].download(), want[c].flat()), f"{name} gradient {c}"
    lap = osys.laplacian(u)
    gl = _upload(prob, _rand_field(osys, 15))
    prob.laplacian(gu)(gl)
    assert np.array_equal(gl.download(), lap.flat()), f"{name} laplacian"
    prob.close()


def test_e2_second_derivative_of_f2_vanishes_on_device(ctx):
    """derivative.t.cpp:515-568 on the device: with the floating sphere the
    E2 second derivative of f2 (linear in each variable) is zero at the fluid
    points and at every cut point."""
    cfg = CUT["objects_E2"]
    osys, prob = H.oracle_system(cfg), H.product_problem(ctx, cfg)
    m = osys.mesh
    f2 = lambda x, y, z: x * (y + z) + y * (x + z) + z * (x + y) + 3 * x * y * z  # noqa: E731
    f2x = lambda x, y, z: 2.0 * (y + z) + 3.0 * y * z  # noqa: E731
    X, Y, Z = np.meshgrid(*[np.asarray(m.m.xyz[d][:m.m.n[d]]) for d in range(3)], indexing="ij")

    def ev(f):
        fld = osys.field()
        fld.D[:] = f(X.ravel(), Y.ravel(), Z.ravel())
        for d in range(3):
            p = m.r_arrays(d)["pos"]
            fld.R[d][:] = f(p[:, 0], p[:, 1], p[:, 2])
        return fld

    u, nu = ev(f2), ev(f2x)
    gu, gnu = _upload(prob, u), _upload(prob, nu)
    scale = np.max(np.abs(u.flat())) / min(m.h()) ** 2
    fluid = prob.fluid()
    for d in range(3):
        gdu = _upload(prob, osys.field())
        prob.derivative(d)(gu, gdu, 0, nu=gnu if d == 0 else None)
        out = gdu.download()
        D, rest = out[:prob.N], out[prob.N:]
        assert np.max(np.abs(D[fluid])) <= 1e-10 * scale, d
        assert np.max(np.abs(rest)) <= 1e-10 * scale, d
        gdu.free()
    prob.close()


def test_transpose_refuses_cut_rows(ctx):
    """The adjoint's transposed apply does not cover cut-point rows."""
    import shoccs_amd as S

    prob = H.product_problem(ctx, CUT["floating_E4_2"])
    f, g = prob.field(), prob.field()
    with pytest.raises(S.ShoccsError):
        prob.derivative_T(0)(f, g, 0)
    prob.close()
