"""Spectral diagnostics (shoccs_amd.spectral): the dense operator of one
direction over its unknowns, as the reference's eigenvalue_visitor forms it
(src/operators/eigenvalue_visitor.cpp:10-52), and the hyperbolic eigenvalue
statistic (src/systems/hyperbolic_eigenvalues.cpp:78-89)."""
import numpy as np
import pytest

import helpers as H
import oracle_ffi as O

pytestmark = pytest.mark.gpu

F, D = O.FLOATING, O.DIRICHLET
E4U_ALPHA = (-0.7733323791884821, 0.1623961700641681)


def _line(n, scheme, grid, alpha=(0.0, 0.0), kind=O.SYS_HEAT, params=(1.0, O.MMS_SINPROD)):
    return dict(kind=kind, n=(n, 1, 1), lo=(0, 0, 0), hi=(1, 0, 0), spheres=[], obj_bc=[],
                grid_bc=list(grid) + [F] * 4, scheme=scheme, alpha=alpha, params=list(params))


def _positions(prob, d):
    import shoccs_amd.spectral as SP

    Dk, Rk = SP.unknowns(prob, d)
    return np.concatenate([Dk, prob.N + sum(prob.nR[:d]) + Rk])


def test_e2_dirichlet_eigenvalues_closed_form(ctx):
    """E2 second derivative with Dirichlet ends on n points: the tridiagonal
    (1, -2, 1)/h^2 on the n - 2 interior points, eigenvalues
    -(4/h^2) sin^2(k pi / (2 (n - 1))), k = 1 .. n - 2."""
    import shoccs_amd.spectral as SP

    n = 21
    prob = H.product_problem(ctx, _line(n, O.E2_2, [D, D]))
    lam = SP.eigenvalues(prob, 0)
    assert lam.size == n - 2 and np.max(np.abs(lam.imag)) == 0.0
    h = prob.h()[0]
    want = -(4 / h ** 2) * np.sin(np.arange(1, n - 1) * np.pi / (2 * (n - 1))) ** 2
    np.testing.assert_allclose(np.sort(lam.real), np.sort(want), rtol=1e-12)
    prob.close()


@pytest.mark.parametrize("case", ["line_E4u", "circle_E4_2_x", "circle_E4_2_y"])
def test_operator_matrix_matches_oracle(ctx, case):
    """The device-assembled matrix over the unknowns (mesh rows, then the
    free cut points of R(dir)) equals the oracle's, bit for bit."""
    import shoccs_amd.spectral as SP

    if case == "line_E4u":
        cfg, d = _line(31, O.E4U_1, [D, F], alpha=E4U_ALPHA, kind=O.SYS_WAVE_RADIAL, params=(-0.25, 0, 0, 0)), 0
    else:
        cfg = dict(kind=O.SYS_HEAT, n=(19, 17, 1), lo=(0, 0, 0), hi=(1, 1, 0), spheres=[[0.5, 0.52, 0.0, 0.22]],
                   obj_bc=[F], grid_bc=[D, F, F, D, F, F], scheme=O.E4_2, alpha=(0.0, 0.0),
                   params=[1.0, O.MMS_SINPROD])
        d = 0 if case.endswith("x") else 1
    prob = H.product_problem(ctx, cfg)
    pos = _positions(prob, d)
    A = SP.operator_matrix(prob, d)
    assert A.shape == (pos.size, pos.size)
    if case != "line_E4u":
        assert pos.size > prob.N // 2 and pos[-1] >= prob.N  # free cut points are unknowns
    assert np.array_equal(A, H.oracle_operator_matrix(cfg, d, pos))
    prob.close()


def test_hyperbolic_eigenvalue_statistic(ctx):
    """-h min Re(eig D_x) for E4u with the reference's scalar-wave closure
    parameters, Dirichlet inflow, floating outflow: not positive (every
    eigenvalue of D has a positive real part, so u_t + u_x = 0 is damped),
    small, shrinking under refinement, and equal to LAPACK on the oracle's
    matrix."""
    import shoccs_amd.spectral as SP

    stats = []
    for n in (21, 41):
        cfg = _line(n, O.E4U_1, [D, F], alpha=E4U_ALPHA, kind=O.SYS_WAVE_RADIAL, params=(-0.25, 0, 0, 0))
        prob = H.product_problem(ctx, cfg)
        st = SP.hyperbolic_eigenvalue_stat(prob)
        A = H.oracle_operator_matrix(cfg, 0, _positions(prob, 0))
        want = -prob.h()[0] * np.min(np.linalg.eigvals(A).real)
        assert st == pytest.approx(want, rel=1e-10)
        stats.append(st)
        prob.close()
    assert stats[0] <= 0.0 and stats[1] <= 0.0
    assert abs(stats[0]) < 0.05 and abs(stats[1]) < 0.25 * abs(stats[0])


HE = [13 / 100, 7 / 50, 3 / 20, 4 / 25, 17 / 100, 9 / 50, 3 / 25, 13 / 100, 7 / 50, 0, 0, 0, 0]


def test_hyperbolic_eigenvalues_known_answer_with_rects(ctx):
    """hyperbolic_eigenvalues.t.cpp:17-60: E2-poly between a Dirichlet yz_rect
    wall (psi 0.001, normal +1) and a floating one (psi 0.9, normal -1) on 21
    points; the statistic is within 1.2e-5 of zero.  The device operator
    equals the oracle's over the same unknowns, bit for bit."""
    import shoccs_amd.spectral as SP

    cfg = H.rect_1d_cfg(0.001, 0.9, n=21, alpha=HE)
    prob = H.product_problem(ctx, cfg)
    stat = SP.hyperbolic_eigenvalue_stat(prob)
    assert stat + 1.0 == pytest.approx(1.0, rel=1.2e-5)
    pos = _positions(prob, 0)
    np.testing.assert_array_equal(SP.operator_matrix(prob, 0), H.oracle_operator_matrix(cfg, 0, pos))
    prob.close()


EIGEN_LUA = """
-- the configuration of the reference's eigenvalues.lua: 21 points, two
-- yz_rect walls at psi = 1 - 1e-8 (Dirichlet, normal +1; floating, normal
-- -1), E2-poly with its floating and Dirichlet parameters
local psi = 1 - 1e-8
simulation = {
    mesh = { index_extents = {21}, domain_bounds = {1} },
    shapes = {
        { type = "yz_rect", psi = psi, normal = 1, boundary_condition = "dirichlet" },
        { type = "yz_rect", psi = psi, normal = -1, boundary_condition = "floating" },
    },
    scheme = { order = 1, type = "E2-poly",
               floating_alpha = { 0.7039278390946743, 0.5390086175376538, -0.647109821986589,
                                  0.2051508287133347, 0.6062051039572746, 0.8148425279273044 },
               dirichlet_alpha = { -0.10739761225713096, 0.8736492896991024, 0.40413606410467495 } },
    system = { type = "eigenvalues" },
}
"""


def test_eigenvalues_config_through_the_cli(ctx, tmp_path):
    """system.type "eigenvalues" run by python -m shoccs_amd: one MaxEigen
    row holding the statistic of the device operator, which equals the one
    computed from the oracle's assembled operator."""
    import shoccs_amd as S
    import shoccs_amd.__main__ as M
    import shoccs_amd.spectral as SP

    path = tmp_path / "eigenvalues.lua"
    path.write_text(EIGEN_LUA)
    csv = tmp_path / "system.csv"
    assert M.main([str(path), "--csv", str(csv)]) == 0
    lines = csv.read_text().splitlines()
    assert lines[0] == "Timestamp,MaxEigen" and len(lines) == 2
    cfg = S.Config(EIGEN_LUA)
    prob = cfg.problem(ctx)
    stat = SP.hyperbolic_eigenvalue_stat(prob)
    assert float(lines[1].split(",")[1]) == stat
    ocfg = H.rect_1d_cfg(1 - 1e-8, 1 - 1e-8, n=21, alpha=H.E2P_LUA[:9] + [0, 0, 0, 0])
    A = H.oracle_operator_matrix(ocfg, 0, _positions(prob, 0))
    lam = np.linalg.eigvals(A)
    assert stat == pytest.approx(float(-prob.h()[0] * np.min(lam.real)), rel=1e-12, abs=1e-14)
    prob.close()
