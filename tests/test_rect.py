"""Rectangular bodies (src/mesh/rect.hpp:11-44, rect.cpp:8-33; yz_rect in the
Lua reader, object_geometry.cpp:268-290), on the oracle and the product's
host assembly (CPU).  Known answers from the reference's tests:

* shapes.t.cpp:64-149 "xy_rect" / "yz_rect": a ray meeting the rectangle at
  t = 2 hits it at the expected point, entering the fluid (normal +1) or
  leaving it (normal -1);
* object_geometry.t.cpp:331-410 "rect_intersections" and "1D
  rect_intersections": two cut points and two solid points;
* hyperbolic_eigenvalues.t.cpp:17-60: the E2-poly x derivative between a
  Dirichlet and a floating wall (psi 0.001 and 0.9) has -h min Re(lambda)
  within 1.2e-5 of zero.
"""
import numpy as np
import pytest

import helpers as H
import oracle_ffi as O

HE = [13 / 100, 7 / 50, 3 / 20, 4 / 25, 17 / 100, 9 / 50, 3 / 25, 13 / 100, 7 / 50, 0, 0, 0, 0]


def _r(mesh, d):
    return mesh.r_arrays(d)


def _solid(mesh, d):
    """Points of the lines along d outside every fluid segment (the count of
    object_geometry's S(d), object_geometry.cpp:136-199): a segment ending at a
    body excludes that end (mesh.cpp:90-111)."""
    n = [int(v) for v in mesh.m.n]
    fluid = 0
    for a in mesh.line_array(d):
        lo = a[0] + (1 if a[2] >= 0 else 0)
        hi = a[1] - (1 if a[3] >= 0 else 0)
        fluid += hi - lo + 1
    return n[0] * n[1] * n[2] - fluid


@pytest.mark.parametrize("normal", [1, -1])
def test_yz_rect_hit(normal):
    """shapes.t.cpp:108-149: c0 (2, 0, 1), c1 (2, 1, 3); the x line through
    (., 0.5, 1.1) meets it at (2, 0.5, 1.1)."""
    m = O.OracleMesh((9, 1, 1), (0.0, 0.5, 1.1), (4.0, 0.5, 1.1),
                     shapes=[O.rect(0, 2.0, (0.0, 1.0), (1.0, 3.0), normal)])
    r = _r(m, 0)
    assert len(r["psi"]) == 1
    assert np.array_equal(r["pos"][0], [2.0, 0.5, 1.1])
    assert bool(r["ray_outside"][0]) == (normal < 0)
    assert r["shape_id"][0] == 0


@pytest.mark.parametrize("normal", [1, -1])
def test_xy_rect_hit(normal):
    """shapes.t.cpp:64-106: c0 (0, 1, 2), c1 (1, 2, 2); the z line through
    (0.5, 1.1, .) meets it at (0.5, 1.1, 2)."""
    m = O.OracleMesh((1, 1, 9), (0.5, 1.1, 0.0), (0.5, 1.1, 4.0),
                     shapes=[O.rect(2, 2.0, (0.0, 1.0), (1.0, 2.0), normal)])
    r = _r(m, 2)
    assert len(r["psi"]) == 1
    assert np.array_equal(r["pos"][0], [0.5, 1.1, 2.0])
    assert bool(r["ray_outside"][0]) == (normal < 0)


def test_rect_intersections():
    """object_geometry.t.cpp:331-367: 11 x 2 mesh on [0, 1]^2, a yz_rect at
    x = 1 - 1e-6 spanning y, z in [-1, 2] with normal -1."""
    loc = 1.0 - 1e-6
    m = O.OracleMesh((11, 2, 1), (0.0, 0.0, 0.0), (1.0, 1.0, 0.0),
                     shapes=[O.rect(0, loc, (-1.0, -1.0), (2.0, 2.0), -1)])
    assert len(_r(m, 0)["psi"]) == 2
    assert _solid(m, 0) == 2
    # the fluid gather comes from the lines of the last direction with more
    # than one point (mesh.cpp:129-131): here y, which the plane never cuts
    assert len(_r(m, 1)["psi"]) == 0 and len(m.fluid()) == 22


def test_1d_rect_intersections():
    """object_geometry.t.cpp:369-410: 11 points on [0, 1], yz_rects at psi
    1e-6 (normal +1) and 0.99 (normal -1); the cut points keep those psi."""
    cfg = H.rect_1d_cfg(1e-6, 0.99, n=11, scheme=O.E4U_1, alpha=(0.0, 0.0))
    osys = H.oracle_system(cfg)
    r = _r(osys.mesh, 0)
    assert len(r["psi"]) == 2
    assert _solid(osys.mesh, 0) == 2
    assert r["psi"] == pytest.approx([1e-6, 0.99], rel=1e-6)
    assert list(r["solid"][:, 0]) == [0, 10]


def test_hyperbolic_eigenvalues_known_answer():
    """hyperbolic_eigenvalues.t.cpp:17-60 on the oracle's assembled operator:
    Dirichlet wall at psi 0.001 (inflow), floating wall at psi 0.9."""
    cfg = H.rect_1d_cfg(0.001, 0.9, n=21, alpha=HE)
    osys = H.oracle_system(cfg)
    m = osys.mesh
    # the unknowns: the points with a row of the x derivative (mesh points of
    # its segments, and the floating wall's cut point); Dirichlet points have
    # neither a row nor a column
    pos = np.arange(m.N + m.nR(0), dtype=np.int64)
    A = H.oracle_operator_matrix(cfg, 0, pos)
    keep = np.any(A != 0, axis=1)
    lam = np.linalg.eigvals(A[np.ix_(keep, keep)])
    stat = -m.h()[0] * np.min(lam.real)
    assert stat + 1.0 == pytest.approx(1.0, rel=1.2e-5)


def test_product_host_assembly_matches_oracle():
    """The product's host assembly over rectangles and a sphere: cut points,
    lines and fluid gather bit for bit (both walls of the 1-D problem, and a
    2-D problem with a sphere behind a yz_rect)."""
    cfgs = [H.rect_1d_cfg(0.001, 0.9, n=21, alpha=HE),
            H.rect_1d_cfg(1e-6, 0.99, n=11, scheme=O.E4U_1, alpha=(0.0, 0.0))]
    c2 = H.sphere_wave_cfg(24, scheme=O.E4U_1, alpha=(0.1, 0.2))
    c2["shapes"] = [O.sphere((0.55, 0.5, 0.5), 0.2), O.rect(0, 0.2 + 0.3 / 23, (-1, -1), (2, 2), 1)]
    c2["obj_bc"] = [O.DIRICHLET, O.DIRICHLET]
    c2.pop("spheres")
    cfgs.append(c2)
    for cfg in cfgs:
        osys = H.oracle_system(cfg)
        prob = H.product_problem(None, cfg)
        for d in range(3):
            ra, rb = osys.mesh.r_arrays(d), prob.r_arrays(d)
            for k in ("psi", "pos", "solid", "ray_outside", "shape_id"):
                assert np.array_equal(ra[k], rb[k]), (d, k)
            assert np.array_equal(osys.mesh.line_array(d), prob.line_array(d)), d
        assert np.array_equal(osys.mesh.fluid(), prob.fluid())


def test_lua_yz_rect_placement():
    """object_geometry.t.cpp "1D rect_intersections" read by the Lua subset:
    the psi placement and default corners give the oracle's shapes."""
    import shoccs_amd as S

    cfg = S.Config("""
simulation = {
    mesh = { index_extents = {11}, domain_bounds = {1} },
    shapes = {
        { type = "yz_rect", psi = 1e-6, normal = 1, boundary_condition = "dirichlet" },
        { type = "yz_rect", psi = 0.99, normal = -1, boundary_condition = "floating" }
    },
    scheme = { order = 1, type = "E4u" },
    system = { type = "scalar wave", velocity = {1, 0, 0}, initial_condition = { center = {0.5, 0, 0} } },
}
""")
    want = H.rect_1d_cfg(1e-6, 0.99, n=11)["shapes"]
    assert np.array_equal(np.array(cfg.shapes), np.array(want))
    prob = cfg.problem(None)
    assert prob.nR[0] == 2
