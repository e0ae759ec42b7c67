"""Sphere intersections (src/mesh/object_geometry.cpp:33-199) on the oracle
and the product's host assembly (CPU), against the reference's own known
answers in object_geometry.t.cpp:

* "sphere intersections" (:11-329): a 21 x 22 x 23 mesh on
  [-1, 1] x [-1, 2] x [0, 2.2] and a sphere of radius 0.25 at
  (0.01, -0.01, 0.5).  Per direction, the cut points' positions (exact
  intersections, computed by the reference's authors in Mathematica, to
  1e-14), their psi and solid coordinates in the reference's R order, and
  the solid points of that direction's lines.  The vectors sit in
  tests/golden/object_geometry_sphere.json (written by
  tests/golden/make_fixtures.py from the test's initializers);
* "grid-aligned sphere - cross-direction consistency" (:412-485): a sphere
  whose surface passes exactly through the grid node (1.245, 0.845) of a
  401 x 401 mesh; the x ray leaves the body there (solid coordinate 249,
  psi 1) and the y ray enters it (solid coordinate 170), so the two
  directions agree on the cell of the same boundary point.
"""
import json
import os

import numpy as np
import pytest

import helpers as H
import oracle_ffi as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "object_geometry_sphere.json")))
SLOW = {0: 1, 1: 0, 2: 0}
FAST = {0: 2, 1: 2, 2: 1}


def _mesh():
    return O.OracleMesh(tuple(GOLD["n"]), tuple(GOLD["lo"]), tuple(GOLD["hi"]),
                        shapes=[O.sphere(tuple(GOLD["center"]), GOLD["radius"])])


def _solid_points(mesh, d):
    """The points of the lines along d that no fluid segment covers (a
    segment ending at a body excludes that end, mesh.cpp:90-111), as sorted
    (i, j, k) triples."""
    n = [int(v) for v in mesh.m.n]
    covered = {}
    for a in mesh.line_array(d):
        lo = int(a[0]) + (1 if a[2] >= 0 else 0)
        hi = int(a[1]) - (1 if a[3] >= 0 else 0)
        covered.setdefault((int(a[4]), int(a[5])), []).append((lo, hi))
    out = []
    for s in range(n[SLOW[d]]):
        for f in range(n[FAST[d]]):
            segs = covered.get((s, f), [])
            for q in range(n[d]):
                if any(lo <= q <= hi for lo, hi in segs):
                    continue
                p = [0, 0, 0]
                p[d], p[SLOW[d]], p[FAST[d]] = q, s, f
                out.append(tuple(p))
    return sorted(out)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_sphere_intersections(d):
    m = _mesh()
    r = m.r_arrays(d)
    want = GOLD["dirs"][d]
    assert len(r["psi"]) == len(want["pos"])
    np.testing.assert_allclose(r["pos"], np.array(want["pos"]), rtol=0, atol=1e-14)
    np.testing.assert_allclose(r["psi"], np.array(want["psi"]), rtol=1e-12, atol=0)
    assert np.array_equal(r["solid"], np.array(want["solid_coord"]))
    assert _solid_points(m, d) == sorted(tuple(p) for p in want["solid"])


def test_sphere_solid_sets_agree_across_directions():
    """Every direction sees the same 44 solid points (the test's three S lists
    hold the same set)."""
    m = _mesh()
    sets = [set(_solid_points(m, d)) for d in range(3)]
    assert sets[0] == sets[1] == sets[2]
    assert len(sets[0]) == 44


def test_product_host_assembly_matches_sphere_intersections():
    """The product's host assembly of the same mesh and sphere: cut points,
    lines and fluid gather equal the oracle's, bit for bit."""
    # the geometry needs a system around it: heat, E2 (the segments below the
    # sphere are too short for the wave's E4u closures)
    cfg = dict(kind=O.SYS_HEAT, n=tuple(GOLD["n"]), lo=tuple(GOLD["lo"]), hi=tuple(GOLD["hi"]),
               shapes=[O.sphere(tuple(GOLD["center"]), GOLD["radius"])], obj_bc=[O.DIRICHLET],
               grid_bc=[O.DIRICHLET] * 6, scheme=O.E2_2, alpha=(0.0, 0.0), params=[1.0, O.MMS_SINPROD])
    osys = H.oracle_system(cfg)
    prob = H.product_problem(None, cfg)
    for d in range(3):
        ra, rb = osys.mesh.r_arrays(d), prob.r_arrays(d)
        for k in ("psi", "pos", "solid", "ray_outside", "shape_id"):
            assert np.array_equal(ra[k], rb[k]), (d, k)
        assert np.array_equal(osys.mesh.line_array(d), prob.line_array(d)), d
    assert np.array_equal(osys.mesh.fluid(), prob.fluid())


def test_grid_aligned_sphere_cross_direction_consistency():
    m = O.OracleMesh((401, 401, 1), (0.0, 0.0, 0.0), (2.0, 2.0, 0.0), shapes=[O.sphere((1.053, 0.901, 0.0), 0.2)])
    found = 0
    for d, outside, solid in ((0, 0, 249), (1, 1, 170)):
        r = m.r_arrays(d)
        hit = np.flatnonzero((np.abs(r["pos"][:, 0] - 1.245) < 1e-14) & (np.abs(r["pos"][:, 1] - 0.845) < 1e-14))
        assert len(hit) == 1, d
        q = hit[0]
        assert int(r["ray_outside"][q]) == outside
        assert int(r["solid"][q][d]) == solid
        if d == 0:
            assert r["psi"][q] == pytest.approx(1.0, abs=1e-8)
        found += 1
    assert found == 2
    # the product's host assembly casts the same rays
    cfg = dict(kind=O.SYS_HEAT, n=(401, 401, 1), lo=(0.0, 0.0, 0.0), hi=(2.0, 2.0, 0.0),
               shapes=[O.sphere((1.053, 0.901, 0.0), 0.2)], obj_bc=[O.DIRICHLET], grid_bc=[O.DIRICHLET] * 6,
               scheme=O.E2_2, alpha=(0.0, 0.0), params=[1.0, O.MMS_SINPROD])
    prob = H.product_problem(None, cfg)
    for d in range(2):
        ra, rb = m.r_arrays(d), prob.r_arrays(d)
        for k in ("psi", "pos", "solid", "ray_outside"):
            assert np.array_equal(ra[k], rb[k]), (d, k)
