"""The product's host assembly (C++20, libshoccs.so, built without a device
context) against the oracle: cut points, masks, line extents, segment tables,
closure coefficients and B entries must be identical (bit for bit)."""
import numpy as np
import pytest

import helpers as H
import oracle_ffi as O

CFGS = {
    "wave48_E4_1": H.sphere_wave_cfg(48),
    "wave22_E4u_short_segments": H.sphere_wave_cfg(22, scheme=O.E4U_1, alpha=(0.1, 0.2)),
    "wave30_radial": H.sphere_wave_cfg(30, linear=False, scheme=O.E4U_1),
    "heat2d_101_config1": H.heat2d_cfg(101),
    "heat3d_noncubic": H.heat3d_cfg(),
    "heat3d_E2": H.heat3d_cfg(scheme=O.E2_2),
    "wave31_E2P_inflow": H.wave_e2p_cfg(31),
}


@pytest.mark.parametrize("name", list(CFGS))
def test_structures_bitexact(name):
    cfg = CFGS[name]
    osys = H.oracle_system(cfg)
    prob = H.product_problem(None, cfg)
    n = prob.n
    assert prob.N == osys.mesh.N
    for d in range(3):
        ra, rb = osys.mesh.r_arrays(d), prob.r_arrays(d)
        for k in ("psi", "pos", "solid", "ray_outside", "shape_id"):
            assert np.array_equal(ra[k], rb[k]), (d, k)
        assert np.array_equal(osys.mesh.line_array(d), prob.line_array(d)), d
        if n[d] < 2:
            continue
        so, Bo = H.oracle_segments(osys.deriv(d), n, d)
        sp, Bp = H.product_segments(prob, d)
        assert len(so) == len(sp)
        for a, b in zip(so, sp):
            for k in a:
                if isinstance(a[k], np.ndarray):
                    assert np.array_equal(a[k], b[k]), (d, k)
                else:
                    assert a[k] == b[k], (d, k, a, b)
        for x, y in zip(Bo, Bp):
            assert np.array_equal(x, y), d
        st = O.stencil(cfg["scheme"], cfg["alpha"])
        assert np.array_equal(O.interior(st, osys.mesh.h()[d]), prob.interior(d))
    assert np.array_equal(osys.mesh.fluid(), prob.fluid())


def test_timestep_matches():
    for cfg in CFGS.values():
        osys = H.oracle_system(cfg)
        prob = H.product_problem(None, cfg)
        assert osys.dt(0.5, 0.8) == prob.dt(0.5, 0.8)


@pytest.mark.parametrize("n,scheme,alpha", [(24, O.E4_1, (0.1, 0.7)), (20, O.E4U_1, (0.1, 0.2)),
                                             (41, O.E4_1, (0.1, 0.7))])
def test_rejects_overlapping_closures(n, scheme, alpha):
    """A segment between the sphere and a wall shorter than its two closure
    blocks would need a circulant with a negative row count
    (derivative.cpp:428-436), undefined in the reference: the oracle and the
    host assembly both refuse the mesh, the latter with a message that names
    the segment."""
    import shoccs_amd as S

    cfg = H.sphere_wave_cfg(n, scheme=scheme, alpha=alpha)
    with pytest.raises(ValueError):
        H.oracle_system(cfg)
    with pytest.raises(S.ShoccsError, match="shorter than its closures"):
        H.product_problem(None, cfg)


CUT = H.cut_cfgs()


@pytest.mark.parametrize("name", list(CUT))
def test_cut_rows_bitexact(name):
    """Cut-point rows of floating and Neumann bodies (cut_discretization,
    derivative.cpp:159-257) and the D-space segments next to them: the host
    assembly against the oracle, bit for bit."""
    cfg = CUT[name]
    osys = H.oracle_system(cfg)
    prob = H.product_problem(None, cfg)
    total = 0
    for d in range(3):
        if prob.n[d] < 2:
            continue
        want, got = H.oracle_cut_rows(osys, d), prob.cut_rows(d)
        for k in ("rset", "rrow", "rptr", "rcol", "rw"):
            assert np.array_equal(want[k], got[k]), (name, d, k)
        total += len(got["rw"])
        so, Bo = H.oracle_segments(osys.deriv(d), prob.n, d)
        sp, Bp = H.product_segments(prob, d)
        assert len(so) == len(sp)
        for a, b in zip(so, sp):
            for k in a:
                assert np.array_equal(a[k], b[k]) if isinstance(a[k], np.ndarray) else a[k] == b[k], (d, k)
        for x, y in zip(Bo, Bp):
            assert np.array_equal(x, y), d
    assert total > 0


def test_operator_only_problem_has_no_system():
    import shoccs_amd as S

    prob = H.product_problem(None, CUT["objects_E2"])
    with pytest.raises(S.ShoccsError, match="device context"):
        prob.sys
