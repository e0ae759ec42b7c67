"""The allocation-free fused RK4 entry point (shx_rk4_fused_steps_noalloc)
against the allocating one (shx_rk4_fused_steps_async), through the C ABI:
the same kernels on registers carved from the per-device workspace, so the
results must be identical bit for bit.  Two calls in a row, the second on a
smaller grid (the workspace keeps the first call's contents), both
precisions, and the size limit."""
import ctypes as C

import numpy as np
import pytest

import helpers as H
import oracle_ffi as O
import shoccs_amd as S
from shoccs_amd._lib import lib

pytestmark = pytest.mark.gpu

SHX_OK, SHX_EINVAL = 0, -1


def _initial(cfg):
    osys = H.oracle_system(cfg)
    u0 = osys.field()
    osys.initialize(u0, 0.0)
    osys.update_boundary(u0, 0.0)
    return osys, u0.flat()


def _noalloc(prob, field, t, dt, nsteps):
    rc = lib().shx_rk4_fused_steps_noalloc(prob.sys, field.ref(), t, dt, nsteps)
    prob.ctx.sync()
    return rc


def _compare(ctx, cfg, nsteps, precision=S.F64):
    osys, flat = _initial(cfg)
    prob = H.product_problem(ctx, cfg)
    if precision != S.F64:
        prob.set_precision(precision)
    dt = osys.dt(0.5, 0.5)
    a, b = prob.field(), prob.field()
    a.upload(flat)
    b.upload(flat)
    prob.rk4_fused(a, 0.0, dt, nsteps)
    prob.stage_points()  # builds the point classes (allocates; set-up)
    assert _noalloc(prob, b, 0.0, dt, nsteps) == SHX_OK, lib().shx_last_error(ctx.p)
    got, want = b.download(), a.download()
    assert np.all(np.isfinite(want))
    assert np.array_equal(got, want)
    return prob


@pytest.mark.parametrize("precision", [S.F64, S.F32_COMP])
def test_noalloc_matches_fused_two_calls(ctx, precision):
    big = H.sphere_wave_cfg(96, scheme=O.E4U_1, alpha=(-0.7733323791884821, 0.1623961700641681))
    small = H.heat3d_cfg(n=(41, 38, 36))
    _compare(ctx, big, 3, precision)
    # second call on a smaller grid: the workspace still holds the first
    # call's registers
    _compare(ctx, small, 2, precision)


def test_noalloc_rejects_grids_beyond_the_workspace(ctx):
    cfg = H.heat2d_cfg(1700)  # 2.89 million points: 69 MB of FP64 registers
    prob = H.product_problem(ctx, cfg)
    prob.stage_points()
    flat = np.random.default_rng(3).standard_normal(prob.flat_size)
    f = prob.field()
    f.upload(flat)
    rc = _noalloc(prob, f, 0.0, 1e-7, 1)
    assert rc == SHX_EINVAL
    assert "workspace" in lib().shx_last_error(ctx.p).decode()
    assert np.array_equal(f.download(), flat)  # nothing ran


def test_noalloc_needs_point_classes(ctx):
    cfg = H.heat3d_cfg(n=(20, 22, 24))
    osys, flat = _initial(cfg)
    prob = H.product_problem(ctx, cfg)
    f = prob.field()
    f.upload(flat)
    rc = _noalloc(prob, f, 0.0, 1e-5, 1)
    assert rc != SHX_OK
    assert np.array_equal(f.download(), flat)
