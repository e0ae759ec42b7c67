"""ctypes binding of libshoccs.so (include/shoccs_hip.h and include/shoccs.h).

The library is the product: HIP kernels for gfx950 behind a C ABI plus the
C++20 host library that assembles the operators.  There is no CPU fallback:
when the library is missing or no GPU is present the calls fail loudly.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libshoccs.so")

SHX_OK = 0
BC_DIRICHLET, BC_FLOATING, BC_NEUMANN, BC_INFLOW = 0, 1, 2, 3
SYS_WAVE_RADIAL, SYS_WAVE_LINEAR, SYS_HEAT = 0, 1, 2
MMS_GAUSS, MMS_SINPROD, MMS_HOST = 0, 1, 2
SHAPE_SPHERE, SHAPE_RECT_X, SHAPE_RECT_Y, SHAPE_RECT_Z = 0, 1, 2, 3
E2_2, E4_2, E4_1, E4U_1, E2P_1 = 0, 1, 2, 3, 4
F64, F32_COMP = 0, 1

dp = C.POINTER(C.c_double)
i64p = C.POINTER(C.c_int64)
i32p = C.POINTER(C.c_int32)


class ShxField(C.Structure):
    _fields_ = [("D", C.c_void_p), ("R", C.c_void_p * 3)]


class ShxLineBlock(C.Structure):
    _fields_ = [(name, C.c_int32) for name in (
        "pair", "lo", "nl", "ni", "nr", "lcols", "lcol0", "lcoef", "rcols", "rcol0", "rcoef",
        "lobj", "lbw", "robj", "rbw", "flags")]


# shx_mms_fns: a manufactured solution's functions of (time, location)
MMS_SCALAR = C.CFUNCTYPE(C.c_double, C.c_void_p, C.c_double, dp)
MMS_GRAD = C.CFUNCTYPE(None, C.c_void_p, C.c_double, dp, dp)


class MmsFns(C.Structure):
    _fields_ = [("ctx", C.c_void_p), ("call", MMS_SCALAR), ("ddt", MMS_SCALAR), ("grad", MMS_GRAD),
                ("lap", MMS_SCALAR)]


class ProblemDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("n", C.c_int32 * 3),
        ("min", C.c_double * 3),
        ("max", C.c_double * 3),
        ("nshapes", C.c_int32),
        ("spheres", dp),
        ("shape_bc", i32p),
        ("grid_bc", C.c_int32 * 6),
        ("scheme", C.c_int32),
        ("alpha", C.c_double * 13),
        ("params", dp),
        ("nparams", C.c_int32),
        ("mesh_only", C.c_int32),
        ("mms_fns", C.POINTER(MmsFns)),
        ("shapes8", dp),
    ]


class RunDesc(C.Structure):
    _fields_ = [
        ("integrator", C.c_int32),
        ("max_step", C.c_int32),
        ("max_time", C.c_double),
        ("h_cfl", C.c_double),
        ("p_cfl", C.c_double),
        ("min_dt", C.c_double),
        ("stats", dp),
        ("stats_cap", C.c_int64),
        ("io", C.c_void_p),
    ]


class IODesc(C.Structure):
    _fields_ = [
        ("dir", C.c_char_p),
        ("xdmf_filename", C.c_char_p),
        ("suffix_length", C.c_int32),
        ("write_every_step", C.c_int32),
        ("write_every_time", C.c_double),
    ]


class RunResult(C.Structure):
    _fields_ = [
        ("time", C.c_double),
        ("step", C.c_int32),
        ("completed", C.c_int32),
        ("linf", C.c_double),
        ("nstats", C.c_int64),
    ]


RK4, RK4_FUSED, EULER = 0, 1, 2


class ShoccsError(RuntimeError):
    pass


_lib = None

# (name, restype, argtypes) of every exported entry point
_SIGS = [
    ("shx_init", C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    ("shx_finalize", C.c_int, [C.c_void_p]),
    ("shx_last_error", C.c_char_p, [C.c_void_p]),
    ("shoccs_last_error", C.c_char_p, []),
    ("shx_sync", C.c_int, [C.c_void_p]),
    ("shx_alloc", C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    ("shx_free", C.c_int, [C.c_void_p, C.c_void_p]),
    ("shx_h2d", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    ("shx_d2h", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    ("shx_fill", C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_int64]),
    ("shx_copy", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]),
    ("shx_time_copy", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int, dp]),
    ("shx_derivative_create", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    ("shx_derivative_destroy", C.c_int, [C.c_void_p]),
    ("shx_derivative_apply", C.c_int, [C.c_void_p, C.POINTER(ShxField), C.POINTER(ShxField),
                                       C.POINTER(ShxField), C.c_int]),
    ("shx_block_apply", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    ("shx_csr_create", C.c_int, [C.c_void_p, C.c_int64, i64p, i64p, dp, C.POINTER(C.c_void_p)]),
    ("shx_csr_destroy", C.c_int, [C.c_void_p]),
    ("shx_csr_apply", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("shx_gradient_apply", C.c_int, [C.c_void_p * 3, C.POINTER(ShxField), C.POINTER(ShxField) * 3]),
    ("shx_laplacian_apply", C.c_int, [C.c_void_p * 3, C.POINTER(ShxField), C.POINTER(ShxField),
                                      C.POINTER(ShxField)]),
    ("shx_system_create", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    ("shx_system_destroy", C.c_int, [C.c_void_p]),
    ("shx_system_rhs", C.c_int, [C.c_void_p, C.POINTER(ShxField), C.POINTER(ShxField), C.c_double]),
    ("shx_system_rhs_unfused", C.c_int, [C.c_void_p, C.POINTER(ShxField), C.POINTER(ShxField),
                                         C.c_double]),
    ("shx_system_update_boundary", C.c_int, [C.c_void_p, C.POINTER(ShxField), C.c_double]),
    ("shx_system_initialize", C.c_int, [C.c_void_p, C.POINTER(ShxField), C.c_double]),
    ("shx_system_exact", C.c_int, [C.c_void_p, C.POINTER(ShxField), C.c_double]),
    ("shx_system_stats", C.c_int, [C.c_void_p, C.POINTER(ShxField), C.c_double, dp]),
    ("shx_euler_step", C.c_int, [C.c_void_p, C.POINTER(ShxField), C.POINTER(ShxField),
                                 C.POINTER(ShxField), C.c_double, C.c_double]),
    ("shx_rk4_step", C.c_int, [C.c_void_p] + [C.POINTER(ShxField)] * 4 + [C.c_double, C.c_double]),
    ("shx_rk4_fused_steps_async", C.c_int, [C.c_void_p, C.POINTER(ShxField), C.c_double, C.c_double,
                                            C.c_int]),
    ("shx_rk4_fused_steps_noalloc", C.c_int, [C.c_void_p, C.POINTER(ShxField), C.c_double, C.c_double,
                                            C.c_int]),
    ("shx_time_fused_steps", C.c_int, [C.c_void_p, C.POINTER(ShxField), C.c_double, C.c_double,
                                       C.c_int, C.c_int, dp]),
    ("shx_time_unfused_steps", C.c_int, [C.c_void_p] + [C.POINTER(ShxField)] * 4 +
     [C.c_double, C.c_double, C.c_int, dp]),
    ("shx_time_fused_stage_kernel", C.c_int, [C.c_void_p, C.POINTER(ShxField), C.c_double,
                                              C.c_double, C.c_int, dp]),
    ("shx_system_set_tuning", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    ("shx_system_get_tuning", C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    ("shx_system_stage_points", C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    ("shx_system_set_precision", C.c_int, [C.c_void_p, C.c_int]),
    ("shx_system_get_precision", C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    ("shx_system_autotune", C.c_int, [C.c_void_p, C.POINTER(ShxField), C.c_double, C.c_double, dp]),
    ("shoccs_problem_create", C.c_int, [C.c_void_p, C.POINTER(ProblemDesc), C.POINTER(C.c_void_p)]),
    ("shoccs_problem_destroy", C.c_int, [C.c_void_p]),
    ("shoccs_problem_system", C.c_void_p, [C.c_void_p]),
    ("shoccs_problem_derivative", C.c_void_p, [C.c_void_p, C.c_int]),
    ("shoccs_problem_size", C.c_int64, [C.c_void_p, C.c_int]),
    ("shoccs_problem_dt", C.c_double, [C.c_void_p, C.c_double, C.c_double]),
    ("shoccs_mesh_r", C.c_int, [C.c_void_p, C.c_int, dp, dp, i32p, i32p, i32p]),
    ("shoccs_mesh_nlines", C.c_int64, [C.c_void_p, C.c_int]),
    ("shoccs_mesh_lines", C.c_int, [C.c_void_p, C.c_int, i64p]),
    ("shoccs_mesh_fluid", C.c_int, [C.c_void_p, i64p]),
    ("shoccs_deriv_nblocks", C.c_int64, [C.c_void_p, C.c_int]),
    ("shoccs_deriv_blocks", C.c_int, [C.c_void_p, C.c_int, C.POINTER(ShxLineBlock)]),
    ("shoccs_deriv_npool", C.c_int64, [C.c_void_p, C.c_int]),
    ("shoccs_deriv_pool", C.c_int, [C.c_void_p, C.c_int, dp]),
    ("shoccs_deriv_interior", C.c_int, [C.c_void_p, C.c_int, dp]),
    ("shoccs_deriv_ncut", C.c_int64, [C.c_void_p, C.c_int, C.POINTER(C.c_int64)]),
    ("shoccs_deriv_cut", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                   C.POINTER(C.c_int64), C.POINTER(C.c_int64), dp]),
    ("shoccs_field_alloc", C.c_int, [C.c_void_p, C.POINTER(ShxField)]),
    ("shoccs_field_free", C.c_int, [C.c_void_p, C.POINTER(ShxField)]),
    ("shoccs_field_upload", C.c_int, [C.c_void_p, C.POINTER(ShxField), dp]),
    ("shoccs_field_download", C.c_int, [C.c_void_p, C.POINTER(ShxField), dp]),
    ("shx_batch_create", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(ShxField), dp, dp,
                                   C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    ("shx_batch_destroy", C.c_int, [C.c_void_p]),
    ("shx_batch_steps_async", C.c_int, [C.c_void_p, C.c_int]),
    ("shx_batch_capture", C.c_int, [C.c_void_p, C.c_int]),
    ("shx_batch_launch_async", C.c_int, [C.c_void_p]),
    ("shx_batch_times", C.c_int, [C.c_void_p, dp]),
    ("shx_rk4_fused_keep_async", C.c_int, [C.c_void_p, C.POINTER(ShxField), C.POINTER(ShxField),
                                           C.POINTER(ShxField), C.POINTER(ShxField), C.POINTER(ShxField),
                                           C.c_double, C.c_double]),
    ("shx_rk4_adjoint_stages_async", C.c_int, [C.c_void_p, C.POINTER(ShxField), C.POINTER(ShxField), C.c_double,
                                               C.c_int]),
    ("shx_adjoint_prepare", C.c_int, [C.c_void_p]),
    ("shx_adjoint_coef_set", C.c_int, [C.c_void_p, dp]),
    ("shx_adjoint_coef_get", C.c_int, [C.c_void_p, dp]),
    ("shx_rk4_graph_create", C.c_int, [C.c_void_p, C.POINTER(ShxField), C.c_double, C.c_double, C.c_int,
                                       C.POINTER(C.c_void_p)]),
    ("shx_rk4_graph_launch", C.c_int, [C.c_void_p]),
    ("shx_rk4_graph_time", C.c_int, [C.c_void_p, dp]),
    ("shx_rk4_graph_destroy", C.c_int, [C.c_void_p]),
    ("shx_time_derivative_apply", C.c_int, [C.c_void_p, C.POINTER(ShxField), C.POINTER(ShxField),
                                            C.c_int, C.c_int, dp]),
    ("shx_time_batch", C.c_int, [C.c_void_p, C.c_int, C.c_int, dp, C.POINTER(C.c_int64)]),
    ("shoccs_run", C.c_int, [C.c_void_p, C.POINTER(RunDesc), C.POINTER(ShxField),
                             C.POINTER(RunResult)]),
    ("shoccs_config_parse", C.c_int, [C.c_char_p, C.POINTER(C.c_void_p), C.c_char_p, C.c_int64]),
    ("shoccs_config_desc", C.c_int, [C.c_void_p, C.POINTER(ProblemDesc), C.POINTER(RunDesc)]),
    ("shoccs_config_io", C.c_int, [C.c_void_p, C.POINTER(IODesc)]),
    ("shoccs_io_create", C.c_int, [C.c_void_p, C.POINTER(IODesc), C.POINTER(C.c_void_p)]),
    ("shoccs_io_destroy", C.c_int, [C.c_void_p]),
    ("shoccs_io_write", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_char_p), C.POINTER(dp), C.c_int32,
                                  C.c_double, C.c_double, C.POINTER(C.c_int32)]),
    ("shoccs_error_field", C.c_int, [C.c_void_p, dp, dp, dp]),
    ("shoccs_euler_create", C.c_int, [C.c_void_p, C.c_double, C.POINTER(C.c_void_p)]),
    ("shx_euler_create", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    ("shx_euler_destroy", C.c_int, [C.c_void_p]),
    ("shx_euler_wall", C.c_int, [C.c_void_p, C.POINTER(ShxField)]),
    ("shx_euler_rhs", C.c_int, [C.c_void_p, C.POINTER(ShxField), C.POINTER(ShxField)]),
    ("shx_euler_rk4_step", C.c_int, [C.c_void_p] + [C.POINTER(ShxField)] * 4 + [C.c_double]),
    ("shx_euler_set_precision", C.c_int, [C.c_void_p, C.c_int]),
    ("shx_time_euler_steps", C.c_int, [C.c_void_p] + [C.POINTER(ShxField)] * 4 +
     [C.c_double, C.c_int, C.POINTER(C.c_double)]),
    ("shoccs_problem_geometry", C.c_int, [C.c_void_p, C.POINTER(C.c_int32), dp, dp]),
    ("shoccs_config_destroy", C.c_int, [C.c_void_p]),
    ("shoccs_config_error", C.c_char_p, [C.c_void_p]),
    ("shoccs_config_eigenvalues", C.c_int, [C.c_void_p]),
    ("shx_derivative_apply_T", C.c_int, [C.c_void_p, C.POINTER(ShxField), C.POINTER(ShxField),
                                         C.c_int]),
    ("shx_system_rhs_T", C.c_int, [C.c_void_p, C.POINTER(ShxField), C.POINTER(ShxField)]),
    ("shx_rk4_adjoint", C.c_int, [C.c_void_p, C.POINTER(ShxField), C.POINTER(ShxField), C.c_double,
                                  C.c_double, dp]),
    ("shx_system_set_interior", C.c_int, [C.c_void_p, C.c_int, dp, C.c_int]),
    ("shx_system_get_interior", C.c_int, [C.c_void_p, C.c_int, dp, C.POINTER(C.c_int)]),
]

EXPORTED = [s[0] for s in _SIGS]


def lib():
    """Load libshoccs.so (building it first only if it is absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        from . import build as _build

        _build.build()
    L = C.CDLL(LIB_PATH)
    for name, res, args in _SIGS:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc: int, ctx=None, what: str = ""):
    if rc != SHX_OK:
        msg = lib().shx_last_error(ctx).decode() if ctx else ""
        raise ShoccsError(f"{what} failed with code {rc}: {msg}")
