"""shoccs_amd -- the SHOCCS discrete-operator path on one MI355X.

Host-side mirror of the reference's operator interface over the C ABI of
libshoccs.so:

* ``Problem``   -- mesh + boundary conditions + stencil + system, as built by
  ``system::from_lua`` (src/systems/system.cpp:103-132);
* ``Field``     -- a device scalar {D, Rx, Ry, Rz} (src/fields/scalar.hpp);
* ``Problem.derivative(dir)(u, du, op)`` -- ``derivative::operator()``
  (src/operators/derivative.cpp:502-519);
* ``Problem.gradient(u)(dux, duy, duz)`` / ``Problem.laplacian(u)(du)``
  (src/operators/gradient.cpp:34-46, laplacian.cpp:37-58);
* ``Problem.rhs``, ``update_boundary``, ``initialize``, ``stats``
  (src/systems/system.hpp:22-65);
* ``Problem.rk4`` -- ``integrators::rk4::operator()`` (src/temporal/rk4.cpp:16-47)
  and ``Problem.rk4_fused`` -- the same step fused into one kernel per stage.

No compute call has a CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import weakref

import numpy as np

from ._lib import (BC_DIRICHLET, BC_FLOATING, BC_INFLOW, BC_NEUMANN, E2_2, E2P_1, E4_1, E4_2, E4U_1, EULER,
                   F32_COMP, F64, MMS_GAUSS, MMS_HOST, MMS_SINPROD, RK4, RK4_FUSED, SHAPE_RECT_X,
                   SHAPE_RECT_Y, SHAPE_RECT_Z, SHAPE_SPHERE, SYS_HEAT,
                   SYS_WAVE_LINEAR, SYS_WAVE_RADIAL, IODesc, MMS_GRAD, MMS_SCALAR, MmsFns, ProblemDesc, RunDesc,
                   RunResult, ShoccsError,
                   SHX_OK, ShxField, ShxLineBlock, check, lib)

__all__ = [
    "Context", "Problem", "Field", "Batch", "Rk4Graph", "FieldIO", "Euler", "ShoccsError", "BC_DIRICHLET", "BC_FLOATING", "BC_NEUMANN",
    "BC_INFLOW", "E2_2", "E4_2", "E4_1", "E4U_1", "E2P_1", "SYS_WAVE_RADIAL", "SYS_WAVE_LINEAR", "SYS_HEAT",
    "MMS_GAUSS", "MMS_SINPROD", "MMS_HOST", "F64", "F32_COMP", "RK4", "RK4_FUSED", "EULER", "eq", "plus_eq",
    "mms_functions", "sphere_shape", "rect_shape", "SHAPE_SPHERE", "SHAPE_RECT_X", "SHAPE_RECT_Y", "SHAPE_RECT_Z",
]


def sphere_shape(center, radius) -> list:
    """A shape record (shoccs_problem_desc.shapes8) of a sphere."""
    c = list(center) + [0.0] * (3 - len(center))
    return [float(SHAPE_SPHERE), c[0], c[1], c[2], float(radius), 0.0, 0.0, 0.0]


def rect_shape(axis: int, plane: float, c0, c1, normal: float) -> list:
    """A shape record of the rectangle normal to `axis` (0: yz_rect, 1:
    xz_rect, 2: xy_rect) at coordinate `plane`, bounded by c0 - c1 in the
    plane's (slow, fast) axes, fluid on the side of `normal` (its sign)."""
    return [float(SHAPE_RECT_X + axis), float(plane), c0[0], c0[1], c1[0], c1[1],
            1.0 if normal > 0 else -1.0, 0.0]


def mms_functions(call, ddt, grad, lap) -> MmsFns:
    """A manufactured solution for heat problems with params [k, MMS_HOST]
    from Python functions of (t, (x, y, z)) (grad returns three values), as
    the reference's lua_mms calls its Lua functions.  Pass the result as
    Problem(..., mms=...); the problem keeps it alive."""
    def scalar(f):
        return MMS_SCALAR(lambda _c, t, loc: float(f(t, (loc[0], loc[1], loc[2]))))

    def vec(_c, t, loc, g):
        gx, gy, gz = grad(t, (loc[0], loc[1], loc[2]))
        g[0], g[1], g[2] = float(gx), float(gy), float(gz)

    return MmsFns(None, scalar(call), scalar(ddt), MMS_GRAD(vec), scalar(lap))

eq = 0
plus_eq = 1


class Context:
    """One device and one HIP stream (shx_init)."""

    def __init__(self, device: int = 0):
        L = lib()
        p = C.c_void_p()
        rc = L.shx_init(device, C.byref(p))
        if rc != 0:
            raise ShoccsError(f"shx_init({device}) failed with code {rc}: no usable GPU")
        self.p = p
        self._problems = weakref.WeakSet()
        self._batches = weakref.WeakSet()

    def sync(self):
        check(lib().shx_sync(self.p), self.p, "shx_sync")

    def time_copy(self, n: int, reps: int = 20) -> float:
        """Mean device milliseconds of one n-double device-to-device copy
        (shx_time_copy) between two scratch buffers allocated here."""
        L = lib()
        a, b = C.c_void_p(), C.c_void_p()
        check(L.shx_alloc(self.p, 8 * n, C.byref(a)), self.p, "alloc")
        try:
            check(L.shx_alloc(self.p, 8 * n, C.byref(b)), self.p, "alloc")
            try:
                check(L.shx_fill(self.p, a, 1.0, n), self.p, "fill")
                ms = C.c_double()
                check(L.shx_time_copy(self.p, b, a, n, reps, C.byref(ms)), self.p, "time_copy")
                return ms.value
            finally:
                L.shx_free(self.p, b)
        finally:
            L.shx_free(self.p, a)

    def close(self):
        """Release every problem made on this context, then the context."""
        if getattr(self, "p", None):
            for bat in list(getattr(self, "_batches", ())):
                bat.close()
            for prob in list(getattr(self, "_problems", ())):
                prob.close()
            lib().shx_finalize(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Field:
    """A device scalar field of a problem's shape."""

    def __init__(self, problem: "Problem"):
        self.problem = problem
        self.c = ShxField()
        check(lib().shoccs_field_alloc(problem.p, C.byref(self.c)), problem.ctx.p, "field alloc")
        problem._fields.add(self)

    def ref(self):
        return C.byref(self.c)

    def upload(self, flat: np.ndarray):
        a = np.ascontiguousarray(flat, dtype=np.float64)
        assert a.size == self.problem.flat_size
        check(lib().shoccs_field_upload(self.problem.p, C.byref(self.c),
                                        a.ctypes.data_as(C.POINTER(C.c_double))),
              self.problem.ctx.p, "field upload")

    def download(self) -> np.ndarray:
        out = np.empty(self.problem.flat_size)
        check(lib().shoccs_field_download(self.problem.p, C.byref(self.c),
                                          out.ctypes.data_as(C.POINTER(C.c_double))),
              self.problem.ctx.p, "field download")
        return out

    def copy_from(self, other: "Field"):
        """Device copy of every part (D, Rx, Ry, Rz) of another field of the
        same problem (shx_copy)."""
        assert other.problem is self.problem
        ctxp = self.problem.ctx.p
        check(lib().shx_copy(ctxp, self.c.D, other.c.D, self.problem.N), ctxp, "copy")
        for d in range(3):
            n = self.problem.nR[d]
            if n:
                check(lib().shx_copy(ctxp, self.c.R[d], other.c.R[d], n), ctxp, "copy")

    def free(self):
        if getattr(self, "c", None) is not None and self.problem.p:
            lib().shoccs_field_free(self.problem.p, C.byref(self.c))
            self.c = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Config:
    """A Lua `simulation` configuration (shoccs_config_parse): the problem
    it describes and the driver settings of its integrator and
    step_controller."""

    def __init__(self, text: str):
        L = lib()
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        rc = L.shoccs_config_parse(text.encode(), C.byref(h), err, len(err))
        if rc != 0:
            raise ShoccsError(f"Lua configuration: {err.value.decode()}")
        self._h = None
        self.mms = None
        try:
            d, r = ProblemDesc(), RunDesc()
            check(L.shoccs_config_desc(h, C.byref(d), C.byref(r)), None, "config")
            # a "lua" manufactured solution runs in the configuration's
            # evaluator: the configuration then lives as long as this object
            self.mms = d.mms_fns if d.mms_fns else None
            # system.type: "eigenvalues" carries only the operators
            self.eigenvalues = L.shoccs_config_eigenvalues(h) == 1
            self.kind = d.kind
            self.n = tuple(d.n[a] for a in range(3))
            self.lo = tuple(d.min[a] for a in range(3))
            self.hi = tuple(d.max[a] for a in range(3))
            self.spheres = [[d.spheres[4 * i + c] for c in range(4)] for i in range(d.nshapes)]
            self.shapes = [[d.shapes8[8 * i + c] for c in range(8)] for i in range(d.nshapes)] \
                if d.shapes8 else None
            self.shape_bc = [d.shape_bc[i] for i in range(d.nshapes)]
            self.grid_bc = tuple(d.grid_bc[f] for f in range(6))
            self.scheme = d.scheme
            self.alpha = tuple(d.alpha[i] for i in range(13 if d.scheme == E2P_1 else 2))
            self.params = [d.params[i] for i in range(d.nparams)]
            self.integrator = r.integrator
            self.max_step = None if r.max_step == 2 ** 31 - 1 else r.max_step
            self.max_time = None if r.max_time >= np.finfo(np.float64).max else r.max_time
            self.h_cfl, self.p_cfl, self.min_dt = r.h_cfl, r.p_cfl, r.min_dt
            io = IODesc()
            has_io = L.shoccs_config_io(h, C.byref(io))
            self.io = None if has_io != 1 else dict(
                dir=io.dir.decode(), xdmf_filename=io.xdmf_filename.decode(),
                suffix_length=io.suffix_length,
                write_every_step=io.write_every_step if io.write_every_step > 0 else None,
                write_every_time=io.write_every_time if io.write_every_time > 0 else None)
        finally:
            if self.mms is None:
                L.shoccs_config_destroy(h)
            else:
                self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            lib().shoccs_config_destroy(self._h)
            self._h = None

    def mms_error(self) -> str:
        """The first error raised by a function of the "lua" manufactured
        solution ("" when none)."""
        return lib().shoccs_config_error(self._h).decode() if self._h else ""

    @classmethod
    def from_file(cls, path: str) -> "Config":
        with open(path) as f:
            return cls(f.read())

    def problem(self, ctx: "Context | None") -> "Problem":
        p = Problem(ctx, self.kind, self.n, self.lo, self.hi, spheres=self.spheres,
                    shape_bc=self.shape_bc, grid_bc=self.grid_bc, scheme=self.scheme,
                    alpha=self.alpha, params=self.params, mms=self.mms, shapes=self.shapes)
        p._config = self  # the solution's functions live in this configuration
        return p

    def run_kwargs(self) -> dict:
        return dict(integrator=self.integrator, max_step=self.max_step, max_time=self.max_time,
                    h_cfl=self.h_cfl, p_cfl=self.p_cfl, min_dt=self.min_dt)


class FieldIO:
    """The reference's field output (shoccs_io_*): raw FP64 dumps of each
    variable (D, then Rx, Ry, Rz), the cut-point geometry and an XDMF3 time
    series, written when the dump interval fires (and at step 0)."""

    def __init__(self, problem: "Problem", dir: str = "io", write_every_step: int | None = None,
                 write_every_time: float | None = None, suffix_length: int = 6,
                 xdmf_filename: str = "view.xmf"):
        d = IODesc()
        self._keep = (dir.encode(), xdmf_filename.encode())
        d.dir, d.xdmf_filename = self._keep
        d.suffix_length = suffix_length
        d.write_every_step = write_every_step or 0
        d.write_every_time = write_every_time or 0.0
        self.problem = problem
        self.h = C.c_void_p()
        check(lib().shoccs_io_create(problem.p, C.byref(d), C.byref(self.h)), None, "io create")

    def write(self, fields: dict, step: int, time: float, dt: float) -> bool:
        """Dump the host flat fields {name: array}; True when written."""
        names = (C.c_char_p * len(fields))(*[k.encode() for k in fields])
        arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in fields.values()]
        for a in arrs:
            assert a.size == self.problem.flat_size
        ptrs = (C.POINTER(C.c_double) * len(arrs))(*[a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs])
        wrote = C.c_int32()
        check(lib().shoccs_io_write(self.h, len(arrs), names, ptrs, step, time, dt, C.byref(wrote)), None,
              "io write")
        return bool(wrote.value)

    def close(self):
        if getattr(self, "h", None):
            lib().shoccs_io_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Problem:
    """Mesh, boundary conditions, stencil and system on one device.

    ``ctx=None`` builds the host structures only (no device, no compute).
    """

    def __init__(self, ctx: Context | None, kind: int, n, lo, hi, spheres=(), shape_bc=(),
                 grid_bc=(BC_FLOATING,) * 6, scheme: int = E4_1, alpha=(0.0, 0.0), params=(),
                 mesh_only: bool = False, mms=None, shapes=None):
        """mesh_only: the mesh and cut points alone (no operators or system).
        mms: for heat with params [k, MMS_HOST], the solution's functions
        (mms_functions(...), or a configuration's pointer).
        shapes: shape records in place of spheres (sphere_shape(),
        rect_shape()); shape_bc then has one condition per record."""
        L = lib()
        self.ctx = ctx
        d = ProblemDesc()
        d.kind = kind
        n = list(n) + [1] * (3 - len(n))
        for a in range(3):
            d.n[a] = int(n[a])
            d.min[a] = float(lo[a])
            d.max[a] = float(hi[a])
        sp = np.ascontiguousarray(np.array(spheres, dtype=np.float64).reshape(-1, 4))
        bc = np.ascontiguousarray(np.array(list(shape_bc) + [0], dtype=np.int32))
        pr = np.ascontiguousarray(np.array(list(params) + [0.0], dtype=np.float64))
        self._keep = (sp, bc, pr)
        d.nshapes = len(sp)
        d.spheres = sp.ctypes.data_as(C.POINTER(C.c_double))
        if shapes is not None:
            s8 = np.ascontiguousarray(np.array(shapes, dtype=np.float64).reshape(-1, 8))
            self._keep = self._keep + (s8,)
            d.nshapes = len(s8)
            d.shapes8 = s8.ctypes.data_as(C.POINTER(C.c_double))
        d.shape_bc = bc.ctypes.data_as(C.POINTER(C.c_int32))
        for f in range(6):
            d.grid_bc[f] = int(grid_bc[f])
        d.scheme = scheme
        for i, a in enumerate(alpha):
            d.alpha[i] = a
        d.params = pr.ctypes.data_as(C.POINTER(C.c_double))
        d.nparams = len(pr) - 1
        d.mesh_only = int(mesh_only)
        self._mms = mms
        if mms is not None:
            d.mms_fns = C.pointer(mms) if isinstance(mms, MmsFns) else mms
        p = C.c_void_p()
        rc = L.shoccs_problem_create(ctx.p if ctx else None, C.byref(d), C.byref(p))
        if rc != SHX_OK and L.shoccs_last_error():
            raise ShoccsError(f"shoccs_problem_create failed with code {rc}: {L.shoccs_last_error().decode()}")
        check(rc, ctx.p if ctx else None, "shoccs_problem_create")
        self.p = p
        self.n = tuple(n)
        self.kind = kind
        self.N = L.shoccs_problem_size(p, 0)
        self.nR = tuple(L.shoccs_problem_size(p, a + 1) for a in range(3))
        self.flat_size = self.N + sum(self.nR)
        self._sys = L.shoccs_problem_system(p)
        self._fields = weakref.WeakSet()
        if ctx is not None:
            ctx._problems.add(self)

    def close(self):
        """Free the problem's live fields, then the problem (device state
        included); safe to call more than once.  Batches that step this
        problem are closed first."""
        if getattr(self, "p", None):
            if self.ctx is not None:
                for bat in list(getattr(self.ctx, "_batches", ())):
                    if any(p is self for p in bat.problems):
                        bat.close()
            for eu in list(getattr(self, "_eulers", ())):
                eu.close()
            for f in list(getattr(self, "_fields", ())):
                f.free()
            lib().shoccs_problem_destroy(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -------------------------------------------------------- structures
    def r_arrays(self, d: int) -> dict:
        n = self.nR[d]
        psi = np.zeros(max(n, 1))
        pos = np.zeros((max(n, 1), 3))
        solid = np.zeros((max(n, 1), 3), np.int32)
        ro = np.zeros(max(n, 1), np.int32)
        sid = np.zeros(max(n, 1), np.int32)
        check(lib().shoccs_mesh_r(self.p, d, psi.ctypes.data_as(C.POINTER(C.c_double)),
                                  pos.ctypes.data_as(C.POINTER(C.c_double)),
                                  solid.ctypes.data_as(C.POINTER(C.c_int32)),
                                  ro.ctypes.data_as(C.POINTER(C.c_int32)),
                                  sid.ctypes.data_as(C.POINTER(C.c_int32))), None, "mesh_r")
        return dict(psi=psi[:n], pos=pos[:n], solid=solid[:n], ray_outside=ro[:n], shape_id=sid[:n])

    def line_array(self, d: int) -> np.ndarray:
        n = lib().shoccs_mesh_nlines(self.p, d)
        a = np.zeros((max(n, 1), 7), np.int64)
        check(lib().shoccs_mesh_lines(self.p, d, a.ctypes.data_as(C.POINTER(C.c_int64))), None, "lines")
        return a[:n]

    def fluid(self) -> np.ndarray:
        n = lib().shoccs_problem_size(self.p, 4)
        a = np.zeros(max(n, 1), np.int64)
        check(lib().shoccs_mesh_fluid(self.p, a.ctypes.data_as(C.POINTER(C.c_int64))), None, "fluid")
        return a[:n]

    def blocks(self, d: int) -> np.ndarray:
        n = lib().shoccs_deriv_nblocks(self.p, d)
        arr = (ShxLineBlock * max(n, 1))()
        check(lib().shoccs_deriv_blocks(self.p, d, arr), None, "blocks")
        names = [f[0] for f in ShxLineBlock._fields_]
        raw = np.frombuffer(arr, dtype=np.int32).reshape(-1, len(names))[:n].copy()
        return np.rec.fromarrays(raw.T, names=names)

    def pool(self, d: int) -> np.ndarray:
        n = lib().shoccs_deriv_npool(self.p, d)
        a = np.zeros(max(n, 1))
        check(lib().shoccs_deriv_pool(self.p, d, a.ctypes.data_as(C.POINTER(C.c_double))), None, "pool")
        return a[:n]

    def h(self) -> tuple:
        """Grid spacing per axis, (max - min) / (n - 1) (0 on an axis of one
        point)."""
        n = (C.c_int32 * 3)()
        lo, hi = (C.c_double * 3)(), (C.c_double * 3)()
        check(lib().shoccs_problem_geometry(self.p, n, lo, hi), None, "geometry")
        return tuple((hi[a] - lo[a]) / (n[a] - 1) if n[a] > 1 else 0.0 for a in range(3))

    def error_field(self, u: np.ndarray, sol: np.ndarray) -> np.ndarray:
        """|u - sol| at the fluid points and the cut points of non-Dirichlet
        bodies, 0 elsewhere and on Dirichlet grid faces (shoccs_error_field)."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        sol = np.ascontiguousarray(sol, dtype=np.float64)
        out = np.empty(self.flat_size)
        dpp = C.POINTER(C.c_double)
        check(lib().shoccs_error_field(self.p, u.ctypes.data_as(dpp), sol.ctypes.data_as(dpp),
                                       out.ctypes.data_as(dpp)), None, "error field")
        return out

    def cut_rows(self, d: int) -> dict:
        """Direction d's R-space rows (cut points of non-Dirichlet bodies):
        rset, rrow, rptr, rcol (>= 0: a point of D; -1 - c: point c of
        R(rset)), rw."""
        L = lib()
        nnz = C.c_int64()
        n = L.shoccs_deriv_ncut(self.p, d, C.byref(nnz))
        out = dict(rset=np.zeros(max(n, 1), np.int32), rrow=np.zeros(max(n, 1), np.int64),
                   rptr=np.zeros(n + 1, np.int64), rcol=np.zeros(max(nnz.value, 1), np.int64),
                   rw=np.zeros(max(nnz.value, 1)))
        args = [out[k].ctypes.data_as(C.POINTER(C.c_int32 if k == "rset" else
                                                C.c_double if k == "rw" else C.c_int64))
                for k in ("rset", "rrow", "rptr", "rcol", "rw")]
        check(L.shoccs_deriv_cut(self.p, d, *args), None, "cut rows")
        out["rset"], out["rrow"] = out["rset"][:n], out["rrow"][:n]
        out["rcol"], out["rw"] = out["rcol"][:nnz.value], out["rw"][:nnz.value]
        return out

    def interior(self, d: int) -> np.ndarray:
        a = np.zeros(9)
        w = lib().shoccs_deriv_interior(self.p, d, a.ctypes.data_as(C.POINTER(C.c_double)))
        return a[:w]

    def dt(self, h_cfl=1.0, p_cfl=1.0) -> float:
        return lib().shoccs_problem_dt(self.p, h_cfl, p_cfl)

    # ---------------------------------------------------------- compute
    @property
    def sys(self):
        """The device system (shx_system*)."""
        self._need_device()
        return self._sys

    def _need_device(self):
        if self.ctx is None:
            raise ShoccsError("problem was built without a device context")

    def field(self) -> Field:
        self._need_device()
        return Field(self)

    def derivative(self, d: int):
        self._need_device()
        dev = lib().shoccs_problem_derivative(self.p, d)

        def apply(u: Field, du: Field, op: int = eq, nu: Field | None = None):
            check(lib().shx_derivative_apply(dev, u.ref(), nu.ref() if nu else None, du.ref(), op),
                  self.ctx.p, "shx_derivative_apply")

        return apply

    def time_derivative(self, d: int, u: Field, du: Field, reps: int = 20, op: int = eq) -> float:
        """Mean device milliseconds of one direction-d derivative apply."""
        self._need_device()
        dev = lib().shoccs_problem_derivative(self.p, d)
        ms = C.c_double()
        check(lib().shx_time_derivative_apply(dev, u.ref(), du.ref(), op, reps, C.byref(ms)),
              self.ctx.p, "time_derivative")
        return ms.value

    def _dev_ops(self):
        arr = (C.c_void_p * 3)()
        for a in range(3):
            arr[a] = lib().shoccs_problem_derivative(self.p, a) if self.n[a] > 1 else None
        return arr

    def gradient(self, u: Field):
        self._need_device()

        def apply(dux: Field, duy: Field, duz: Field):
            outs = (C.POINTER(ShxField) * 3)(C.pointer(dux.c), C.pointer(duy.c), C.pointer(duz.c))
            check(lib().shx_gradient_apply(self._dev_ops(), u.ref(), outs), self.ctx.p, "gradient")

        return apply

    def laplacian(self, u: Field):
        self._need_device()

        def apply(du: Field):
            check(lib().shx_laplacian_apply(self._dev_ops(), u.ref(), None, du.ref()), self.ctx.p,
                  "laplacian")

        return apply

    def derivative_T(self, d: int):
        """Transpose of direction d's derivative (shx_derivative_apply_T):
        apply(dbar, ubar, op) sets (op = eq) or adds (op = plus_eq)
        ubar.D = O^T dbar.D and ubar.R[d] = B^T dbar.D."""
        self._need_device()
        dev = lib().shoccs_problem_derivative(self.p, d)

        def apply(dbar: Field, ubar: Field, op: int = eq):
            check(lib().shx_derivative_apply_T(dev, dbar.ref(), ubar.ref(), op), self.ctx.p,
                  "shx_derivative_apply_T")

        return apply

    def rhs_T(self, kbar: Field, wbar: Field):
        """wbar = (d rhs / d u)^T kbar (shx_system_rhs_T)."""
        self._need_device()
        check(lib().shx_system_rhs_T(self.sys, kbar.ref(), wbar.ref()), self.ctx.p, "rhs_T")

    def rk4_adjoint(self, u: Field, ubar: Field, t: float, dt: float,
                    coef_grad: np.ndarray | None = None):
        """Adjoint of one RK4 step from u at time t (shx_rk4_adjoint): ubar
        holds dJ/du_{n+1} on entry and dJ/du_n on exit; the step's
        interior-coefficient gradient is added to coef_grad (3 x 9) when given."""
        self._need_device()
        gp = None
        if coef_grad is not None:
            assert coef_grad.dtype == np.float64 and coef_grad.size == 27 and coef_grad.flags.c_contiguous
            gp = coef_grad.ctypes.data_as(C.POINTER(C.c_double))
        check(lib().shx_rk4_adjoint(self.sys, u.ref(), ubar.ref(), t, dt, gp), self.ctx.p,
              "rk4_adjoint")

    def rk4_fused_keep(self, u: Field, w1: Field, w2: Field, w3: Field, out: Field, t: float, dt: float):
        """One fused FP64 step u -> out keeping the stage inputs w1..w3
        (shx_rk4_fused_keep_async; asynchronous)."""
        self._need_device()
        check(lib().shx_rk4_fused_keep_async(self.sys, u.ref(), w1.ref(), w2.ref(), w3.ref(), out.ref(), t, dt),
              self.ctx.p, "rk4_fused_keep")

    def rk4_adjoint_stages(self, w, ubar: Field, dt: float, with_coef: bool = True):
        """Adjoint of one RK4 step over its kept stage inputs w = (u, w1, w2,
        w3) (shx_rk4_adjoint_stages_async; asynchronous); the coefficient
        gradient accumulates on the device (adjoint_coef_set / _get)."""
        self._need_device()
        arr = (ShxField * 4)(*[f.c for f in w])
        check(lib().shx_rk4_adjoint_stages_async(self.sys, arr, ubar.ref(), dt, int(with_coef)), self.ctx.p,
              "rk4_adjoint_stages")

    def adjoint_prepare(self):
        """Builds the adjoint structures now (shx_adjoint_prepare) instead of
        on the first adjoint call."""
        self._need_device()
        check(lib().shx_adjoint_prepare(self.sys), self.ctx.p, "adjoint_prepare")

    def adjoint_coef_set(self, g: np.ndarray):
        g = np.ascontiguousarray(g, dtype=np.float64).reshape(27)
        check(lib().shx_adjoint_coef_set(self.sys, g.ctypes.data_as(C.POINTER(C.c_double))), self.ctx.p,
              "adjoint_coef_set")

    def adjoint_coef_get(self) -> np.ndarray:
        g = np.zeros(27)
        check(lib().shx_adjoint_coef_get(self.sys, g.ctypes.data_as(C.POINTER(C.c_double))), self.ctx.p,
              "adjoint_coef_get")
        return g

    def device_interior(self, d: int) -> np.ndarray:
        """Interior stencil coefficients that direction d's device operator
        applies (those of the assembly unless replaced by set_interior)."""
        self._need_device()
        c = np.zeros(9)
        w = C.c_int()
        check(lib().shx_system_get_interior(self.sys, d, c.ctypes.data_as(C.POINTER(C.c_double)),
                                            C.byref(w)), self.ctx.p, "get_interior")
        return c[:w.value].copy()

    def set_interior(self, d: int, coef) -> None:
        self._need_device()
        c = np.ascontiguousarray(coef, dtype=np.float64)
        check(lib().shx_system_set_interior(self.sys, d, c.ctypes.data_as(C.POINTER(C.c_double)),
                                            c.size), self.ctx.p, "set_interior")

    def rhs(self, u: Field, out: Field, t: float, fused: bool = True):
        self._need_device()
        f = lib().shx_system_rhs if fused else lib().shx_system_rhs_unfused
        check(f(self.sys, u.ref(), out.ref(), t), self.ctx.p, "rhs")

    def update_boundary(self, u: Field, t: float):
        self._need_device()
        check(lib().shx_system_update_boundary(self.sys, u.ref(), t), self.ctx.p, "update_boundary")

    def initialize(self, u: Field, t: float = 0.0):
        self._need_device()
        check(lib().shx_system_initialize(self.sys, u.ref(), t), self.ctx.p, "initialize")

    def exact(self, out: Field, t: float):
        self._need_device()
        check(lib().shx_system_exact(self.sys, out.ref(), t), self.ctx.p, "exact")

    def stats(self, u: Field, t: float) -> np.ndarray:
        self._need_device()
        s = np.zeros(11)
        check(lib().shx_system_stats(self.sys, u.ref(), t, s.ctypes.data_as(C.POINTER(C.c_double))),
              self.ctx.p, "stats")
        return s

    def rk4(self, u0: Field, out: Field, rk: Field, srhs: Field, t: float, dt: float):
        self._need_device()
        check(lib().shx_rk4_step(self.sys, u0.ref(), out.ref(), rk.ref(), srhs.ref(), t, dt),
              self.ctx.p, "rk4")

    def euler(self, u0: Field, out: Field, srhs: Field, t: float, dt: float):
        """integrators::euler, the reference's slot schedule (shx_euler_step)."""
        self._need_device()
        check(lib().shx_euler_step(self.sys, u0.ref(), out.ref(), srhs.ref(), t, dt), self.ctx.p,
              "euler")

    def run(self, u: Field, integrator: int = RK4, max_step: int | None = None,
            max_time: float | None = None, h_cfl: float = 1.0, p_cfl: float = 1.0,
            min_dt: float = 1e-6, keep_stats: bool = True, io: "FieldIO | None" = None) -> dict:
        """simulation_cycle::run without I/O (shoccs_run): returns the final
        time, step count, whether the controller ran out (`completed`), the
        final L-inf error (NaN when ended prematurely) and the statistics of
        every step (row 0 at t = 0)."""
        self._need_device()
        d = RunDesc()
        d.integrator = integrator
        d.max_step = 2 ** 31 - 1 if max_step is None else max_step
        d.max_time = float(np.finfo(np.float64).max) if max_time is None else max_time
        d.h_cfl, d.p_cfl, d.min_dt = h_cfl, p_cfl, min_dt
        self.sys  # raises for a problem without a system
        cap = 0
        if keep_stats:
            if max_step is None and max_time is not None:
                # steps the controller takes to reach max_time (plus slack)
                dt = self.dt(h_cfl, p_cfl)
                cap = int(min(max_time / dt, 10 ** 7)) + 2 if dt > 0 else 100000
            else:
                cap = (max_step if max_step is not None else 100000)
            cap += 1
        buf = np.zeros((max(cap, 1), 11))
        d.stats = buf.ctypes.data_as(C.POINTER(C.c_double)) if cap else None
        d.stats_cap = cap
        d.io = io.h if io is not None else None
        r = RunResult()
        check(lib().shoccs_run(self.p, C.byref(d), u.ref(), C.byref(r)), self.ctx.p, "run")
        return dict(time=r.time, step=r.step, completed=bool(r.completed), linf=r.linf,
                    stats=buf[: r.nstats].copy())

    def rk4_fused(self, u0: Field, t: float, dt: float, nsteps: int = 1):
        self._need_device()
        check(lib().shx_rk4_fused_steps_async(self.sys, u0.ref(), t, dt, nsteps), self.ctx.p,
              "rk4_fused")
        self.ctx.sync()

    def set_tuning(self, variant: int, march_ty: int = 4, march_xchunk: int = 16):
        """Fused-stage kernel: 0 point-wise tiles; 1-4 x-marching tiles with the
        special-point work in a second launch (1) or in the same launch before
        (2), after (3) or among (4) the tiles."""
        self._need_device()
        check(lib().shx_system_set_tuning(self.sys, variant, march_ty, march_xchunk), self.ctx.p,
              "set_tuning")

    def set_precision(self, precision: int):
        """Fused-form precision: F64 (bit-identical to the reference's sums) or
        F32_COMP (FP32 stage register with compensated interior sums; u0, the
        accumulator, closures and B terms in FP64)."""
        self._need_device()
        check(lib().shx_system_set_precision(self.sys, precision), self.ctx.p, "set_precision")

    def precision(self) -> int:
        out = C.c_int()
        check(lib().shx_system_get_precision(self.sys, C.byref(out)), self.ctx.p, "get_precision")
        return out.value

    def tuning(self) -> tuple:
        out = (C.c_int * 3)()
        check(lib().shx_system_get_tuning(self.sys, out), self.ctx.p, "get_tuning")
        return tuple(out)

    def stage_points(self) -> tuple:
        """(points of D updated by the timed fused-stage launch, points among
        them written by the special-point work: those with a record and the
        Dirichlet grid-face points) for the current variant."""
        self._need_device()
        out = (C.c_int64 * 2)()
        check(lib().shx_system_stage_points(self.sys, out), self.ctx.p, "stage_points")
        return tuple(out)

    def autotune(self, u0: Field, t: float, dt: float) -> tuple:
        """Time every fused-stage candidate on a copy of u0 and keep the
        fastest; returns (variant, march_ty, march_xchunk, ms_per_step)."""
        self._need_device()
        ms = C.c_double()
        check(lib().shx_system_autotune(self.sys, u0.ref(), t, dt, C.byref(ms)), self.ctx.p, "autotune")
        return self.tuning() + (ms.value,)

    def time_fused(self, u0: Field, t: float, dt: float, nsteps: int, reps: int = 1) -> float:
        ms = C.c_double()
        check(lib().shx_time_fused_steps(self.sys, u0.ref(), t, dt, nsteps, reps, C.byref(ms)),
              self.ctx.p, "time_fused")
        return ms.value

    def time_fused_stage_kernel(self, u0: Field, t: float, dt: float, nsteps: int) -> float:
        ms = C.c_double()
        check(lib().shx_time_fused_stage_kernel(self.sys, u0.ref(), t, dt, nsteps, C.byref(ms)),
              self.ctx.p, "time_fused_stage_kernel")
        return ms.value

    def time_unfused(self, u0: Field, out: Field, rk: Field, srhs: Field, t: float, dt: float,
                     nsteps: int) -> float:
        ms = C.c_double()
        check(lib().shx_time_unfused_steps(self.sys, u0.ref(), out.ref(), rk.ref(), srhs.ref(), t, dt,
                                           nsteps, C.byref(ms)), self.ctx.p, "time_unfused")
        return ms.value


class Euler:
    """Compressible Euler (shx_euler_*) over a problem's mesh and
    first-derivative operators: five conserved fields (rho, rho u, rho v,
    rho w, E), slip-wall bodies, Dirichlet grid faces held.  The problem's
    bodies must be Dirichlet (its operators then carry no cut-point rows)."""

    def __init__(self, problem: "Problem", gamma: float = 1.4):
        problem._need_device()
        self.problem = problem
        self.gamma = gamma
        self.h = C.c_void_p()
        check(lib().shoccs_euler_create(problem.p, gamma, C.byref(self.h)), problem.ctx.p, "euler create")
        # the system uses the problem's operators: closing the problem closes it first
        if not hasattr(problem, "_eulers"):
            problem._eulers = weakref.WeakSet()
        problem._eulers.add(self)

    def fields(self) -> list:
        return [self.problem.field() for _ in range(5)]

    @staticmethod
    def _refs(U):
        return (ShxField * 5)(*[u.c for u in U])

    def upload(self, U, arrays):
        for u, a in zip(U, arrays):
            u.upload(a)

    def download(self, U) -> list:
        return [u.download() for u in U]

    def wall(self, U):
        check(lib().shx_euler_wall(self.h, self._refs(U)), self.problem.ctx.p, "euler wall")

    def rhs(self, U, out):
        check(lib().shx_euler_rhs(self.h, self._refs(U), self._refs(out)), self.problem.ctx.p, "euler rhs")

    def rk4_step(self, u0, out, rk, srhs, dt: float):
        check(lib().shx_euler_rk4_step(self.h, self._refs(u0), self._refs(out), self._refs(rk), self._refs(srhs),
                                       dt), self.problem.ctx.p, "euler rk4")

    def set_precision(self, precision: int):
        """F64 (default) or F32_COMP (the mixed form) for rk4_step and
        time_steps."""
        check(lib().shx_euler_set_precision(self.h, precision), self.problem.ctx.p, "euler precision")

    def time_steps(self, u0, out, rk, srhs, dt: float, nsteps: int) -> float:
        """Mean device milliseconds per step over nsteps steps (u0 <- out
        after each)."""
        ms = C.c_double()
        check(lib().shx_time_euler_steps(self.h, self._refs(u0), self._refs(out), self._refs(rk), self._refs(srhs),
                                         dt, nsteps, C.byref(ms)), self.problem.ctx.p, "euler time")
        return ms.value

    def close(self):
        if getattr(self, "h", None):
            lib().shx_euler_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Problems stepped together, one launch per RK4 stage (shx_batch_*).

    ``problems`` share the context, system family, stencil width and
    precision; ``fields[p]`` (one per problem) is advanced in place from time
    ``t0[p]`` with step ``dt[p]``.  ``capture(n)`` records n steps as a
    hipGraph that ``launch()`` replays.
    """

    def __init__(self, ctx: Context, problems, fields, t0, dt, march_ty: int = 4):
        n = len(problems)
        assert n == len(fields) == len(t0) == len(dt) and n > 0
        self.ctx = ctx
        self.problems = list(problems)
        self.fields = list(fields)  # the batch refers to their device memory
        syss = (C.c_void_p * n)(*[p.sys for p in problems])
        flds = (ShxField * n)(*[f.c for f in fields])
        t0a = np.ascontiguousarray(t0, dtype=np.float64)
        dta = np.ascontiguousarray(dt, dtype=np.float64)
        h = C.c_void_p()
        check(lib().shx_batch_create(ctx.p, syss, flds, t0a.ctypes.data_as(C.POINTER(C.c_double)),
                                     dta.ctypes.data_as(C.POINTER(C.c_double)), n, march_ty,
                                     C.byref(h)), ctx.p, "batch create")
        self.h = h
        self.n = n
        ctx._batches.add(self)

    def steps(self, nsteps: int):
        check(lib().shx_batch_steps_async(self.h, nsteps), self.ctx.p, "batch steps")
        self.ctx.sync()

    def capture(self, nsteps: int):
        check(lib().shx_batch_capture(self.h, nsteps), self.ctx.p, "batch capture")

    def launch(self):
        check(lib().shx_batch_launch_async(self.h), self.ctx.p, "batch launch")
        self.ctx.sync()

    def times(self) -> np.ndarray:
        out = np.zeros(self.n)
        check(lib().shx_batch_times(self.h, out.ctypes.data_as(C.POINTER(C.c_double))), self.ctx.p,
              "batch times")
        return out

    def time_steps(self, use_graph: bool, reps: int) -> tuple:
        """(ms per step, total points) over `reps` steps or graph replays."""
        ms = C.c_double()
        pts = C.c_int64()
        check(lib().shx_time_batch(self.h, int(use_graph), reps, C.byref(ms), C.byref(pts)),
              self.ctx.p, "batch time")
        return ms.value, pts.value

    def close(self):
        if getattr(self, "h", None) and getattr(self.ctx, "p", None):
            lib().shx_batch_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Rk4Graph:
    """One problem's fused RK4 steps captured as a hipGraph
    (shx_rk4_graph_*): ``launch()`` replays `nsteps` steps of ``field`` in
    place, the clock on the device advancing with every replay."""

    def __init__(self, problem: Problem, field: Field, t0: float, dt: float, nsteps: int):
        problem._need_device()
        self.ctx = problem.ctx
        self.problems = [problem]
        self.field = field  # the graph refers to its device memory
        h = C.c_void_p()
        check(lib().shx_rk4_graph_create(problem.sys, C.byref(field.c), t0, dt, nsteps, C.byref(h)),
              self.ctx.p, "rk4 graph create")
        self.h = h
        self.ctx._batches.add(self)

    def launch(self, sync: bool = True):
        check(lib().shx_rk4_graph_launch(self.h), self.ctx.p, "rk4 graph launch")
        if sync:
            self.ctx.sync()

    def time(self) -> float:
        t = C.c_double()
        check(lib().shx_rk4_graph_time(self.h, C.byref(t)), self.ctx.p, "rk4 graph time")
        return t.value

    def close(self):
        if getattr(self, "h", None) and getattr(self.ctx, "p", None):
            lib().shx_rk4_graph_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
