"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure only: the oracle restates the reference's algorithm on the
CPU and is the checker the HIP path is compared against.  Only tests/,
``__graft_entry__.smoke()`` and the cpu_baseline leg of ``bench.py`` use it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "_build", "liboracle.so")

DIRICHLET, FLOATING, NEUMANN, INFLOW = 0, 1, 2, 3
E2_2, E4_2, E4_1, E4U_1, E2P_1 = 0, 1, 2, 3, 4
SYS_WAVE_RADIAL, SYS_WAVE_LINEAR, SYS_HEAT = 0, 1, 2
MMS_GAUSS, MMS_SINPROD, MMS_HOST = 0, 1, 2

dp = C.POINTER(C.c_double)
lp = C.POINTER(C.c_long)
ip = C.POINTER(C.c_int)


class Info(C.Structure):
    _fields_ = [("p", C.c_int), ("r", C.c_int), ("t", C.c_int), ("x", C.c_int)]


class Stencil(C.Structure):
    _fields_ = [("kind", C.c_int), ("alpha", C.c_double * 13)]


class RInfo(C.Structure):
    _fields_ = [
        ("psi", C.c_double),
        ("pos", C.c_double * 3),
        ("normal", C.c_double * 3),
        ("ray_outside", C.c_int),
        ("solid", C.c_int * 3),
        ("shape_id", C.c_int),
    ]


class Bound(C.Structure):
    _fields_ = [
        ("mesh", C.c_int * 3),
        ("has_obj", C.c_int),
        ("obj_coord", C.c_long),
        ("obj_id", C.c_long),
        ("psi", C.c_double),
    ]


class Line(C.Structure):
    _fields_ = [("stride", C.c_long), ("start", Bound), ("end", Bound)]


class Mesh(C.Structure):
    _fields_ = [
        ("n", C.c_int * 3),
        ("min", C.c_double * 3),
        ("max", C.c_double * 3),
        ("h", C.c_double * 3),
        ("dims", C.c_int),
        ("xyz", dp * 3),
        ("nshapes", C.c_int),
        ("spheres", dp),
        ("R", C.POINTER(RInfo) * 3),
        ("nR", C.c_long * 3),
        ("lines", C.POINTER(Line) * 3),
        ("nlines", C.c_long * 3),
        ("fluid", lp),
        ("nfluid", C.c_long),
        ("shapes8", dp),
    ]


class IBlock(C.Structure):
    _fields_ = [
        ("row_offset", C.c_long),
        ("col_offset", C.c_long),
        ("stride", C.c_long),
        ("left_rows", C.c_long),
        ("left_cols", C.c_long),
        ("left", dp),
        ("interior_rows", C.c_long),
        ("right_rows", C.c_long),
        ("right_cols", C.c_long),
        ("right_col_offset", C.c_long),
        ("right", dp),
        ("flags", C.c_int),
    ]


class Csr(C.Structure):
    _fields_ = [("rows", C.c_long), ("nnz", C.c_long), ("u", lp), ("v", lp), ("w", dp)]


class Deriv(C.Structure):
    _fields_ = [
        ("dir", C.c_int),
        ("N", C.c_long),
        ("w", C.c_int),
        ("interior", C.c_double * 9),
        ("blocks", C.POINTER(IBlock)),
        ("nblocks", C.c_long),
        ("B", Csr),
        ("Nm", Csr),
        ("Bf", Csr * 3),
        ("Br", Csr * 3),
    ]


class Field(C.Structure):
    _fields_ = [("D", dp), ("R", dp * 3)]


# orc_mms_fns: a manufactured solution given by functions of (time, location)
MMS_SCALAR = C.CFUNCTYPE(C.c_double, C.c_void_p, C.c_double, dp)
MMS_GRAD = C.CFUNCTYPE(None, C.c_void_p, C.c_double, dp, dp)


class MmsFns(C.Structure):
    _fields_ = [("ctx", C.c_void_p), ("call", MMS_SCALAR), ("ddt", MMS_SCALAR), ("grad", MMS_GRAD),
                ("lap", MMS_SCALAR)]


def mms_fns(call, ddt, grad, lap) -> MmsFns:
    """MmsFns from Python functions of (t, (x, y, z)); grad returns three
    values.  The caller keeps the result alive while it is used."""
    def scalar(f):
        return MMS_SCALAR(lambda _c, t, loc: float(f(t, (loc[0], loc[1], loc[2]))))

    def vec(_c, t, loc, g):
        gx, gy, gz = grad(t, (loc[0], loc[1], loc[2]))
        g[0], g[1], g[2] = float(gx), float(gy), float(gz)

    return MmsFns(None, scalar(call), scalar(ddt), MMS_GRAD(vec), scalar(lap))


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    L.orc_mesh_create.restype = C.POINTER(Mesh)
    L.orc_mesh_create.argtypes = [ip, dp, dp, C.c_int, dp]
    L.orc_mesh_create8.restype = C.POINTER(Mesh)
    L.orc_mesh_create8.argtypes = [ip, dp, dp, C.c_int, dp]
    L.orc_system_create8.restype = C.c_void_p
    L.orc_system_create8.argtypes = [C.c_int, ip, dp, dp, C.c_int, dp, ip, ip, C.POINTER(Stencil), dp, C.c_int]
    L.orc_mesh_destroy.argtypes = [C.POINTER(Mesh)]
    L.orc_query.restype = Info
    L.orc_query.argtypes = [C.POINTER(Stencil), C.c_int]
    L.orc_query_max.restype = Info
    L.orc_query_max.argtypes = [C.POINTER(Stencil)]
    L.orc_interior.argtypes = [C.POINTER(Stencil), C.c_double, dp]
    L.orc_nbs.argtypes = [C.POINTER(Stencil), C.c_double, C.c_int, C.c_double, C.c_int, dp, dp]
    L.orc_deriv_create.restype = C.POINTER(Deriv)
    L.orc_deriv_create.argtypes = [C.POINTER(Mesh), C.c_int, C.POINTER(Stencil), ip, ip]
    L.orc_deriv_destroy.argtypes = [C.POINTER(Deriv)]
    L.orc_deriv_apply.argtypes = [C.POINTER(Deriv), C.POINTER(Field), C.POINTER(Field),
                                  C.POINTER(Field), C.c_int]
    L.orc_block_apply.argtypes = [C.POINTER(Deriv), dp, dp, C.c_int]
    L.orc_system_create.restype = C.c_void_p
    L.orc_system_create.argtypes = [C.c_int, ip, dp, dp, C.c_int, dp, ip, ip,
                                    C.POINTER(Stencil), dp, C.c_int]
    L.orc_system_destroy.argtypes = [C.c_void_p]
    L.orc_system_set_mms.argtypes = [C.c_void_p, C.POINTER(MmsFns)]
    L.orc_system_mesh.restype = C.POINTER(Mesh)
    L.orc_system_mesh.argtypes = [C.c_void_p]
    L.orc_system_deriv.restype = C.POINTER(Deriv)
    L.orc_system_deriv.argtypes = [C.c_void_p, C.c_int]
    for name in ("orc_system_update_boundary", "orc_system_initialize", "orc_system_exact"):
        getattr(L, name).argtypes = [C.c_void_p, C.POINTER(Field), C.c_double]
    L.orc_system_rhs.argtypes = [C.c_void_p, C.POINTER(Field), C.POINTER(Field), C.c_double]
    L.orc_system_stats.argtypes = [C.c_void_p, C.POINTER(Field), C.c_double, dp]
    L.orc_system_dt.restype = C.c_double
    L.orc_system_dt.argtypes = [C.c_void_p, C.c_double, C.c_double]
    L.orc_rk4_step.argtypes = [C.c_void_p] + [C.POINTER(Field)] * 4 + [C.c_double, C.c_double]
    L.orc_euler_step.argtypes = [C.c_void_p] + [C.POINTER(Field)] * 3 + [C.c_double, C.c_double]
    L.orc_gradient_apply.argtypes = [C.c_void_p] + [C.POINTER(Field)] * 4
    L.orc_laplacian_apply.argtypes = [C.c_void_p] + [C.POINTER(Field)] * 3
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_euler_create.restype = C.c_void_p
    L.orc_euler_create.argtypes = [C.POINTER(Mesh), C.POINTER(C.POINTER(Deriv)), C.c_double, ip]
    L.orc_euler_destroy.argtypes = [C.c_void_p]
    L.orc_euler_wall.argtypes = [C.c_void_p, C.POINTER(Field)]
    L.orc_euler_rhs.argtypes = [C.c_void_p, C.POINTER(Field), C.POINTER(Field)]
    L.orc_euler_rk4_step.argtypes = [C.c_void_p] + [C.POINTER(Field)] * 4 + [C.c_double]
    _lib = L
    return L


def _arr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def stencil(kind: int, alpha=(0.0, 0.0)) -> Stencil:
    s = Stencil()
    s.kind = kind
    for i, a in enumerate(alpha):
        s.alpha[i] = a
    return s


def nbs(st: Stencil, h: float, bc: int, psi: float, right: bool):
    L = lib()
    q = L.orc_query(C.byref(st), bc)
    c = np.zeros(64)
    ex = np.zeros(8)
    L.orc_nbs(C.byref(st), h, bc, psi, int(right), _arr(c, C.c_double), _arr(ex, C.c_double))
    return c[: q.r * q.t].copy(), ex[: q.x].copy(), q


def interior(st: Stencil, h: float):
    L = lib()
    q = L.orc_query_max(C.byref(st))
    c = np.zeros(16)
    L.orc_interior(C.byref(st), h, _arr(c, C.c_double))
    return c[: 2 * q.p + 1].copy()


class OracleField:
    """Owns numpy buffers D, Rx, Ry, Rz and a ctypes Field pointing at them."""

    def __init__(self, N: int, nR):
        self.D = np.zeros(N)
        self.R = [np.zeros(max(int(n), 1)) for n in nR]
        self.nR = [int(n) for n in nR]
        self.c = Field()
        self.c.D = _arr(self.D, C.c_double)
        for d in range(3):
            self.c.R[d] = _arr(self.R[d], C.c_double)

    def ref(self):
        return C.byref(self.c)

    def flat(self) -> np.ndarray:
        return np.concatenate([self.D] + [self.R[d][: self.nR[d]] for d in range(3)])


def sphere(center, radius):
    """A shape record (orc_mesh.shapes8) of a sphere."""
    c = list(center) + [0.0] * (3 - len(center))
    return [0.0, c[0], c[1], c[2], radius, 0.0, 0.0, 0.0]


def rect(axis, plane, c0, c1, normal):
    """A shape record of the rectangle normal to `axis` at `plane`, bounded
    by c0 - c1 in that plane's (slow, fast) axes, fluid on the side of
    `normal` (+1 or -1; rect.cpp:8-19)."""
    return [1.0 + axis, plane, c0[0], c0[1], c1[0], c1[1], 1.0 if normal > 0 else -1.0, 0.0]


def _shapes8(shapes):
    return np.ascontiguousarray(np.array(shapes, dtype=np.float64).reshape(-1, 8)) if len(shapes) \
        else np.zeros((0, 8))


class OracleMesh:
    def __init__(self, n, lo, hi, spheres=(), shapes=None):
        """spheres: n x {cx, cy, cz, r}; or shapes: shape records (sphere(), rect())."""
        L = lib()
        self.n = np.array(list(n) + [1] * (3 - len(n)), dtype=np.intc)
        self.lo = np.array(lo, dtype=np.float64)
        self.hi = np.array(hi, dtype=np.float64)
        if shapes is not None:
            self._sp = _shapes8(shapes)
            self.p = L.orc_mesh_create8(_arr(self.n, C.c_int), _arr(self.lo, C.c_double),
                                        _arr(self.hi, C.c_double), len(self._sp), _arr(self._sp, C.c_double))
        else:
            sp = np.array(spheres, dtype=np.float64).reshape(-1, 4) if len(spheres) else np.zeros((0, 4))
            self._sp = np.ascontiguousarray(sp)
            self.p = L.orc_mesh_create(_arr(self.n, C.c_int), _arr(self.lo, C.c_double),
                                       _arr(self.hi, C.c_double), len(sp), _arr(self._sp, C.c_double))
        self.own = True

    @classmethod
    def borrowed(cls, ptr):
        self = cls.__new__(cls)
        self.p = ptr
        self.own = False
        return self

    def __del__(self):
        if getattr(self, "own", False) and self.p:
            lib().orc_mesh_destroy(self.p)
            self.p = None

    @property
    def m(self) -> Mesh:
        return self.p.contents

    @property
    def N(self) -> int:
        n = self.m.n
        return n[0] * n[1] * n[2]

    def nR(self, d):
        return self.m.nR[d]

    def R(self, d):
        m = self.m
        out = []
        for i in range(m.nR[d]):
            r = m.R[d][i]
            out.append(dict(psi=r.psi, pos=tuple(r.pos), normal=tuple(r.normal),
                            ray_outside=bool(r.ray_outside), solid=tuple(r.solid),
                            shape_id=r.shape_id))
        return out

    def r_arrays(self, d):
        """R(d) as numpy arrays: psi, pos[n,3], solid[n,3], ray_outside, shape_id."""
        m = self.m
        n = m.nR[d]
        if n == 0:
            return dict(psi=np.zeros(0), pos=np.zeros((0, 3)), solid=np.zeros((0, 3), np.int32),
                        ray_outside=np.zeros(0, np.int32), shape_id=np.zeros(0, np.int32))
        raw = np.ctypeslib.as_array(C.cast(m.R[d], C.POINTER(C.c_byte)),
                                    shape=(n * C.sizeof(RInfo),)).copy()
        dt = np.dtype([("psi", "f8"), ("pos", "f8", 3), ("normal", "f8", 3), ("ray_outside", "i4"),
                       ("solid", "i4", 3), ("shape_id", "i4")], align=True)
        assert dt.itemsize == C.sizeof(RInfo)
        rec = raw.view(dt)
        return dict(psi=rec["psi"].copy(), pos=rec["pos"].copy(), solid=rec["solid"].copy(),
                    ray_outside=rec["ray_outside"].copy(), shape_id=rec["shape_id"].copy())

    def lines(self, d):
        m = self.m
        out = []
        for i in range(m.nlines[d]):
            ln = m.lines[d][i]

            def b(x):
                return dict(mesh=tuple(x.mesh), obj=(x.obj_coord, x.obj_id, x.psi) if x.has_obj else None)

            out.append(dict(stride=ln.stride, start=b(ln.start), end=b(ln.end)))
        return out

    def line_array(self, d):
        """Lines as an int64 array [n, 7]: start_mesh[d], end_mesh[d], start_obj, end_obj,
        slow, fast, stride (obj = R index or -1)."""
        m = self.m
        slow = {0: 1, 1: 0, 2: 0}[d]
        fast = {0: 2, 1: 2, 2: 1}[d]
        a = np.zeros((m.nlines[d], 7), dtype=np.int64)
        for i in range(m.nlines[d]):
            ln = m.lines[d][i]
            a[i] = (ln.start.mesh[d], ln.end.mesh[d],
                    ln.start.obj_coord if ln.start.has_obj else -1,
                    ln.end.obj_coord if ln.end.has_obj else -1,
                    ln.start.mesh[slow], ln.start.mesh[fast], ln.stride)
        return a

    def fluid(self):
        m = self.m
        return np.ctypeslib.as_array(m.fluid, shape=(m.nfluid,)).copy() if m.nfluid else np.zeros(0, np.int64)

    def h(self):
        return tuple(self.m.h)


def csr_arrays(A: Csr):
    rows = A.rows
    u = np.ctypeslib.as_array(A.u, shape=(rows + 1,)).copy()
    nnz = A.nnz
    v = np.ctypeslib.as_array(A.v, shape=(max(nnz, 1),))[:nnz].copy()
    w = np.ctypeslib.as_array(A.w, shape=(max(nnz, 1),))[:nnz].copy()
    return u, v, w


def deriv_blocks(dp_) -> list:
    d = dp_.contents
    out = []
    for i in range(d.nblocks):
        b = d.blocks[i]
        L = np.ctypeslib.as_array(b.left, shape=(b.left_rows * b.left_cols,)).copy() if b.left_rows * b.left_cols else np.zeros(0)
        R = np.ctypeslib.as_array(b.right, shape=(b.right_rows * b.right_cols,)).copy() if b.right_rows * b.right_cols else np.zeros(0)
        out.append(dict(row_offset=b.row_offset, col_offset=b.col_offset, stride=b.stride,
                        left_rows=b.left_rows, left_cols=b.left_cols, left=L,
                        interior_rows=b.interior_rows, right_rows=b.right_rows,
                        right_cols=b.right_cols, right_col_offset=b.right_col_offset, right=R,
                        flags=b.flags))
    return out


class OracleSystem:
    """A heat or scalar-wave system restated on the CPU."""

    def __init__(self, kind, n, lo, hi, spheres, obj_bc, grid_bc, st: Stencil, params, shapes=None):
        """spheres: n x {cx, cy, cz, r}; or shapes: shape records (sphere(), rect())."""
        L = lib()
        self.kind = kind
        n = np.array(list(n) + [1] * (3 - len(n)), dtype=np.intc)
        lo = np.array(lo, dtype=np.float64)
        hi = np.array(hi, dtype=np.float64)
        ob = np.array(list(obj_bc) + [0], dtype=np.intc)
        gb = np.array(grid_bc, dtype=np.intc)
        pr = np.array(params, dtype=np.float64)
        if shapes is not None:
            sp = _shapes8(shapes)
            self.p = L.orc_system_create8(kind, _arr(n, C.c_int), _arr(lo, C.c_double), _arr(hi, C.c_double),
                                          len(sp), _arr(sp, C.c_double), _arr(ob, C.c_int), _arr(gb, C.c_int),
                                          C.byref(st), _arr(pr, C.c_double), len(pr))
        else:
            sp = np.ascontiguousarray(np.array(spheres, dtype=np.float64).reshape(-1, 4))
            self.p = L.orc_system_create(kind, _arr(n, C.c_int), _arr(lo, C.c_double), _arr(hi, C.c_double),
                                         len(sp), _arr(sp, C.c_double), _arr(ob, C.c_int), _arr(gb, C.c_int),
                                         C.byref(st), _arr(pr, C.c_double), len(pr))
        self._keep = (n, lo, hi, sp, ob, gb, pr, st)
        if not self.p:
            raise ValueError("oracle rejected the configuration")
        self.mesh = OracleMesh.borrowed(L.orc_system_mesh(self.p))

    def __del__(self):
        if getattr(self, "p", None):
            lib().orc_system_destroy(self.p)
            self.p = None

    def field(self) -> OracleField:
        m = self.mesh.m
        return OracleField(self.mesh.N, [m.nR[0], m.nR[1], m.nR[2]])

    def deriv(self, d):
        return lib().orc_system_deriv(self.p, d)

    def set_mms(self, fns: MmsFns):
        """Functions of a MMS_HOST manufactured solution (kept alive here)."""
        self._mms = fns
        lib().orc_system_set_mms(self.p, C.byref(fns))

    def initialize(self, u: OracleField, t: float):
        lib().orc_system_initialize(self.p, u.ref(), t)

    def update_boundary(self, u: OracleField, t: float):
        lib().orc_system_update_boundary(self.p, u.ref(), t)

    def exact(self, t: float) -> OracleField:
        f = self.field()
        lib().orc_system_exact(self.p, f.ref(), t)
        return f

    def rhs(self, u: OracleField, t: float) -> OracleField:
        out = self.field()
        lib().orc_system_rhs(self.p, u.ref(), out.ref(), t)
        return out

    def stats(self, u: OracleField, t: float) -> np.ndarray:
        s = np.zeros(11)
        lib().orc_system_stats(self.p, u.ref(), t, _arr(s, C.c_double))
        return s

    def dt(self, h_cfl=1.0, p_cfl=1.0) -> float:
        return lib().orc_system_dt(self.p, h_cfl, p_cfl)

    def rk4_step(self, u0, out, rk, srhs, t, dt):
        lib().orc_rk4_step(self.p, u0.ref(), out.ref(), rk.ref(), srhs.ref(), t, dt)

    def euler_step(self, u0, out, srhs, t, dt):
        lib().orc_euler_step(self.p, u0.ref(), out.ref(), srhs.ref(), t, dt)

    def gradient(self, u: OracleField):
        outs = [self.field() for _ in range(3)]
        lib().orc_gradient_apply(self.p, u.ref(), outs[0].ref(), outs[1].ref(), outs[2].ref())
        return outs

    def laplacian(self, u: OracleField, nu: OracleField | None = None) -> OracleField:
        out = self.field()
        lib().orc_laplacian_apply(self.p, u.ref(), nu.ref() if nu is not None else None, out.ref())
        return out

    def deriv_apply(self, d: int, u: OracleField, du: OracleField, plus_eq=False):
        lib().orc_deriv_apply(self.deriv(d), u.ref(), None, du.ref(), int(plus_eq))


class OracleEuler:
    """The compressible Euler system on the CPU (oracle/euler.c): five
    conserved fields (rho, rho u, rho v, rho w, E) over the mesh and
    derivative operators of a configuration (bodies Dirichlet in the
    operators' sense: their cut points carry the slip-wall state)."""

    def __init__(self, cfg, gamma=1.4):
        L = lib()
        st = stencil(cfg["scheme"], cfg["alpha"])
        self.sys = OracleSystem(SYS_HEAT, cfg["n"], cfg["lo"], cfg["hi"], cfg["spheres"],
                                [DIRICHLET] * len(cfg["spheres"]), cfg["grid_bc"], st, [1.0, MMS_SINPROD])
        self.mesh = self.sys.mesh
        ds = (C.POINTER(Deriv) * 3)(*[self.sys.deriv(d) for d in range(3)])
        gb = np.array(cfg["grid_bc"], dtype=np.intc)
        self._keep = (ds, gb)
        self.p = L.orc_euler_create(L.orc_system_mesh(self.sys.p), ds, gamma, _arr(gb, C.c_int))

    def __del__(self):
        if getattr(self, "p", None):
            lib().orc_euler_destroy(self.p)
            self.p = None

    def fields(self):
        return [self.sys.field() for _ in range(5)]

    @staticmethod
    def _refs(U):
        return (Field * 5)(*[u.c for u in U])

    def wall(self, U):
        lib().orc_euler_wall(self.p, self._refs(U))

    def rhs(self, U):
        out = self.fields()
        lib().orc_euler_rhs(self.p, self._refs(U), self._refs(out))
        return out

    def rk4_step(self, u0, out, rk, srhs, dt):
        lib().orc_euler_rk4_step(self.p, self._refs(u0), self._refs(out), self._refs(rk), self._refs(srhs), dt)

