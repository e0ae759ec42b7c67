"""Spectral diagnostics of the derivative operators: the dense operator of one
direction over its unknowns and its eigenvalues.

Mirrors the reference's eigenvalue_visitor (src/operators/eigenvalue_visitor.cpp:10-52)
and the hyperbolic_eigenvalues system's statistic
(src/systems/hyperbolic_eigenvalues.cpp:78-89).  The unknowns are the points
that carry a row of the operator (unit_stride_visitor.cpp:49-140): the mesh
points of its line segments (a Dirichlet wall or body point has neither a row
nor a column) and the cut points of non-Dirichlet bodies in R(dir), in that
order.  Column j of the matrix is the device operator applied to the unit
vector of unknown j (one apply per column, so the matrix is exactly what the
product computes); the eigenvalues come from LAPACK's general solver through
numpy, as the reference calls geev.
"""
from __future__ import annotations

import numpy as np


def unknowns(problem, d: int) -> tuple[np.ndarray, np.ndarray]:
    """(mesh points with a row of direction d, cut points of R(d) with a row):
    flat D indices and R(d) indices, ascending."""
    n = problem.n
    blocks = problem.blocks(d)
    slow, fast = (1, 2) if d == 0 else (0, 2) if d == 1 else (0, 1)
    stride = [n[1] * n[2], n[2], 1]
    rows = []
    for b in blocks:
        s_, f_ = divmod(int(b.pair), n[fast])
        base = s_ * stride[slow] + f_ * stride[fast]
        total = int(b.nl) + int(b.ni) + int(b.nr)
        rows.extend(base + (int(b.lo) + q) * stride[d] for q in range(total))
    cut = problem.cut_rows(d)
    rr = np.unique(cut["rrow"][cut["rset"] == d])
    return np.unique(np.asarray(rows, dtype=np.int64)), rr.astype(np.int64)


def operator_matrix(problem, d: int) -> np.ndarray:
    """Dense matrix of direction d's derivative over its unknowns (rows and
    columns in the order of `unknowns`)."""
    D, R = unknowns(problem, d)
    N = problem.N
    off = N + sum(problem.nR[:d])
    pos = np.concatenate([D, off + R])
    n = pos.size
    A = np.zeros((n, n))
    apply = problem.derivative(d)
    u, du = problem.field(), problem.field()
    e = np.zeros(problem.flat_size)
    zero = np.zeros(problem.flat_size)
    for j in range(n):
        e[pos[j]] = 1.0
        u.upload(e)
        du.upload(zero)
        apply(u, du, 0)
        A[:, j] = du.download()[pos]
        e[pos[j]] = 0.0
    u.free()
    du.free()
    return A


def eigenvalues(problem, d: int) -> np.ndarray:
    """Eigenvalues of direction d's derivative over its unknowns."""
    return np.linalg.eigvals(operator_matrix(problem, d))


def hyperbolic_eigenvalue_stat(problem) -> float:
    """hyperbolic_eigenvalues::stats (src/systems/hyperbolic_eigenvalues.cpp:78-89)
    of a one-dimensional problem: -h_x times the least real part of the x
    derivative's eigenvalues (at most a small positive number for a scheme
    whose semi-discretisation of u_t + u_x = 0 is stable)."""
    if problem.n[1] != 1 or problem.n[2] != 1:
        raise ValueError("the eigenvalue statistic is defined for one-dimensional meshes")
    lam = eigenvalues(problem, 0)
    return float(-problem.h()[0] * np.min(lam.real))
