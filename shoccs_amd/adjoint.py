"""Gradients of a final-time misfit through an RK4 solve on the device.

J(u0, coef) = 1/2 ||u(T) - u_obs||^2 over every entry of the state (D, Rx, Ry,
Rz), where u(T) is nsteps fused FP64 RK4 steps of a system from u0 with the
interior stencil coefficients `coef` of its three operators.  The gradient
comes from the discrete adjoint of the step: the forward run keeps a
checkpoint every `every` steps; the backward run replays each segment from
its checkpoint, keeping every step's input and stage inputs
(shx_rk4_fused_keep_async), and runs the step adjoints in reverse over them
(shx_rk4_adjoint_stages_async).  No reference counterpart (the reference has no
adjoint).

`Solve` works on numpy arrays at its boundary; `solve_fn` wraps it as a
torch.autograd.Function of (u0, coef) when torch is importable.
"""
from __future__ import annotations

import numpy as np


class Solve:
    """nsteps fused FP64 RK4 steps of `problem` from t0 with step dt."""

    def __init__(self, problem, t0: float, dt: float, nsteps: int, every: int = 20):
        if nsteps < 0 or every < 1:
            raise ValueError("nsteps >= 0 and every >= 1")
        self.p = problem
        self.t0, self.dt, self.nsteps, self.every = float(t0), float(dt), int(nsteps), int(every)
        self._ckpt = []  # (step, Field)
        self._seg = []  # the backward sweep's segment buffers
        self._work = problem.field()
        # the transposed operators are assembled once, with the solver (as the
        # forward operators are with the problem)
        problem.adjoint_prepare()

    # ---------------------------------------------------------- helpers
    def set_coef(self, coef: np.ndarray) -> None:
        """Interior coefficients of the three operators, shape (3, W); rows of
        absent directions are ignored."""
        coef = np.asarray(coef, dtype=np.float64)
        for d in range(3):
            if self.p.n[d] > 1:
                self.p.set_interior(d, coef[d])

    def coef(self) -> np.ndarray:
        rows = [self.p.device_interior(d) if self.p.n[d] > 1 else None for d in range(3)]
        w = max(len(r) for r in rows if r is not None)
        out = np.zeros((3, w))
        for d, r in enumerate(rows):
            if r is not None:
                out[d] = r
        return out

    def _time(self, step: int) -> float:
        return self.t0 + step * self.dt

    def _free_ckpt(self):
        for _, f in self._ckpt:
            f.free()
        self._ckpt = []

    # ---------------------------------------------------------- forward
    def run(self, u0: np.ndarray) -> np.ndarray:
        """u(T) from u0 (flat D, Rx, Ry, Rz), keeping the checkpoints."""
        self._free_ckpt()
        u = self._work
        u.upload(u0)
        for s in range(0, self.nsteps, self.every):
            c = self.p.field()
            c.copy_from(u)
            self._ckpt.append((s, c))
            n = min(self.every, self.nsteps - s)
            self.p.rk4_fused(u, self._time(s), self.dt, n)
        return u.download()

    def misfit(self, u0: np.ndarray, u_obs: np.ndarray) -> float:
        r = self.run(u0) - u_obs
        return 0.5 * float(r @ r)

    # --------------------------------------------------------- backward
    def gradient(self, uT_bar: np.ndarray, with_coef: bool = True):
        """dJ/du0 and dJ/dcoef (3 x 9, per direction's stored interior
        coefficients) given dJ/du(T), after `run`.  Each segment is replayed
        from its checkpoint keeping every step's input and stage inputs
        (shx_rk4_fused_keep_async), then the step adjoints run in reverse over
        them (shx_rk4_adjoint_stages_async), all on the device; the
        coefficient gradient is read back once."""
        if len(self._ckpt) * self.every < self.nsteps:
            raise RuntimeError("gradient() needs a preceding run()")
        ubar = self.p.field()
        ubar.upload(uT_bar)
        if with_coef:
            self.p.adjoint_coef_set(np.zeros(27))
        seg = self._seg  # per step of a segment: (input, w1, w2, w3), kept for later calls
        try:
            for s0, c in reversed(self._ckpt):
                n = min(self.every, self.nsteps - s0)
                while len(seg) < n:
                    seg.append(tuple(self.p.field() for _ in range(4)))
                seg[0][0].copy_from(c)
                for k in range(n):
                    u, w1, w2, w3 = seg[k]
                    nxt = seg[k + 1][0] if k + 1 < n else self._work
                    self.p.rk4_fused_keep(u, w1, w2, w3, nxt, self._time(s0 + k), self.dt)
                for k in range(n - 1, -1, -1):
                    self.p.rk4_adjoint_stages(seg[k], ubar, self.dt, with_coef)
            g_u0 = ubar.download()
            grad = self.p.adjoint_coef_get().reshape(3, 9) if with_coef else None
        finally:
            ubar.free()
        return g_u0, grad

    def close(self):
        self._free_ckpt()
        for fs in self._seg:
            for f in fs:
                f.free()
        self._seg = []
        self._work.free()


def solve_fn(solver: Solve):
    """torch.autograd.Function of (u0, coef) -> u(T) for `solver` (float64
    tensors; coef has shape (3, W)).  The backward pass runs the discrete
    adjoint; tensors cross to the device library through host memory."""
    import torch

    class _RK4Solve(torch.autograd.Function):
        @staticmethod
        def forward(ctx, u0, coef):
            solver.set_coef(coef.detach().cpu().numpy())
            uT = solver.run(u0.detach().cpu().numpy().astype(np.float64))
            ctx.coef_width = coef.shape[1]
            return torch.from_numpy(uT).to(u0.device)

        @staticmethod
        def backward(ctx, gT):
            g_u0, g_coef = solver.gradient(gT.detach().cpu().numpy().astype(np.float64))
            w = ctx.coef_width
            return (torch.from_numpy(g_u0).to(gT.device),
                    torch.from_numpy(np.ascontiguousarray(g_coef[:, :w])).to(gT.device))

    return _RK4Solve.apply
