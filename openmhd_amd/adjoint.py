"""Discrete adjoint of the step (libomhd's omhd_adjoint_*: 2D and 3D, periodic
and wall boundaries, resistive flux) and its torch.autograd.Function.

The forward pass advances a Solver nsteps steps from W0 keeping the state
every `every` steps; the backward pass maps dL/dW(T) to dL/dW(0) with the
library's hand-written reverse kernels (u2v, limiter, HLLD, glm_ss2, RK
stages), replaying one segment at a time from its checkpoint.  dt and ch are
frozen at the forward pass's values (the CFL dt is not differentiated).  No
counterpart in the reference, which has no derivatives.

    solver = Solver(512, 512, dx, lm_type=2)           # periodic 2D
    roll = Rollout(solver, nsteps=200, every=20)
    WT = roll(W0)                                      # W0: (9, ny, nx) cuda f64, requires_grad
    loss = ((WT[6] - target) ** 2).sum()
    loss.backward()                                    # W0.grad via omhd_adjoint_backward

A resistive solver (Solver(..., resistivity=(eta0, eta1, w))) also takes the
resistivity parameters as a tensor, so they can be fitted the same way:

    eta = torch.tensor([1e-3, 1e-2, 2.0], dtype=torch.float64, requires_grad=True)
    WT = roll(W0, eta=eta)                             # eta.grad via omhd_adjoint_eta_grad
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import DP, check, lib
from .solver import _cptr


class Adjoint:
    """omhd_adjoint over a Solver: forward (with checkpoints) and backward."""

    def __init__(self, solver, nsteps: int, every: int = 20):
        self.solver, self.nsteps, self.every = solver, int(nsteps), int(every)
        a = C.c_void_p()
        check(lib.omhd_adjoint_create(solver.h, self.nsteps, self.every, C.byref(a)), "omhd_adjoint_create")
        self.a = a
        self.dt = self.ch = None

    def forward(self, dt=None, ch=None):
        """Run nsteps from the solver's current state; with dt/ch given they
        are imposed, else the CFL values are used.  Returns (dt, ch)."""
        dts, chs = np.zeros(self.nsteps), np.zeros(self.nsteps)
        if (dt is None) != (ch is None):
            raise ValueError("give both dt and ch, or neither")
        if dt is not None:
            dt = np.ascontiguousarray(dt, dtype=np.float64)
            ch = np.ascontiguousarray(ch, dtype=np.float64)
            if dt.shape != (self.nsteps,) or ch.shape != (self.nsteps,):
                raise ValueError(f"dt, ch: {self.nsteps} values each")
        rc = lib.omhd_adjoint_forward(self.a, dt.ctypes.data_as(DP) if dt is not None else None,
                                      ch.ctypes.data_as(DP) if ch is not None else None, dts.ctypes.data_as(DP),
                                      chs.ctypes.data_as(DP))
        check(rc, "omhd_adjoint_forward")
        self.solver.t += float(dts.sum())
        self.dt, self.ch = dts, chs
        return dts, chs

    def backward(self, lam_T, out=None):
        """dL/dW(0) from dL/dW(T) (both of the solver's state_shape(), numpy or
        torch, host or device)."""
        n = int(np.prod(self.solver.state_shape()))
        if out is None:
            out = np.zeros(self.solver.state_shape())
        check(lib.omhd_adjoint_backward(self.a, _cptr(lam_T, n), _cptr(out, n)), "omhd_adjoint_backward")
        return out

    def eta_grad(self):
        """dL/d(eta0, eta1, eta_w) of the last backward pass (resistive step)."""
        g = np.zeros(3)
        check(lib.omhd_adjoint_eta_grad(self.a, g.ctypes.data_as(DP)), "omhd_adjoint_eta_grad")
        return g

    def close(self):
        if getattr(self, "a", None):
            lib.omhd_adjoint_destroy(self.a)
            self.a = None

    def __del__(self):
        self.close()


def _function():
    import torch

    class RolloutFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, W0, eta, roll, dt, ch):
            s = roll.solver
            if eta is not None:  # (eta0, eta1, w) of the localized resistivity
                s.set_resistivity(*(float(v) for v in eta.detach().cpu().reshape(3)))
            s.set_state(W0.detach().contiguous())  # waits for torch's stream (the copy included)
            adj = Adjoint(s, roll.nsteps, roll.every)
            roll.dt, roll.ch = adj.forward(dt, ch)
            WT = torch.empty(W0.shape, dtype=W0.dtype, device=W0.device)  # dense, whatever W0's strides
            s.get_state(WT)  # synchronises the library's stream
            ctx.adj = adj
            ctx.eta_like = None if eta is None else (eta.dtype, eta.device, eta.shape)
            return WT

        @staticmethod
        def backward(ctx, gWT):
            g = gWT.contiguous()
            out = torch.empty_like(g)
            ctx.adj.backward(g, out)  # waits for torch's stream, returns with the library's drained
            geta = None
            if ctx.eta_like is not None:  # dL/d(eta0, eta1, w) of this backward pass
                dtype, device, shape = ctx.eta_like
                geta = torch.tensor(ctx.adj.eta_grad(), dtype=dtype, device=device).reshape(shape)
            return out, geta, None, None, None

    return RolloutFunction


class Rollout:
    """nsteps steps of `solver` as a differentiable function of the initial
    state W0 (a (9, ny, nx) or (9, nz, ny, nx) float64 CUDA tensor) and, for
    a resistive solver, of eta = (eta0, eta1, w), the parameters of
    eta0 + eta1 exp(-(x^2 + y^2)/w^2) (a float64 tensor of 3 values; the
    solver keeps them after the call).  After a call, .dt and .ch hold the
    step sequence used; pass dt=, ch= to impose one (e.g. to repeat a run with
    the same frozen sequence, as finite-difference checks do).  The sequence
    is not differentiated, including the resistive limit on dt."""

    def __init__(self, solver, nsteps: int, every: int = 20):
        self.solver, self.nsteps, self.every = solver, int(nsteps), int(every)
        self.dt = self.ch = None
        self._fn = _function()

    def __call__(self, W0, dt=None, ch=None, eta=None):
        if eta is not None and eta.numel() != 3:
            raise ValueError("eta: the three values (eta0, eta1, w)")
        return self._fn.apply(W0, eta, self, dt, ch)
