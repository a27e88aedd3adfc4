"""Eigenvector tangent dy/dalpha of the alpha sweep and its adjoint, by deflated MINRES on the GPU.

No counterpart in the reference: a third predictor beside the diagonal and secant ones of
utils/predictors.py:6-56. For the unit leading eigenvector y of N(alpha) = s (.) (J/c) (.) s,
s = d^(-alpha/2), with eigenvalue lambda:

    dN/dalpha = -1/2 (L N + N L),  L = diag(ln d);   u = 1/2 L y,  P = I - y y^T
    dy/dalpha = P u + 2 lambda z,   z = (N - lambda I)^+ P u
    dlambda/dalpha = y^T N' y = -2 lambda (y . u)

(differentiate N y = lambda y; N' y = -(lambda u + N u) and (N - lambda I)^+ (N - lambda I) = P). The
first term P u is the first-order part of the diagonal predictor D^(da/2) y, so the tangent
predictor is the diagonal predictor plus the correction 2 lambda da z, with z from a few MINRES
iterations (sa_minres_*; every vector op and scalar on the device).

The adjoint of a cotangent g on y is g . dy/dalpha = g . P u + 2 lambda (a . u) with
a = (N - lambda I)^+ P g: the same deflated MINRES with right-hand side g. ``LeadingEigvec`` exposes
alpha -> (y, lambda) as a torch.autograd.Function whose backward is that adjoint solve.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import SA_BROKEN, SA_CONVERGED, SA_RUNNING, EigensolverBreakdown, SaMinresBufs, SaMinresState, call
from .device import DeviceProblem, Work, ptr, stream
from .eigsh import LanczosSolver, use_shadow

CLAMP = 1e-12


@dataclass
class MinresResult:
    status: int
    iterations: int
    beta1: float        # ||P r||
    rnorm: float        # ||P r - A x|| (MINRES's phibar)


class DeflatedMinres:
    """Device buffers and launch sequence of one deflated MINRES solve on a DeviceProblem."""

    def __init__(self, prob: DeviceProblem, chunk: int = 4):
        self.prob = prob
        self.n = prob.n
        self.chunk = max(1, int(chunk))
        dev, dt = prob.device, prob.dtype
        self.r = [torch.empty(self.n, dtype=dt, device=dev) for _ in range(2)]
        self.w = [torch.empty(self.n, dtype=dt, device=dev) for _ in range(3)]
        self.x = torch.empty(self.n, dtype=dt, device=dev)
        self.z = torch.empty(self.n, dtype=dt, device=dev)
        self.q = torch.empty(self.n, dtype=dt, device=dev)
        self.bufs = SaMinresBufs()
        self.bufs.r[0], self.bufs.r[1] = (t.data_ptr() for t in self.r)
        for i in range(3):
            self.bufs.w[i] = self.w[i].data_ptr()
        self.bufs.x, self.bufs.z, self.bufs.q = self.x.data_ptr(), self.z.data_ptr(), self.q.data_ptr()
        self.work = Work(dev)
        self._host = torch.empty(ctypes.sizeof(SaMinresState), dtype=torch.uint8, pin_memory=True)

    def state(self) -> SaMinresState:
        nbytes = ctypes.sizeof(SaMinresState)
        self._host.copy_(self.work.buf[:nbytes], non_blocking=True)
        torch.cuda.current_stream().synchronize()
        st = SaMinresState()
        ctypes.memmove(ctypes.addressof(st), self._host.data_ptr(), nbytes)
        return st

    def solve(self, s: torch.Tensor | None, y: torch.Tensor, r: torch.Tensor, shift: float, tol: float,
              maxit: int) -> MinresResult:
        """x = (N - shift I)^+ P r restricted to y's complement (y unit), into self.x."""
        sfx = self.prob.sfx
        bufs = ctypes.byref(self.bufs)
        call(f"sa_minres_begin_{sfx}", self.work.ptr, self.n, ptr(s), ptr(y), ptr(r), float(shift), float(tol),
             int(maxit), bufs, stream())
        done = 0
        while done < maxit:
            k = min(self.chunk, maxit - done)
            call(f"sa_minres_iterate_{sfx}", self.prob.csr_ref, self.work.ptr, ptr(s), ptr(y), k, bufs, stream())
            done += k
            if self.state().status != SA_RUNNING:
                break
        call(f"sa_minres_end_{sfx}", self.work.ptr, self.n, ptr(y), bufs, stream())
        st = self.state()
        if st.status == SA_BROKEN:
            raise EigensolverBreakdown("MINRES breakdown: non-finite recurrence")
        return MinresResult(int(st.status), int(st.itn), float(st.beta1), float(st.rnorm))


class TangentPredictor:
    """z = (N - lambda I)^+ P u for the sweep's tangent predictor (u = 1/2 ln d (.) y)."""

    def __init__(self, prob: DeviceProblem, steps: int = 8, tol: float = 1e-2):
        self.prob = prob
        self.steps = int(steps)
        self.tol = float(tol)
        self.minres = DeflatedMinres(prob)
        self.u = torch.empty(prob.n, dtype=prob.dtype, device=prob.device)
        self.z = torch.empty_like(self.u)
        self.z0 = torch.empty_like(self.u)
        self.w = torch.empty_like(self.u)
        self.scratch = torch.empty_like(self.u)
        self.lam = None
        self.last: MinresResult | None = None
        self.matvecs = 0

    def rhs(self, y: torch.Tensor) -> torch.Tensor:
        call(f"sa_tangent_rhs_{self.prob.sfx}", self.prob.n, ptr(y), ptr(self.prob.d), CLAMP, ptr(self.u), stream())
        return self.u

    def compute(self, s: torch.Tensor | None, y: torch.Tensor, lam: float,
                history: tuple[torch.Tensor, float] | None = None) -> MinresResult:
        """Tangent correction at (alpha, y, lambda): self.z = (N - lambda I)^+ P u (y unit, s = s(alpha)).

        With ``history = (y_prev, dalpha)`` (the previous alpha's unit eigenvector, sign-aligned with y,
        and alpha - alpha_prev), MINRES starts from the secant estimate of z,
        z0 = ((y - y_prev) / dalpha - P u) / (2 lambda): the backward difference carries the components
        along the eigenvectors next to lambda, which a few Krylov steps cannot resolve, and MINRES
        corrects the rest (one extra operator application for the initial residual)."""
        prob = self.prob
        self.rhs(y)
        self.lam = float(lam)
        if history is None or not (history[1] > 0) or lam == 0.0:
            res = self.minres.solve(s, y, self.u, lam, self.tol, self.steps)
            self.z.copy_(self.minres.x)
            self.matvecs = res.iterations
            self.last = res
            return res
        y_prev, dprev = history
        _, yu = prob.norm_dot(y, self.u)
        inv = 1.0 / (2.0 * lam)
        prob.combine3((1.0 / dprev + yu) * inv, y, -inv / dprev, y_prev, -inv, self.u, self.z0)
        # residual of the derivative equation at z0: r = u - (N - lambda I) z0 (deflated inside MINRES)
        prob.matvec(s, self.z0, -lam, 1.0, self.w, self.scratch)
        prob.combine3(1.0, self.u, -1.0, self.w, 0.0, None, self.w)
        res = self.minres.solve(s, y, self.w, lam, self.tol, self.steps)
        prob.combine3(1.0, self.z0, 1.0, self.minres.x, 0.0, None, self.z)
        self.matvecs = res.iterations + 1
        self.last = res
        return res

    def full_tangent(self, y: torch.Tensor) -> torch.Tensor:
        """dy/dalpha = P u + 2 lambda z (after compute)."""
        u = self.u.double()
        yd = y.double()
        pu = u - (yd @ u) * yd
        return pu + 2.0 * self.lam * self.z.double()


# ---------------------------------------------------------------------------------------------
# alpha -> leading eigenpair as a differentiable function
# ---------------------------------------------------------------------------------------------
class EigenContext:
    """Everything LeadingEigvec needs for one problem: the Lanczos solver, the scaling vector and
    the MINRES buffers of the adjoint solve."""

    def __init__(self, prob: DeviceProblem, tol: float | None = None, maxiter: int = 5000, ncv: int | None = None,
                 adjoint_steps: int = 200, adjoint_tol: float = 1e-10, seed: int = 0):
        self.prob = prob
        eps = np.finfo(np.float32 if prob.dtype == torch.float32 else np.float64).eps
        self.tol = float(tol) if tol else float(np.sqrt(eps) * 1e-2 if prob.dtype == torch.float32 else 1e-12)
        self.lz = LanczosSolver(prob, ncv, shadow=use_shadow(prob.dtype, self.tol, prob.n))
        self.s = torch.empty(prob.n, dtype=prob.dtype, device=prob.device)
        self.x0 = torch.empty_like(self.s)
        self.tangent = TangentPredictor(prob, adjoint_steps, adjoint_tol)
        self.maxiter = int(maxiter)
        self.seed = int(seed)
        self.alpha = None

    def solve(self, alpha: float, sign_ref: torch.Tensor | None = None):
        """Unit leading eigenvector y of N(alpha) (sign: y . sign_ref >= 0, else its largest entry > 0)
        and lambda; leaves s = s(alpha) in self.s."""
        prob = self.prob
        _, g = prob.alpha_scaling(alpha, CLAMP, self.s)
        prob.fill_normal(self.seed, 0, self.x0)
        self.lz.begin(self.tol, self.maxiter, bound=2.0 * g + 2.0)
        self.lz.start(self.s, self.x0)
        res = self.lz.run(self.s, self.maxiter)
        if res.status == SA_BROKEN:
            raise EigensolverBreakdown("Lanczos breakdown")
        nrm, _ = self.lz.final_vector(None)
        y = self.lz.Y.double() / nrm
        if sign_ref is not None:
            flip = float(y @ sign_ref.double()) < 0
        else:
            flip = float(y[torch.argmax(torch.abs(y))]) < 0
        if flip:
            y = -y
        self.alpha = float(alpha)
        return y.to(prob.dtype), float(res.theta), res


class LeadingEigvec(torch.autograd.Function):
    """(y, lambda) = leading unit eigenvector and eigenvalue of N(alpha), differentiable in alpha.

    forward(alpha, ctx_obj, sign_ref) with alpha a 0-dim float64 tensor; backward solves the deflated
    adjoint system (N - lambda I) a = P g on the GPU and returns
    dL/dalpha = g . P u + 2 lambda (a . u) + g_lambda (-2 lambda y . u).
    """

    @staticmethod
    def forward(ctx, alpha: torch.Tensor, eig: EigenContext, sign_ref: torch.Tensor | None = None):
        y, lam, _ = eig.solve(float(alpha), sign_ref)
        ctx.eig = eig
        ctx.alpha = float(alpha)
        ctx.lam = lam
        ctx.alpha_dtype = alpha.dtype
        ctx.save_for_backward(y)
        return y, torch.tensor(lam, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gy, glam):
        (y,) = ctx.saved_tensors
        eig: EigenContext = ctx.eig
        prob = eig.prob
        lam = ctx.lam
        # s(alpha) may have been overwritten by a later forward call: recompute it
        prob.alpha_scaling(ctx.alpha, CLAMP, eig.s)
        tp = eig.tangent
        u = tp.rhs(y).double()
        yd = y.double()
        yu = float(yd @ u)
        grad = torch.zeros((), dtype=torch.float64)
        if gy is not None:
            g = gy.to(prob.dtype).contiguous()
            tp.minres.solve(eig.s, y, g, lam, tp.tol, tp.steps)
            a = tp.minres.x.double()
            gd = gy.double()
            grad = grad + float(gd @ u - (gd @ yd) * yu) + 2.0 * lam * float(a @ u)
        if glam is not None:
            grad = grad + float(glam) * (-2.0 * lam * yu)
        return grad.to(ctx.alpha_dtype), None, None
