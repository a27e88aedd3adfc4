"""Warm-started Lanczos eigensolver on the GPU: the counterpart of utils/eigsh.py:10-17.

The reference calls CuPy's thick-restart ``eigsh(S_rho, k=1, which='SA', tol, maxiter, v0, ncv)``
(models/spectral_annealing.py:671-677), which restarts from the Ritz vector every ``ncv`` steps and
reorthogonalises each new vector against the whole basis. This solver keeps ``keep`` Ritz vectors
per thick restart (keep = 1 is CuPy's restart) and differs in four deliberate ways:

* the recurrence runs on N_alpha (the reference's S_rho = rho I - N has the same Krylov space and
  T_S = rho I - T_N), which keeps the FP32 basis orthogonal; mu = rho - theta is reported;
* the stopping test |beta_j s_j| <= tol is applied after every step, not once per restart;
* ``iterations`` counts operator applications (CuPy's counter is ncv + (ncv - 1) r);
* reorthogonalisation is selective: the projections on every basis row are measured each step, and
  the update subtracts all of them only when one on a row older than j - 1 exceeds orth_tau * beta_j
  (orth_tau = sqrt(eps) / 2: a semi-orthogonal basis, whose Ritz values are as accurate as with full
  reorthogonalisation); orth_tau = 0 restores the full reorthogonalisation of every step.

Every step is a fused scaled SpMV, the projections and the Gram-Schmidt update with the FP64 Ritz
solve; all launches read the solve's status on the device and do nothing once it stopped, so the
host queues several steps between reads of the status.
"""
from __future__ import annotations

import math

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import (SA_BROKEN, SA_CONVERGED, SA_MAXITER, SA_REJECTED, SA_RESTART, SA_RUNNING, EigensolverBreakdown,
                   call)
from .device import DeviceProblem, ptr, stream

TERMINAL = (SA_CONVERGED, SA_MAXITER, SA_BROKEN, SA_REJECTED)


@dataclass
class LanczosResult:
    status: int
    theta: float          # largest Ritz value of N
    res: float            # residual estimate |beta s_last|
    steps: int            # operator applications
    cycles: int           # restarts
    first_beta: float     # beta_0 of a gated (predicted) start: the predictor residual
    start_norm: float


GRAPH_MAX_N = 1 << 18
# Thick-restart default with ncv = 33 (sweeps of 128 alphas): keep 12 takes 4.57 s on ER 2^22 and
# 1.83 s on BA 2^20 against 4.65 s and 1.85 s with 16 (fewer rows per projection, about the same
# operator applications); keep 10 needed 10% more applications.
DEFAULT_KEEP = 12


# The half-width shadow copy of the basis (FP16 for FP32 vectors, FP32 for FP64 vectors) rounds the
# projection coefficients on old rows to ~eps_shadow / sqrt(n) of ||W||; through the occasional full
# Gram-Schmidt update that error enters the Lanczos relation. FP64 solves asked for a tolerance below
# this bound therefore run without the shadow (measured on Barabasi-Albert n = 1500 at tol 1e-10: true
# residuals up to 2.4e-9 with the FP32 shadow, against estimates below 1e-10).
SHADOW_MIN_TOL = {torch.float32: 0.0, torch.float64: 1e-8}
# The rounding of a delocalised unit row in FP16 moves a projection by about 2.8e-4 ||W|| / sqrt(n): below
# 1.1e-6 ||W|| from n = 2^16 on, but 8.7e-6 ||W|| at n = 1024, where FP32 solves then ended with true
# residuals up to 1.6e-5 against estimates below their 1e-6 tolerance. Small problems gain nothing
# from the narrower reads either, so they keep the full-precision basis only.
SHADOW_MIN_N = 1 << 16


def use_shadow(dtype: torch.dtype, tol: float, n: int | None = None) -> bool:
    """Whether solves of this precision, tolerance and size read the half-width shadow basis."""
    return tol >= SHADOW_MIN_TOL[dtype] and (n is None or n >= SHADOW_MIN_N)


def default_orth_tau(dtype: torch.dtype) -> float:
    """sqrt(eps) / 2 of the vectors' type: the semi-orthogonality level the selective
    reorthogonalisation keeps."""
    return 0.5 * math.sqrt(torch.finfo(dtype).eps)


def default_ncv(n: int, ncv: int | None) -> int:
    """CuPy's default for k = 1 is min(max(2k, k + 32), n - 1) = min(33, n - 1)."""
    v = min(33, n - 1) if ncv is None else int(ncv)
    return max(2, min(v, _lib.SA_MAX_NCV, n))


class LanczosSolver:
    """Buffers and launch sequence of the warm-started Lanczos for one DeviceProblem."""

    def __init__(self, prob: DeviceProblem, ncv: int | None = None, chunk: int = 8, keep: int | None = None,
                 graphs: bool | None = None, orth_tau: float | None = None, shadow: bool = True):
        """orth_tau: selective reorthogonalisation threshold (None: sqrt(eps) / 2 of the vectors' type;
        0: reorthogonalise against the whole basis at every step). shadow: keep a half-width copy of the
        basis for the projections on rows older than j - 1."""
        self.prob = prob
        self.orth_tau = default_orth_tau(prob.dtype) if orth_tau is None else float(orth_tau)
        self.n = prob.n
        self.ncv = default_ncv(prob.n, ncv)
        self.chunk = max(1, int(chunk))
        # Ritz vectors a thick restart keeps: half the basis, at most DEFAULT_KEEP; 1 = restart from the
        # Ritz vector alone, the reference eigsh's k = 1 thick restart.
        want = min(DEFAULT_KEEP, self.ncv // 2) if keep is None else int(keep)
        self.keep = max(1, min(_lib.SA_MAX_KEEP, want, self.ncv - 1))
        dev, dt = prob.device, prob.dtype
        # Basis rows padded so that consecutive rows do not start at the same HBM channel offset
        # (n = 2^k would put every row start in one channel group); V is the n-wide view.
        self.ldv = _lib.padded_ld(self.n, 4 if dt == torch.float32 else 8)
        self.Vbuf = torch.empty((self.ncv + 1, self.ldv), dtype=dt, device=dev)
        self.V = self.Vbuf[:, : self.n]
        self.Vh = (torch.empty((self.ncv + 1, self.ldv), dtype=torch.float16 if dt == torch.float32 else torch.float32,
                               device=dev) if shadow else None)
        self.Z = torch.empty(self.n, dtype=dt, device=dev)
        self.W = torch.empty(self.n, dtype=dt, device=dev)
        self.Y = torch.empty(self.n, dtype=dt, device=dev)
        self.sfx = prob.sfx
        self.work = prob.work
        # Small problems are launch-bound: a chunk of steps is then captured once as a HIP graph and
        # replayed (every launch reads the solve's state on the device, so one graph serves every
        # solve on these buffers).
        self.use_graphs = (self.n <= GRAPH_MAX_N) if graphs is None else bool(graphs)
        self._graph = None
        self._graph_key = None

    # -- launch helpers ------------------------------------------------------------------------
    def begin(self, tol: float, maxiter: int, bound: float, pred_tol: float = float("inf")) -> None:
        call("sa_lanczos_begin", self.work.ptr, self.ncv, int(maxiter), float(tol), float(bound), float(pred_tol),
             self.keep, self.orth_tau, stream())

    def start(self, s: torch.Tensor | None, x: torch.Tensor, *, d: torch.Tensor | None = None, pexp: float = 0.0,
              cx: float = 1.0, y: torch.Tensor | None = None, cy: float = 0.0, gate: int = 0,
              only_if: int = -1) -> None:
        call(f"sa_lanczos_start_{self.sfx}", self.work.ptr, self.n, ptr(s), ptr(x), ptr(d), float(pexp), float(cx),
             ptr(y), float(cy), ptr(self.V), self.ldv, ptr(self.Vh), ptr(self.Z), int(gate), int(only_if), stream())

    def _step(self, s: torch.Tensor | None) -> None:
        call(f"sa_lanczos_step_{self.sfx}", self.prob.csr_ref, self.work.ptr, ptr(s), ptr(self.V), self.ldv,
             ptr(self.Vh), ptr(self.Z), ptr(self.W), stream())

    def thick_restart(self, s: torch.Tensor | None) -> None:
        call(f"sa_lanczos_thick_restart_{self.sfx}", self.work.ptr, self.n, ptr(s), ptr(self.V), self.ldv,
             ptr(self.Vh), ptr(self.Z), stream())

    def ritz(self, y_last: torch.Tensor | None = None, only_if: int = -1) -> None:
        call(f"sa_lanczos_ritz_{self.sfx}", self.work.ptr, self.n, ptr(self.V), self.ldv, ptr(y_last), ptr(self.Y),
             int(only_if), stream())

    def _group(self, s: torch.Tensor | None) -> None:
        """One step and the thick-restart launches after it (no-ops unless that step filled the basis)."""
        self._step(s)
        self.thick_restart(s)

    def _queue(self, s: torch.Tensor | None, count: int) -> None:
        if self.use_graphs and count == self.chunk:
            c = self.prob.csr
            # the captured launches carry the layout's pointers: any layout change needs a new graph
            key = (None if s is None else s.data_ptr(), c.tile_rows, c.npanels, c.tile_nnz, c.bmode, c.bent,
                   c.pbent, c.partial)
            if self._graph is None or self._graph_key != key:
                g = torch.cuda.CUDAGraph()
                torch.cuda.synchronize()
                with torch.cuda.graph(g):
                    for _ in range(count):
                        self._group(s)
                self._graph, self._graph_key = g, key
            self._graph.replay()
            return
        # eager launches: the host tracks j and queues the restart pair only when the basis fills
        for _ in range(count):
            self._step(s)
            self._j += 1
            if self._j >= self.ncv:
                self.thick_restart(s)
                self._j = self.keep

    # -- the solve ------------------------------------------------------------------------------
    def run(self, s: torch.Tensor | None, maxiter: int, first_check: int | None = None) -> LanczosResult:
        """Iterate from the started vector until converged, stopped at maxiter, rejected or broken.

        The device stops the solve by itself at maxiter operator applications; the host only queues
        chunks of steps and reads the state between them. It gives up when a whole chunk (plus a
        restart) made no progress on the device's step counter, so launches the device skipped (for
        example after a thick restart that kept fewer vectors than the host assumed) never count
        against the budget."""
        check = first_check or self.chunk
        first_beta = None
        self._j = 0
        last_steps, idle = -1, 0
        while True:
            self._queue(s, check)
            st = self.work.head()
            if first_beta is None:
                first_beta = st.first_beta
            if st.status in TERMINAL:
                return LanczosResult(st.status, st.theta, st.res, st.steps, st.cycles, first_beta, st.r[3])
            if st.status == SA_RESTART:
                # the basis filled earlier than the host expected (the restart kept fewer vectors)
                self.thick_restart(s)
                st = self.work.head()
            self._j = st.j
            idle = idle + 1 if st.steps == last_steps else 0
            last_steps = st.steps
            if idle >= 2:
                raise EigensolverBreakdown(
                    f"Lanczos made no progress over two chunks (status {st.status}, {st.steps} steps)")
            check = self.chunk

    def final_vector(self, y_last: torch.Tensor | None) -> tuple[float, float]:
        """Ritz vector of the last cycle into Y; returns (||Y||, Y . y_last)."""
        self.ritz(y_last, only_if=-1)
        st = self.work.head()
        return float(st.r[4]), float(st.r[5])


class ScaledOperator:
    """S_rho = rho I - N_alpha (or N_alpha itself with rho = None), N = s (.) (J/c) (.) s.

    The counterpart of SRhoOperator / NAlphaOperator (models/spectral_annealing.py:29-53).
    """

    def __init__(self, prob: DeviceProblem, s: torch.Tensor | None, rho: float | None):
        self.prob, self.s, self.rho = prob, s, rho
        self.shape = (prob.n, prob.n)
        self.dtype = prob.dtype
        self._z = torch.empty(prob.n, dtype=prob.dtype, device=prob.device)

    def matvec(self, x: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        out = torch.empty_like(x) if out is None else out
        if self.rho is None:
            return self.prob.matvec(self.s, x, 0.0, 1.0, out, self._z)
        return self.prob.matvec(self.s, x, self.rho, -1.0, out, self._z)

    __call__ = matvec


def eigsh(A: ScaledOperator, k: int = 1, which: str = "SA", tol: float = 0.0, maxiter: int | None = None,
          v0: torch.Tensor | None = None, ncv: int | None = None, seed: int = 0):
    """Leading eigenpair on the GPU with the reference wrapper's return convention (w, v, iterations).

    For S_rho (A.rho set) ``which='SA'`` gives mu = rho - lambda_max(N); for N (A.rho None) ``which='LA'``
    gives lambda_max(N). v is an (n, 1) device tensor of unit norm; v0 = None starts from a seeded
    normal vector.
    """
    if k != 1:
        raise ValueError("only k = 1 is supported (the reference calls eigsh with k = 1)")
    expect = "SA" if A.rho is not None else "LA"
    if which != expect:
        raise ValueError(f"which must be {expect!r} for this operator")
    prob = A.prob
    maxiter = int(maxiter) if maxiter is not None else 10 * prob.n
    tol = float(tol) if tol and tol > 0 else float(np.finfo(np.float32 if prob.dtype == torch.float32
                                                            else np.float64).eps)
    solver = LanczosSolver(prob, ncv, shadow=use_shadow(prob.dtype, tol, prob.n))
    x = v0 if v0 is not None else prob.fill_normal(seed, 0, torch.empty(prob.n, dtype=prob.dtype,
                                                                        device=prob.device))
    rho = A.rho if A.rho is not None else 0.0
    solver.begin(tol, maxiter, bound=abs(rho) + 2.0)
    solver.start(A.s, x)
    res = solver.run(A.s, maxiter)
    if res.status == SA_BROKEN:
        raise EigensolverBreakdown("Lanczos breakdown: non-finite or oversized tridiagonal entry")
    nrm, _ = solver.final_vector(None)
    v = (solver.Y / nrm).reshape(-1, 1)
    w = np.array([rho - res.theta if A.rho is not None else res.theta])
    return w, v, res.steps
