"""SpecAdj (alpha = 0) and SpecSignedLap (alpha = 1): the counterparts of models/spec_adj.py and
models/spec_signed_lap.py, run on the same kernels as SpecAnn.

SpecAdj takes the smallest eigenpair of S = I - J/c (spec_adj.py:156-171); SpecSignedLap the
smallest of S = D^-1/2 (D - J) D^-1/2 with d clamped at 1e-7 (spec_signed_lap.py:183-206). Both are
I - N with N = s (.) (J/c) (.) s, s = 1 or s = max(d, 1e-7)^(-1/2), so the Lanczos runs on N from a
seeded random start, mu = 1 - theta is reported, and a non-finite result is retried once
(_eigsh_finite, spec_adj.py:351-363). The rounded vector is refined by the batched local search and
kept if that lowers the energy (spec_adj.py:122-142).
"""
from __future__ import annotations

import time
from datetime import datetime

import numpy as np
import torch

from . import guards
from . import local_search as LS
from ._lib import SA_BROKEN, SA_CONVERGED, EigensolverBreakdown, call
from .device import DeviceProblem, ptr, stream
from .eigsh import LanczosSolver
from .failures import NumericalFailure
from .results import store_general
from .spectral_annealing import LS_POLICIES, unpack_bits


class _Endpoint:
    ALPHA = 0.0
    CLAMP = 1e-12
    DEFAULT_RUN = "endpoint"

    def __init__(self, instance_name: str, dataset: str, random_seed: int = None, mode: int = 0,
                 beta: float = None, use_local_search: bool = True, max_local_iterations: int = 1000,
                 eigsh_tol: float = 1e-6, eigsh_maxiter: int = 1000, run_name: str = None,
                 ls_policy: str = "default", precision: int = 32, verbose: bool = True, **kwargs):
        if ls_policy not in LS_POLICIES:
            raise ValueError(f"ls_policy must be one of {LS_POLICIES}, got {ls_policy!r}")
        if precision not in (32, 64):
            raise ValueError(f"precision must be 32 or 64, got {precision}")
        self.instance_name, self.dataset, self.random_seed = instance_name, dataset, random_seed
        self.mode, self.beta = mode, beta
        self.use_local_search = use_local_search
        self.max_local_iterations = max_local_iterations
        self.eigsh_tol, self.eigsh_maxiter = float(eigsh_tol), int(eigsh_maxiter)
        self.ls_policy = ls_policy
        self.precision = int(precision)
        self.dtype = np.float32 if self.precision == 32 else np.float64
        self.tdtype = torch.float32 if self.precision == 32 else torch.float64
        self.solver_name = run_name if run_name else self.DEFAULT_RUN
        self.verbose = verbose
        if verbose:
            print(f"Initialized {self.solver_name} solver (GPU sparse). Results will be saved under "
                  f"'{self.solver_name}'.")
            if kwargs:
                print(f"  additional kwargs: {kwargs}")

    def _seed(self) -> int:
        return 0 if self.random_seed is None else int(self.random_seed)

    def solve(self, J):
        """Solve the instance with couplings J. Returns (energy, spins, time in seconds)."""
        start = time.time()
        prob = DeviceProblem(J, self.tdtype)
        n = prob.n
        s = torch.empty(n, dtype=self.tdtype, device=prob.device)
        prob.alpha_scaling(self.ALPHA, self.CLAMP, s)
        s_arg = None if self.ALPHA == 0.0 else s
        lz = LanczosSolver(prob, None)
        x = torch.empty(n, dtype=self.tdtype, device=prob.device)
        res = None
        for attempt in (1, 2):
            prob.fill_normal(self._seed(), 1000 + attempt, x)
            lz.begin(self.eigsh_tol, self.eigsh_maxiter, bound=3.0)
            lz.start(s_arg, x)
            res = lz.run(s_arg, self.eigsh_maxiter)
            if res.status != SA_BROKEN and np.isfinite(res.theta):
                break
            if self.verbose:
                print(f"Warning: eigsh returned a non-finite eigenpair (try {attempt} of 2)")
        else:
            raise NumericalFailure("eigsh returned a non-finite eigenpair twice")
        self._eigsh_iterations = res.steps
        nrm, _ = lz.final_vector(None)
        self._eigsh_converged = res.status == SA_CONVERGED
        if not self._eigsh_converged:
            w = torch.empty_like(lz.Y)
            prob.matvec(s_arg, lz.Y, 0.0, 1.0, w, lz.W)
            r, _ = prob.norm_dot(w, lz.Y, 1.0, -res.theta)
            self._eigsh_converged = (r / nrm) <= self.eigsh_tol
        nwords = (n + 63) // 64
        bits = torch.empty((1, nwords), dtype=torch.int64, device=prob.device)
        call(f"sa_round_{prob.sfx}", prob.work.ptr, n, ptr(lz.Y), 1.0 / nrm, None, ptr(bits[0]), stream())
        energy = float(LS.energies_of_bits(prob, bits, 1)[0])
        spins = None
        local_iterations = 0
        if self.use_local_search:
            if self.ls_policy == "default":
                st = LS.flip_refinement(prob, bits, 1, max_iterations=self.max_local_iterations, seed=self._seed())
            else:  # refine_candidates (models/spec_adj.py:125-133)
                st = guards.refine(self.ls_policy, prob, bits, 1, max_iterations=self.max_local_iterations,
                                   seed=self._seed())
            local_iterations = st.passes
            refined = float(st.energies_final[0])
            if refined < energy:
                energy = refined
                spins = LS.extract(prob, st, 0).cpu().numpy().astype(self.dtype)
        if spins is None:
            spins = unpack_bits(bits[0].cpu().numpy(), n).astype(self.dtype)
        total = time.time() - start
        store_general(self.dataset, self.solver_name, self.instance_name, {
            'instance_name': self.instance_name, 'dataset': self.dataset, 'seed': self.random_seed,
            'solver_name': self.solver_name, 'energy': energy, 'time': total,
            'local_search': self.use_local_search, 'local_search_iterations': local_iterations,
            'alpha': self.ALPHA, 'beta': self._beta_column(), 'mode': self.mode, 'precision': self.precision,
            'eigsh_maxiter': self.eigsh_maxiter, 'eigsh_iterations': self._eigsh_iterations,
            'eigsh_converged': self._eigsh_converged, 'timestamp': datetime.now().strftime('%Y-%m-%d %H:%M:%S'),
        })
        self._mu = 1.0 - res.theta
        return energy, spins, total

    def _beta_column(self):
        return 1.0


class Alpha0GPU(_Endpoint):
    """SpecAdj solver (alpha = 0), models/spec_adj.py:21-154."""
    ALPHA = 0.0
    CLAMP = 1e-12
    DEFAULT_RUN = "alpha0_gpu"

    def __init__(self, instance_name: str, dataset: str, random_seed: int = None,
                 mode: int = 0, beta: float = None, alpha: float = 0,
                 use_local_search: bool = True, max_local_iterations: int = 1000,
                 eigsh_tol: float = 1e-6, eigsh_maxiter: int = 1000, run_name: str = None,
                 ls_policy: str = "default", precision: int = 32, **kwargs):
        super().__init__(instance_name, dataset, random_seed, mode, beta, use_local_search, max_local_iterations,
                         eigsh_tol, eigsh_maxiter, run_name, ls_policy, precision, **kwargs)
        self.alpha = 0


class Alpha1GPU(_Endpoint):
    """SpecSignedLap solver (alpha = 1), models/spec_signed_lap.py:21-169."""
    ALPHA = 1.0
    CLAMP = 1e-7
    DEFAULT_RUN = "alpha1_gpu"

    def __init__(self, instance_name: str, dataset: str, random_seed: int = None,
                 mode: int = 0, beta: float = None,
                 use_local_search: bool = True, max_local_iterations: int = 1000,
                 eigsh_tol: float = 1e-6, eigsh_maxiter: int = 1000, run_name: str = None,
                 ls_policy: str = "default", precision: int = 32, **kwargs):
        super().__init__(instance_name, dataset, random_seed, mode, beta, use_local_search, max_local_iterations,
                         eigsh_tol, eigsh_maxiter, run_name, ls_policy, precision, **kwargs)
        self.alpha = 1.0
