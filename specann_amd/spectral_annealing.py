"""SpecAnn (Algorithm 1) on the MI355X: the counterpart of models/spectral_annealing.py.

Same constructor (:172-260), same per-alpha logic (:311-421), same post-sweep selection
(:423-528), same results columns; the compute runs in libspecann:

* normalisation, per-alpha scaling and the Gershgorin bound: sa_abs_degree / sa_normalize /
  sa_alpha_scaling (:270-288, :316, :579);
* predictor + warm start: sa_lanczos_start (the diagonal predictor is applied on the fly, and its
  residual gate is Lanczos step 0, :589-612);
* eigensolve: the warm-started Lanczos of specann_amd.eigsh (:568-677);
* sign alignment, normalisation, rounding and the dedup signature: sa_round (:331-370, :557-561);
* energies: exact int64 (:709-720, :962-976) from the local search's field initialisation;
* local search: the batched deterministic descent of specann_amd.local_search (:423-481).

Energies are exact for integer couplings, and the reported energy is that exact value (the
reference recomputes it in float64 from the input J after the timer, :483-488, which is the same
number for integer J). Other couplings take the floating-point local search (FP32 J / c, exact FP64
fields) and FP64 energies, and the reported energy is recomputed from the input J as the reference
does. ``solve`` returns the best spins as a host numpy array of +-1 in the
solver's dtype.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from datetime import datetime

import numpy as np
import torch

from . import guards
from . import local_search as LS
from ._lib import SA_BROKEN, SA_CONVERGED, SA_MAXITER, SA_REJECTED, EigensolverBreakdown, call
from .device import DeviceProblem, ptr, require_gpu, stream
from .eigsh import LanczosSolver, use_shadow
from .failures import NumericalFailure, is_numerical, is_out_of_memory
from .results import store_alphas, store_general
from .solution_io import ising_energy

PRED_MODES = ("diagonal", "secant", "prev", "tangent")
LS_POLICIES = ("default", "fixed_guard", "adaptive_guard", "independent_set", "sequential_oracle")
SCHEDULES = ("sqrt", "linear")
# eigsh tolerances below this get a true-residual check at convergence (FP64 accuracy asked for), and
# up to REFINE_ROUNDS continuations from the Ritz vector while that residual is above the tolerance
REFINE_TOL = 1e-8
REFINE_ROUNDS = 2


def compute_shift_rho(lambda1_prev, gershgorin_bound, eps=1e-3):
    """rho = (1 + eps) * max(lambda1_prev, g) (models/spectral_annealing.py:56-67)."""
    if lambda1_prev is None or not np.isfinite(lambda1_prev) or lambda1_prev <= 0:
        base = gershgorin_bound
    else:
        base = max(lambda1_prev, gershgorin_bound)
    rho = (1.0 + eps) * base
    if rho <= 0:
        rho = (1.0 + eps) * max(gershgorin_bound, 1.0)
    return rho


@dataclass
class EigenOutcome:
    lambda1: float
    mu: float
    rho: float
    g: float
    degree_scale: float
    steps: int
    retry: bool
    converged: bool
    res_estimate: float
    residual: float | None
    pred_residual: float
    pred_accepted: bool | None
    cycles: int
    y_norm: float = 1.0       # ||Y|| of the final Ritz vector (unnormalised, in the solver's Y buffer)
    y_dot: float = 0.0        # Y . y_last


@dataclass
class SweepOutcome:
    best_energy: float
    best_spins: np.ndarray | None
    best_alpha: float | None
    rows: list = field(default_factory=list)
    num_candidates: int | None = None
    local_passes: int = 0
    sweep_time: float = 0.0
    post_sweep_time: float = 0.0
    matvecs: int = 0
    tangent_matvecs: int = 0          # operator applications of the tangent predictor's MINRES
    orth_full_steps: int = 0          # Lanczos steps orthogonalised against the whole basis
    orth_selective_steps: int = 0     # ... against the last two basis rows only
    cand_bits: object = None          # device int64 [k][nwords]: the distinct rounded candidates
    raw_energies: list = field(default_factory=list)
    refined_energies: list = field(default_factory=list)


class SpectralAnnealing:
    """SpecAnn solver (see the reference's docstring for the parameters,
    models/spectral_annealing.py:132-170). Extra keyword-only tuning: ``tile_nnz`` and
    ``lanes_per_row`` fix the SpMV tiling (default: chosen from the mean row length), ``autotune``
    times the SpMV layouts, ``pred_mode="tangent"`` (with ``tangent_steps``, ``tangent_tol``) adds the
    eigenvector-tangent predictor, ``reorth`` ("selective" or "full") sets the Gram-Schmidt policy,
    ``eigsh_keep`` the Ritz vectors a thick restart keeps (default min(12, ncv / 2)), and
    ``lanczos_graphs`` the HIP-graph replay of Lanczos step chunks (default for n <= 2^18)."""

    def __init__(self, instance_name: str, dataset: str, random_seed: int = None,
                 mode: int = 0, beta: float = None, alpha: float = None,
                 use_local_search: bool = True, max_local_iterations: int = 1000,
                 eigsh_tol: float = 1e-6, eigsh_maxiter: int = 1000, run_name: str = None,
                 disable_warm_start: bool = False, precision: int = 32,
                 pred_mode: str = "diagonal", pred_tol: float = 1e-3, shift_eps: float = 1e-3,
                 schedule: str = "sqrt", num_alphas: int = 128, schedule_interval=(0.0, 1.0),
                 eigsh_ncv: int = None, fp32_spmv: bool = None, ls_policy: str = "default",
                 ls_batch: bool = True, verbose: bool = True, tile_nnz: int = None, lanes_per_row: int = 0,
                 autotune: bool = True, tangent_steps: int = 2, tangent_tol: float = 1e-2,
                 lanczos_graphs: bool = None, reorth: str = "selective", eigsh_keep: int = None, **kwargs):
        self.instance_name = instance_name
        self.dataset = dataset
        self.random_seed = random_seed
        self.mode = mode
        self.beta = beta
        self.alpha = alpha
        self.use_local_search = use_local_search
        self.max_local_iterations = max_local_iterations
        self.eigsh_tol = float(eigsh_tol)
        self.eigsh_maxiter = int(eigsh_maxiter)
        self.eigsh_ncv = None if eigsh_ncv is None else int(eigsh_ncv)
        self.eigsh_keep = None if eigsh_keep is None else int(eigsh_keep)  # thick-restart Ritz vectors kept
        self.disable_warm_start = disable_warm_start
        if pred_mode not in PRED_MODES:
            raise ValueError(f"pred_mode must be one of {PRED_MODES}, got {pred_mode!r}")
        self.pred_mode = pred_mode
        self.pred_tol = float(pred_tol)
        self.shift_eps = float(shift_eps)
        t_min, t_max = (float(v) for v in schedule_interval)
        if schedule not in SCHEDULES:
            raise ValueError(f"schedule must be one of {SCHEDULES}, got {schedule!r}")
        if int(num_alphas) < 2 or not 0.0 <= t_min < t_max:
            raise ValueError(f"need num_alphas >= 2 and 0 <= start < end, got {num_alphas}, {schedule_interval}")
        self.schedule = schedule
        self.num_alphas = int(num_alphas)
        self.schedule_interval = (t_min, t_max)
        if precision == 32:
            self.dtype = np.float32
        elif precision == 64:
            self.dtype = np.float64
        else:
            raise ValueError(f"precision must be 32 or 64, got {precision}")
        self.precision = precision
        self.fp32_spmv = (precision == 32) if fp32_spmv is None else bool(fp32_spmv)
        if ls_policy not in LS_POLICIES:
            raise ValueError(f"ls_policy must be one of {LS_POLICIES}, got {ls_policy!r}")
        self.ls_policy = ls_policy
        self.ls_batch = bool(ls_batch)
        if not self.ls_batch and ls_policy != "default":
            raise ValueError("ls_batch = False is only defined for ls_policy 'default'")
        if self.fp32_spmv and precision != 32:
            raise ValueError("fp32_spmv needs precision 32")
        # Without fp32_spmv the reference's J/c is float64, so its eigsh runs in FP64 (:217-219).
        self.tdtype = torch.float32 if self.fp32_spmv else torch.float64
        self.solver_name = run_name if run_name else "spectral_annealing_pred"
        self.tile_nnz = tile_nnz
        self.lanes_per_row = lanes_per_row
        self.autotune = bool(autotune) and tile_nnz is None
        # pred_mode "tangent" (no reference counterpart): the diagonal predictor plus the eigenvector
        # tangent's correction, from `tangent_steps` deflated MINRES iterations (specann_amd.tangent)
        self.lanczos_graphs = lanczos_graphs      # None: HIP-graph replay of Lanczos chunks for n <= 2^18
        # called after each alpha's eigensolve as hook(alpha_idx, alpha, prob, lz, s, scale, outcome), where
        # lz.Y * scale is the sign-aligned unit eigenvector (None: no call)
        self.eigvec_hook = None
        if reorth not in ("selective", "full"):
            raise ValueError("reorth must be 'selective' or 'full'")
        self.reorth = reorth                      # full: the reference eigsh's reorthogonalisation every step
        self.tangent_steps = int(tangent_steps)
        self.tangent_tol = float(tangent_tol)
        if self.tangent_steps < 0:
            raise ValueError("tangent_steps must be >= 0")
        self.verbose = verbose
        if verbose:
            print(f"Initialized {self.solver_name} solver. Results will be saved under '{self.solver_name}'.")
            print("Parameters:")
            for key in ("instance_name", "dataset", "random_seed", "mode", "beta", "alpha", "use_local_search",
                        "max_local_iterations", "eigsh_tol", "eigsh_maxiter", "eigsh_ncv", "disable_warm_start",
                        "precision", "fp32_spmv", "pred_mode", "pred_tol", "shift_eps", "schedule", "num_alphas",
                        "schedule_interval", "ls_policy"):
                print(f"  {key}: {getattr(self, key)}")
            print(f"  run_name: {run_name}")
            if kwargs:
                print(f"  additional kwargs: {kwargs}")
            print("-" * 50)

    # ------------------------------------------------------------------------------------------
    def _get_alpha_values(self, mode: int) -> list:
        """The alpha values of the schedule (mode 0 only; :549-555)."""
        if mode != 0:
            raise ValueError(f"Unknown mode {mode}")
        t = np.linspace(*self.schedule_interval, self.num_alphas)
        return [float(x) for x in (np.sqrt(t) if self.schedule == "sqrt" else t)]

    def _seed(self) -> int:
        return 0 if self.random_seed is None else int(self.random_seed)

    # ------------------------------------------------------------------------------------------
    def solve(self, J):
        """Run SpecAnn on the couplings J (scipy sparse) and write the results CSVs.
        Returns (best energy, best spins, time in seconds)."""
        if self.alpha is not None:
            raise ValueError("Alpha is not set")  # the reference's message for a fixed alpha (:527-528)
        start_time = time.time()
        prob = DeviceProblem(J, self.tdtype, tile_nnz=self.tile_nnz, lanes_per_row=self.lanes_per_row)
        if self.autotune and prob.n >= (1 << 17):
            prob.autotune()
        out = self.run_on_device(prob, start_time=start_time)
        total_time = time.time() - start_time
        if out.best_spins is not None and prob.jint is None:
            # non-integer couplings: the reported energy is -1/2 s^T J s in FP64 from the input couplings,
            # as the reference reports it after the timer (:483-488)
            out.best_energy = ising_energy(J, out.best_spins)
        self._sweep_time = out.sweep_time
        self._post_sweep_time = total_time - out.sweep_time
        self._gpu_pool_bytes = int(torch.cuda.memory_reserved())
        if out.best_alpha is None:
            store_alphas(self.dataset, self.solver_name, self.instance_name, self.random_seed, out.rows)
            message = f"no alpha produced a solution ({self._alphas_skipped} of {self.num_alphas} skipped)"
            if self._alpha_errors:
                last = self._last_alpha_error
                raise RuntimeError(f"{message}; {self._alpha_errors} of them raised a non-numerical error, "
                                   f"the last {type(last).__name__}: {last}") from last
            raise NumericalFailure(message)
        store_general(self.dataset, self.solver_name, self.instance_name, self._general_row(
            out.best_energy, out.best_alpha, None, total_time, out.local_passes, out.num_candidates))
        store_alphas(self.dataset, self.solver_name, self.instance_name, self.random_seed, out.rows)
        return out.best_energy, out.best_spins, total_time

    def _general_row(self, energy, alpha, beta, time_taken, local_iterations, num_candidates):
        return {
            'instance_name': self.instance_name, 'dataset': self.dataset, 'seed': self.random_seed,
            'solver_name': self.solver_name, 'energy': energy, 'time': time_taken,
            'local_search': self.use_local_search, 'local_search_iterations': local_iterations, 'alpha': alpha,
            'beta': beta, 'mode': self.mode, 'precision': self.precision, 'fp32_spmv': self.fp32_spmv,
            'ls_policy': self.ls_policy, 'alphas_skipped': self._alphas_skipped,
            'alphas_retried': self._alphas_retried, 'num_candidates': num_candidates,
            'sweep_time': self._sweep_time, 'post_sweep_time': self._post_sweep_time,
            'gpu_pool_bytes': self._gpu_pool_bytes, 'timestamp': datetime.now().strftime('%Y-%m-%d %H:%M:%S'),
            'eigsh_matvecs': self._matvecs,
        }

    # ------------------------------------------------------------------------------------------
    def run_on_device(self, prob: DeviceProblem, start_time: float | None = None) -> SweepOutcome:
        """The alpha sweep and the post-sweep local search on a problem already in HBM."""
        start_time = time.time() if start_time is None else start_time
        n, dev, tdt = prob.n, prob.device, prob.dtype
        self._normalization_factor = prob.c
        alpha_list = self._get_alpha_values(self.mode)
        nwords = (n + 63) // 64
        lz = LanczosSolver(prob, self.eigsh_ncv, graphs=self.lanczos_graphs, keep=self.eigsh_keep,
                           orth_tau=0.0 if self.reorth == "full" else None,
                           shadow=use_shadow(prob.dtype, self.eigsh_tol, prob.n))
        s = torch.empty(n, dtype=tdt, device=dev)
        ybuf = [torch.empty(n, dtype=tdt, device=dev) for _ in range(3)]
        x_rand = torch.empty(n, dtype=tdt, device=dev)
        cand_bits = torch.empty((len(alpha_list) + 1, nwords), dtype=torch.int64, device=dev)
        rows, cand_of_row = [], []
        by_hash: dict = {}
        n_cand = 0
        y_last = y_prev = None
        lambda1_prev = alpha_prev = None
        self._alphas_skipped = self._alphas_retried = self._alpha_errors = 0
        self._last_alpha_error = None
        self._matvecs = 0
        self._tangent_matvecs = 0
        self._orth_steps = [0, 0]                 # Lanczos steps orthogonalised fully / selectively
        self._tangent = None
        if self.pred_mode == "tangent" and not self.disable_warm_start:
            from .tangent import TangentPredictor
            self._tangent = TangentPredictor(prob, self.tangent_steps, self.tangent_tol)
        self._tangent_alpha = None
        self._last_ok_idx = None
        raw_known: dict = {}

        for alpha_idx, alpha in enumerate(alpha_list):
            try:
                alpha_start = time.time()
                eo = self._eigensolve(prob, lz, s, x_rand, alpha_idx, alpha, alpha_prev, y_last, y_prev,
                                      lambda1_prev)
                self._alphas_retried += int(eo.retry)
                self._matvecs += eo.steps
                oc = lz.work.full().counters
                self._orth_steps[0] += int(oc[0])
                self._orth_steps[1] += int(oc[1])
                bound = (-0.5 * eo.lambda1 * eo.degree_scale * prob.c) if eo.converged else float('nan')
                # sign alignment with y_last and normalisation of the final Ritz vector (:653-660)
                nrm, dot = eo.y_norm, eo.y_dot
                sign = -1.0 if (y_last is not None and dot < 0) else 1.0
                free = next(b for b in ybuf if b is not y_last and b is not y_prev)
                slot = n_cand
                call(f"sa_round_{prob.sfx}", prob.work.ptr, n, ptr(lz.Y), sign / nrm, ptr(free),
                     ptr(cand_bits[slot]), stream())
                if self.eigvec_hook is not None:
                    # the aligned unit eigenvector is lz.Y * (sign / nrm), for measurements beside the sweep
                    self.eigvec_hook(alpha_idx, alpha, prob, lz, s, sign / nrm, eo)
                sig = int(prob.work.head().hash)
                alpha_before = alpha_prev
                if not self.disable_warm_start:
                    y_prev, y_last = y_last, free
                    lambda1_prev, alpha_prev = eo.lambda1, alpha
                tangent_iters = None
                if self._tangent is not None and alpha_idx + 1 < len(alpha_list) and self.tangent_steps > 0:
                    # tangent at this alpha (s = s(alpha), y_last unit) for the next prediction, its MINRES
                    # started from the backward difference when the previous alpha is the one before
                    hist = None
                    if y_prev is not None and alpha_before is not None and alpha_idx >= 1 and \
                            self._last_ok_idx == alpha_idx - 1:
                        hist = (y_prev, alpha - alpha_before)
                    self._tangent.compute(s, y_last, eo.lambda1, hist)
                    tangent_iters = self._tangent.matvecs
                    self._tangent_matvecs += tangent_iters
                    self._tangent_alpha = alpha
                self._last_ok_idx = alpha_idx
                cand = None
                for c in by_hash.get(sig, []):
                    call("sa_bits_compare", prob.work.ptr, nwords, ptr(cand_bits[slot]), ptr(cand_bits[c]), stream())
                    if prob.work.head().i64[0] == 0:
                        cand = c
                        break
                if cand is None:
                    cand = n_cand
                    by_hash.setdefault(sig, []).append(cand)
                    n_cand += 1
                    if not self.use_local_search:
                        raw_known[cand] = float(LS.energies_of_bits(prob, cand_bits[cand:cand + 1], 1)[0])
                alpha_time = time.time() - alpha_start
                rows.append({
                    'instance_name': self.instance_name, 'dataset': self.dataset, 'seed': self.random_seed,
                    'solver_name': self.solver_name, 'alpha': alpha, 'beta': None, 'energy': None,
                    'raw_energy': None, 'time': alpha_time, 'mode': self.mode,
                    'use_local_search': self.use_local_search, 'local_search_iterations': 0,
                    'eigsh_iterations': eo.steps, 'eigsh_retry': eo.retry, 'eigsh_converged': eo.converged,
                    'warm_start': not self.disable_warm_start, 'alpha_index': alpha_idx,
                    'precision': self.precision, 'fp32_spmv': self.fp32_spmv, 'lambda1': eo.lambda1,
                    'degree_scale': eo.degree_scale, 'energy_bound': bound, 'pred_mode': self.pred_mode,
                    'pred_residual': eo.pred_residual, 'pred_accepted': eo.pred_accepted,
                    'candidate': cand if self.use_local_search else None,
                    'timestamp': datetime.now().strftime('%Y-%m-%d %H:%M:%S'),
                    'residual_estimate': eo.res_estimate, 'residual': eo.residual, 'lanczos_restarts': eo.cycles,
                    'tangent_iterations': tangent_iters,
                })
                cand_of_row.append(cand)
            except Exception as e:
                if is_out_of_memory(e):
                    raise
                if self.verbose:
                    print(f"Alpha {alpha} failed: {e}")
                self._alphas_skipped += 1
                if not is_numerical(e):
                    self._alpha_errors += 1
                    self._last_alpha_error = e.with_traceback(None)
                continue
        torch.cuda.synchronize()
        sweep_end = time.time()

        # ---- post-sweep: raw energies, local search, selection (:423-481) ----------------------
        k = n_cand
        out = SweepOutcome(float("inf"), None, None, rows, k if self.use_local_search else None)
        if k == 0:
            out.sweep_time = sweep_end - start_time
            return out
        state = None
        if self.use_local_search:
            if self.verbose:
                print(f"Applying local search to {k} unique solutions (out of {len(alpha_list)} total)")
            state = self._refine(prob, cand_bits[:k], k)
            as_energy = int if state.exact else float
            raw = [as_energy(v) for v in state.energies_start]
            refined = [as_energy(v) for v in state.energies_final]
            out.local_passes = state.passes
        else:
            as_energy = int if prob.jint is not None else float
            raw = [as_energy(raw_known[c]) for c in range(k)]
            refined = raw
        best_e, best_row = float("inf"), None
        for i, r in enumerate(rows):                      # the sweep's own best (:405-407)
            r['raw_energy'] = float(raw[cand_of_row[i]])
            r['energy'] = r['raw_energy']
            if r['energy'] < best_e:
                best_e, best_row = r['energy'], i
        best_source = ('original', cand_of_row[best_row]) if best_row is not None else None
        if self.use_local_search:
            improved = {}
            for c in range(k):
                if refined[c] > raw[c]:
                    improved[c] = (('original', c), raw[c])
                else:
                    improved[c] = (('refined', c), refined[c])
            for i, r in enumerate(rows):
                src, e = improved[cand_of_row[i]]
                r['energy'] = float(e)
                r['local_search_iterations'] = state.passes
                if r['energy'] < best_e:
                    best_e, best_row, best_source = r['energy'], i, src
        out.best_energy = best_e
        out.best_alpha = rows[best_row]['alpha'] if best_row is not None else None
        out.cand_bits = cand_bits[:k]
        out.raw_energies, out.refined_energies = raw, refined
        if best_source is not None:
            kind, c = best_source
            if kind == 'refined':
                spins = LS.extract(prob, state, c).cpu().numpy()
            else:
                spins = unpack_bits(cand_bits[c].cpu().numpy(), n)
            out.best_spins = spins.astype(self.dtype)
        torch.cuda.synchronize()
        out.sweep_time = sweep_end - start_time
        out.post_sweep_time = time.time() - sweep_end
        out.matvecs = self._matvecs
        out.tangent_matvecs = self._tangent_matvecs
        out.orth_full_steps, out.orth_selective_steps = self._orth_steps
        return out

    def _refine(self, prob: DeviceProblem, cand_bits: torch.Tensor, k: int):
        """Local search of the distinct candidates (_refine_spins, :530-547): the default search, or the
        guarded policy ls_policy names (specann_amd.guards)."""
        seed = self._seed()
        if self.ls_policy != "default":
            st = guards.refine(self.ls_policy, prob, cand_bits, k, max_iterations=self.max_local_iterations,
                               seed=seed)
            self._last_ls_summary = st.summary
            return st
        col = LS.colour_graph(prob, seed)
        if self.ls_batch:
            return LS.flip_refinement(prob, cand_bits, k, max_iterations=self.max_local_iterations, seed=seed,
                                      colouring=col)
        # SpecAnn[-FastL]: the same search, one candidate per launch
        parts = [LS.flip_refinement(prob, cand_bits[c:c + 1], 1, max_iterations=self.max_local_iterations,
                                    seed=seed, colouring=col) for c in range(k)]
        spins = torch.cat([p.spins[:, :1] for p in parts], dim=1)
        kp = 64 * ((k + 63) // 64)
        spins = torch.nn.functional.pad(spins, (0, kp - k))
        return LS.LocalSearchState(spins, None, kp // 64, k, max(p.passes for p in parts),
                                   sum(p.total_passes for p in parts),
                                   np.concatenate([p.e2_start for p in parts]),
                                   np.concatenate([p.e2_final for p in parts]), scale=parts[0].scale)

    # ------------------------------------------------------------------------------------------
    def _eigensolve(self, prob, lz, s, x_rand, alpha_idx, alpha, alpha_prev, y_last, y_prev, lambda1_prev
                    ) -> EigenOutcome:
        """Predict, gate, solve, retry cold once on failure, check (:568-669)."""
        degree_scale, g = prob.alpha_scaling(alpha, 1e-12, s)
        rho = compute_shift_rho(lambda1_prev, g, eps=self.shift_eps)
        bound = rho + g
        seed = self._seed()
        pred_residual, pred_accepted = float('nan'), None
        retry = False
        cold = alpha_idx == 0 or self.disable_warm_start or y_last is None

        def cold_start(stream_id):
            prob.fill_normal(seed, stream_id, x_rand)
            lz.begin(self.eigsh_tol, self.eigsh_maxiter, bound)
            lz.start(s, x_rand)
            return lz.run(s, self.eigsh_maxiter)

        try:
            if cold:
                res = cold_start(alpha_idx * 4)
            elif self.pred_mode == "prev":
                lz.begin(self.eigsh_tol, self.eigsh_maxiter, bound)
                lz.start(s, y_last)
                res = lz.run(s, self.eigsh_maxiter)
            else:
                lz.begin(self.eigsh_tol, self.eigsh_maxiter, bound, pred_tol=self.pred_tol)
                self._predict_start(prob, lz, s, alpha, alpha_prev, y_last, y_prev)
                res = lz.run(s, self.eigsh_maxiter, first_check=1)
                if res.status == SA_BROKEN and not (res.start_norm >= 1e-12):
                    # zero predicted vector: random start (utils/predictors.py:22-25)
                    prob.fill_normal(seed, alpha_idx * 4 + 1, x_rand)
                    lz.begin(self.eigsh_tol, self.eigsh_maxiter, bound, pred_tol=self.pred_tol)
                    lz.start(s, x_rand, gate=1)
                    res = lz.run(s, self.eigsh_maxiter, first_check=1)
                pred_residual = float(res.first_beta)
                pred_accepted = not pred_residual > self.pred_tol
                if res.status == SA_REJECTED:
                    lz.begin(self.eigsh_tol, self.eigsh_maxiter, bound)
                    lz.start(s, y_last)
                    res = lz.run(s, self.eigsh_maxiter)
            self._check(res, g)
            nrm, dot = self._final_vector(lz, y_last)
        except Exception as e:
            if is_out_of_memory(e):
                raise
            if self.verbose:
                print(f"Warning: eigsh failed for alpha={alpha}, rho_k={rho}, error: {e}")
            retry = True
            res = cold_start(alpha_idx * 4 + 2)
            self._check(res, g)
            nrm, dot = self._final_vector(lz, y_last)
        converged = res.status == SA_CONVERGED
        residual = None
        steps = res.steps
        if converged and self.eigsh_tol < REFINE_TOL:
            # tight tolerances: the Lanczos estimate |beta s_j| can undershoot the true residual by a small
            # factor (rounding in the recurrence and the restarts); check it and continue from the Ritz
            # vector while it is above tol (at most REFINE_ROUNDS times, within the maxiter budget)
            residual = self._true_residual(prob, lz, s, res.theta, nrm)
            for _ in range(REFINE_ROUNDS):
                if residual <= self.eigsh_tol or steps >= self.eigsh_maxiter:
                    break
                lz.begin(self.eigsh_tol, self.eigsh_maxiter - steps, bound)
                lz.start(s, lz.Y)
                res = lz.run(s, self.eigsh_maxiter - steps)
                self._check(res, g)
                nrm, dot = self._final_vector(lz, y_last)
                steps += res.steps
                residual = self._true_residual(prob, lz, s, res.theta, nrm)
        if not converged:
            # at maxiter: the true residual ||N y - theta y|| / ||y|| (:640-647)
            residual = self._true_residual(prob, lz, s, res.theta, nrm)
            converged = residual <= self.eigsh_tol
            if not converged and self.verbose:
                print(f"Warning: eigsh did not converge for alpha={alpha}: residual {residual:.3g} > tol "
                      f"{self.eigsh_tol:g}; energy bound not certified")
        return EigenOutcome(res.theta, rho - res.theta, rho, g, degree_scale, steps, retry, converged, res.res,
                            residual, pred_residual, pred_accepted, res.cycles, nrm, dot)

    @staticmethod
    def _true_residual(prob, lz, s, theta, nrm) -> float:
        """||N y - theta y|| / ||y|| of the Ritz vector in lz.Y (norm nrm): one operator application."""
        w = torch.empty_like(lz.Y)
        prob.matvec(s, lz.Y, 0.0, 1.0, w, lz.W)
        r, _ = prob.norm_dot(w, lz.Y, 1.0, -theta)
        return r / nrm

    @staticmethod
    def _final_vector(lz, y_last) -> tuple[float, float]:
        """The final Ritz vector into lz.Y; (||Y||, Y . y_last). A zero or non-finite vector is a
        breakdown, so the caller's cold retry applies to it as to any other (:616-636, :679-691)."""
        nrm, dot = lz.final_vector(y_last)
        if not (nrm > 0 and np.isfinite(nrm) and np.isfinite(dot)):
            raise EigensolverBreakdown("zero or non-finite Ritz vector")
        return nrm, dot

    def _predict_start(self, prob, lz, s, alpha, alpha_prev, y_last, y_prev):
        """Queue the predicted start vector of a warm-started alpha (gated: step 0 tests it)."""
        da = alpha - alpha_prev
        if self.pred_mode == "tangent" and self._tangent is not None and self._tangent_alpha == alpha_prev:
            # D^(da/2) (y + 2 lambda da z): the diagonal predictor plus the tangent correction
            lz.start(s, y_last, d=prob.d, pexp=0.5 * da, y=self._tangent.z, cy=2.0 * self._tangent.lam * da, gate=1)
        elif self.pred_mode in ("diagonal", "tangent"):
            lz.start(s, y_last, d=prob.d, pexp=0.5 * da, gate=1)
        else:
            self._secant_start(prob, lz, s, y_last, y_prev)

    def _secant_start(self, prob, lz, s, y_last, y_prev):
        """Secant predictor (utils/predictors.py:29-56) as the start vector cx y_last + cy y_prev."""
        if y_prev is None:
            lz.start(s, y_last, gate=1)
            return
        _, dot = prob.norm_dot(y_last, y_prev)
        sgn = -1.0 if dot < 0 else 1.0
        diff, _ = prob.norm_dot(y_last, y_prev, 1.0, -sgn)
        if diff < 1e-12:
            lz.start(s, y_last, gate=1)
            return
        lz.start(s, y_last, cx=2.0, y=y_prev, cy=-sgn, gate=1)

    def _check(self, res, g):
        """_check_eigenpair (:679-691) plus the breakdown test of _checked_ritz_matrix (:85-120)."""
        if res.status == SA_BROKEN:
            raise EigensolverBreakdown("Lanczos breakdown: non-finite or oversized tridiagonal entry")
        if not np.isfinite(res.theta):
            raise EigensolverBreakdown(f"non-finite eigenvalue {res.theta}")
        if res.theta > g * (1.0 + max(0.5 * self.shift_eps, 1e-5)):
            raise EigensolverBreakdown(f"spurious eigenvalue {res.theta:.6g} above the Gershgorin bound {g:.6g}")


def _spin_bits_of(prob: DeviceProblem, cols: torch.Tensor) -> torch.Tensor:
    """Candidate-major int8 spins [k][n] on the device -> candidate bits [k][nwords] (sa_pack_spins)."""
    k = cols.shape[0]
    nwords = (prob.n + 63) // 64
    bits = torch.empty((k, nwords), dtype=torch.int64, device=prob.device)
    call("sa_pack_spins", prob.n, k, ptr(cols), nwords, ptr(bits), stream())
    return bits


def _problem_of(J_csr, dtype) -> DeviceProblem:
    if isinstance(J_csr, DeviceProblem):
        return J_csr
    tdt = torch.float64 if np.dtype(dtype) == np.float64 else torch.float32
    return DeviceProblem(J_csr, tdt)


def flip_refinement(J_csr, spins_list, max_iterations=1000, final_sweep_fraction=1.0, final_sweep_iters=25,
                    num_rounds=25, subset_fraction=0.9, dtype=np.float32, seed: int = 0):
    """The reference's local-search callable (models/spectral_annealing.py:1039-1112): the k candidates
    of spins_list (+-1 vectors, host or device, not modified) refined on the GPU; returns (n x k int8
    device tensor, main passes). J_csr: the couplings (scipy sparse, or a DeviceProblem already on the
    GPU). The search is the deterministic colour-class descent (exact int32 fields for integer
    couplings, FP32 J / c with exact FP64 fields otherwise), so dtype only sets the precision of the
    problem this builds; seed fixes the node selection and class order."""
    prob = _problem_of(J_csr, dtype)
    cols = torch.stack([t if isinstance(t, torch.Tensor) else torch.from_numpy(np.asarray(t))
                        for t in spins_list]).to(device=prob.device, dtype=torch.int8).contiguous()
    k = cols.shape[0]
    st = LS.flip_refinement(prob, _spin_bits_of(prob, cols), k, max_iterations=max_iterations,
                            final_sweep_fraction=final_sweep_fraction, final_sweep_iters=final_sweep_iters,
                            num_rounds=num_rounds, subset_fraction=subset_fraction, seed=seed)
    return st.spins[:, :k], st.passes


def fp32_energies(J_csr, spins, normalization=1.0) -> torch.Tensor:
    """The reference's energy callable (models/spectral_annealing.py:962-976): E = -1/2 s^T J s *
    normalization for a +-1 vector or every column of an n x k matrix, with the couplings in FP32 and
    FP64 accumulation (each row's field in row order, then a fixed-order sum over the rows); a float64
    device tensor of k energies. J_csr: scipy sparse (its own values), or a DeviceProblem (then its
    normalised couplings J / c, as the reference calls it with the normalised matrix and c)."""
    if isinstance(J_csr, DeviceProblem):
        prob, fv = J_csr, J_csr.vals.to(torch.float32)
    else:
        prob = DeviceProblem(J_csr, torch.float32)
        from .device import upload
        fv = upload(np.asarray(sp_csr(J_csr).data, dtype=np.float32), prob.device)
    S = spins if isinstance(spins, torch.Tensor) else torch.from_numpy(np.asarray(spins))
    S = S.reshape(-1, 1) if S.dim() == 1 else S
    cols = S.t().to(device=prob.device, dtype=torch.int8).contiguous()
    k = cols.shape[0]
    _, _, _, e2 = LS.init_fields(prob, _spin_bits_of(prob, cols), k, fvals=fv)
    return torch.from_numpy(-0.5 * float(normalization) * e2).to(prob.device)


def sp_csr(J):
    import scipy.sparse as sp
    J = sp.csr_matrix(J)
    if not J.has_sorted_indices:
        J = J.copy()
        J.sort_indices()
    return J


def unpack_bits(words: np.ndarray, n: int) -> np.ndarray:
    """+-1 spins from uint64 words (bit l of word w = spin 64 w + l)."""
    b = np.unpackbits(words.view(np.uint8), bitorder="little")[:n]
    return np.where(b > 0, 1, -1).astype(np.int8)


__all__ = ["SpectralAnnealing", "compute_shift_rho", "flip_refinement", "fp32_energies", "unpack_bits", "require_gpu"]
