"""Guarded local-search policies on the GPU: the counterpart of models/local_search_guards/* and
models/spectral_annealing_safe.py, chosen with ``ls_policy`` (policies.py:13-18).

* fixed_guard / adaptive_guard (Fixed / Adaptive Guard, safe_flip_refinement :1035-1335): rounds of
  Jacobi batch flips, each kept per candidate only if the exact energy strictly drops; the adaptive form
  shrinks a rejected candidate's batch share and grows an accepted one's (safe_update_policy.py:62-95);
  a candidate leaves a phase after `patience` rounds without progress. Main phase at subset_fraction,
  then a final phase of final_sweep_iters * final_sweep_rounds rounds at final_sweep_fraction.
* independent_set (Independent-Set, conflict_free.py:82-296): one class of the largest-first colouring
  per step, classes in a seeded permutation per sweep; a candidate stops after a sweep without a flip.
* sequential_oracle (Sequential-Oracle, sequential_oracle.py:22-121): exact one-flip steepest descent
  to a one-flip local optimum.

The candidates live in the local search's layout with exact fields (int32 for integer couplings, FP64
for FP32 couplings J / c), so every energy and energy change is exact (integer) or an exact-field sum
(FP64), and the acceptance tolerances abs + rel * max(1, |E|) apply to the normalised energy E / c as
in the reference. Differences from the reference, by design:
* the Guard's batch is a hashed share f of the nodes (node i in round r iff
  (mix3(seed ^ GUARD_SALT, r, i) >> 11) < f 2^53) instead of a cupy random permutation prefix of
  exactly int(f n) nodes, so that a seed fixes the run and the CPU oracle reproduces it;
* a proposal's energy change is computed from the fields of its flipped rows (including the pairs of
  adjacent flipped nodes) instead of a full energy evaluation per round;
* the colouring runs on the GPU (the same largest-first colours as the reference's host colouring).
The guard invariants (safe_flip_refinement :1316-1325: no candidate ends above its start energy; every
committed update is a strict decrease) are checked at the end, together with the maintained energies
against energies recomputed from the fields; a failure raises GuardInvariantError.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from . import local_search as LS
from ._lib import call
from .device import DeviceProblem, ptr, stream

# policies.py:13-18 (ls_policy value -> the reference's refiner name)
POLICIES = {
    "fixed_guard": "Safe-Fixed",
    "adaptive_guard": "Safe-Adaptive",
    "independent_set": "Conflict-Free",
    "sequential_oracle": "Sequential-Oracle",
}
LS_POLICIES = ("default",) + tuple(POLICIES)
# policies that need the couplings on the host in the reference (its colouring); here none does
NEEDS_HOST_MATRIX = {"independent_set"}

# policies.py:21-37
PAPER_LOCAL_SEARCH = {
    "subset_fraction": 0.9,
    "final_sweep_fraction": 1.0,
    "final_sweep_iters": 25,
    "final_sweep_rounds": 25,
    "energy_chunk_size": 16,
    "absolute_tolerance": 1e-10,
    "relative_tolerance": 1e-12,
    "identity_absolute_tolerance": 1e-9,
    "identity_relative_tolerance": 1e-10,
    "safe_patience": 3,
    "minimum_subset_fraction": 0.01,
    "batch_shrink_factor": 0.5,
    "batch_growth_factor": 1.25,
    "rejection_threshold": 0.25,
    "oracle_max_steps_multiplier": 100,
}

GUARD_SALT = 0x6A09E667F3BCC908
TWO53 = 1 << 53
_M = (1 << 64) - 1


class GuardInvariantError(RuntimeError):
    """A declared invariant of a guarded policy does not hold (local_search_guards/common.py:13-14)."""


@dataclass
class GuardState:
    spins: torch.Tensor          # int8 [n][kp], node-major
    fields: torch.Tensor
    groups: int
    k: int
    iterations: int              # the count run_policy reports (policies.py:47-49)
    energies_start: np.ndarray   # -s^T J s / 2 in the couplings' units: int64 (integer J) or float64
    energies_final: np.ndarray
    summary: dict = field(default_factory=dict)

    @property
    def exact(self) -> bool:
        return self.energies_start.dtype.kind == "i"

    @property
    def passes(self) -> int:
        return self.iterations


def _state(prob, spins, fields, groups, k, iterations, E0, E, exact, summary) -> GuardState:
    """The result with energies in the couplings' units (the FP64 path tracks them in units of J / c)."""
    if not exact:
        E0, E = E0 * prob.c, E * prob.c
    return GuardState(spins, fields, groups, k, int(iterations), E0, E, summary)


def _tolerance(E: np.ndarray, scale: float, cfg: dict) -> np.ndarray:
    En = np.asarray(E, dtype=np.float64) / scale
    return cfg["absolute_tolerance"] + cfg["relative_tolerance"] * np.maximum(1.0, np.abs(En))


def _bits_per_group(mask: np.ndarray, groups: int) -> np.ndarray:
    """Boolean per candidate -> one uint64 word per 64-candidate group (bit l = candidate 64 g + l)."""
    kp = 64 * groups
    full = np.zeros(kp, dtype=bool)
    full[:mask.size] = mask
    words = np.packbits(full.reshape(groups, 64)[:, ::-1], axis=1).view(">u8").ravel()
    return words.astype(np.uint64)


def _upload_words(words: np.ndarray, dev) -> torch.Tensor:
    return torch.from_numpy(words.astype(np.uint64).view(np.int64).copy()).to(dev)


def _start(prob: DeviceProblem, cand_bits: torch.Tensor, k: int):
    spins, fields, groups, e2 = LS.init_fields(prob, cand_bits, k)
    exact = fields.dtype == torch.int32
    if exact:
        e2h = e2[:k].cpu().numpy()
        E = -(e2h // 2)
        return spins, fields, groups, E, exact, float(prob.c), None
    E = -0.5 * np.asarray(e2[:k], dtype=np.float64)
    return spins, fields, groups, E, exact, 1.0, prob.float_couplings()


def _fields_energies(prob, spins, fields, groups, k, exact) -> np.ndarray:
    """-sum_i s_i h_i / 2 from the maintained fields (the energies the policies must have tracked)."""
    if exact:
        e2 = torch.zeros(64 * groups, dtype=torch.int64, device=prob.device)
        call("sa_ls_energy", prob.n, k, groups, ptr(spins), ptr(fields), ptr(e2), stream())
        return -(e2[:k].cpu().numpy() // 2)
    return -0.5 * LS._float_energies(prob, spins, fields, groups, k)


def _check(prob, spins, fields, groups, k, exact, E0, E, name) -> None:
    if np.any(E > E0):
        raise GuardInvariantError(f"{name}: a candidate ended above its start energy")
    Ef = _fields_energies(prob, spins, fields, groups, k, exact)
    if exact:
        ok = np.array_equal(Ef, E)
    else:
        ok = np.allclose(Ef, E, rtol=1e-9, atol=1e-9 * max(1.0, float(np.abs(E).max())))
    if not ok:
        raise GuardInvariantError(f"{name}: the tracked energies differ from the fields' energies")


def safe_refine(prob: DeviceProblem, cand_bits: torch.Tensor, k: int, *, adaptive: bool, max_iterations: int,
                seed: int, cfg: dict | None = None) -> GuardState:
    """Fixed / Adaptive Guard (safe_flip_refinement with the policies' settings, safe_common.py:16-44)."""
    cfg = dict(PAPER_LOCAL_SEARCH, **(cfg or {}))
    spins, fields, groups, E, exact, scale, fv = _start(prob, cand_bits, k)
    E0 = E.copy()
    dev, kp, n = prob.device, 64 * groups, prob.n
    gseed = (int(seed) ^ GUARD_SALT) & _M
    mask = torch.empty(n * groups, dtype=torch.int64, device=dev)
    counts = torch.zeros(kp, dtype=torch.int32, device=dev)
    delta = torch.zeros(kp, dtype=torch.int64 if exact else torch.float64, device=dev)
    patience = int(cfg["safe_patience"])
    fmin, shrink, grow = cfg["minimum_subset_fraction"], cfg["batch_shrink_factor"], cfg["batch_growth_factor"]
    rid = 0
    rounds = {}
    totals = dict(accepted_updates=0, rejected_updates=0, proposed_flips=0, accepted_flips=0)
    phases = (("main", int(max_iterations), float(cfg["subset_fraction"])),
              ("final", int(cfg["final_sweep_iters"]) * int(cfg["final_sweep_rounds"]),
               float(cfg["final_sweep_fraction"])))
    for phase, nrounds, f0 in phases:
        done = 0
        active = np.ones(k, dtype=bool)
        streak = np.zeros(k, dtype=np.int64)
        frac = np.full(k, f0)
        for _ in range(max(nrounds, 0)):
            if not active.any():
                break
            thresh = np.zeros(kp, dtype=np.uint64)
            thresh[:k] = [int(f * TWO53) if a else 0 for f, a in zip(frac, active)]
            counts.zero_()
            delta.zero_()
            args = (k, groups, ptr(spins), ptr(fields), gseed, rid, ptr(_upload_words(thresh, dev)), ptr(mask),
                    ptr(counts), ptr(delta), stream())
            if exact:
                call("sa_guard_propose", prob.csr_ref, *args)
            else:
                call("sa_guard_propose_f64", prob.csr_ref, ptr(fv), *args)
            cnt = counts[:k].cpu().numpy()
            dE = delta[:k].cpu().numpy()
            has = active & (cnt > 0)
            accepted = has & np.isfinite(dE) & (dE / scale < -_tolerance(E, scale, cfg))
            rejected = has & ~accepted
            no_flips = active & ~has
            if accepted.any():
                acc = _upload_words(_bits_per_group(accepted, groups), dev)
                if exact:
                    call("sa_guard_commit", prob.csr_ref, groups, ptr(spins), ptr(fields), ptr(mask), ptr(acc),
                         stream())
                else:
                    call("sa_guard_commit_f64", prob.csr_ref, ptr(fv), groups, ptr(spins), ptr(fields), ptr(mask),
                         ptr(acc), stream())
                E = E + np.where(accepted, dE, 0).astype(E.dtype)
            if adaptive:
                newf = np.where(rejected, np.maximum(fmin, frac * shrink), frac)
                newf = np.where(accepted, np.minimum(f0, frac * grow), newf)
                stalled = no_flips | (rejected & ~(newf < frac))
                frac = newf
            else:
                stalled = active & ~accepted
            streak = np.where(active, np.where(accepted, 0, np.where(stalled, streak + 1, 0)), streak)
            active = active & (streak < patience)
            totals["accepted_updates"] += int(accepted.sum())
            totals["rejected_updates"] += int(rejected.sum())
            totals["proposed_flips"] += int(cnt[has].sum())
            totals["accepted_flips"] += int(cnt[accepted].sum())
            rid += 1
            done += 1
        rounds[phase] = done
    name = POLICIES["adaptive_guard" if adaptive else "fixed_guard"]
    _check(prob, spins, fields, groups, k, exact, E0, E, name)
    summary = dict(implementation="safe_adaptive" if adaptive else "safe_fixed", main_iterations=rounds["main"],
                   final_iterations=rounds["final"], candidate_energy_increases=0,
                   total_monotonicity_violations=0, **totals)
    return _state(prob, spins, fields, groups, k, rounds["main"], E0, E, exact, summary)


def independent_set_refine(prob: DeviceProblem, cand_bits: torch.Tensor, k: int, *, max_iterations: int,
                           seed: int, cfg: dict | None = None) -> GuardState:
    """Independent-Set (conflict_free.refine with the policies' settings)."""
    cfg = dict(PAPER_LOCAL_SEARCH, **(cfg or {}))
    spins, fields, groups, E, exact, scale, fv = _start(prob, cand_bits, k)
    E0 = E.copy()
    dev, kp = prob.device, 64 * groups
    col = LS.colour_graph(prob, 0, largest_first=True)
    C = col.num_classes
    edt = torch.int64 if exact else torch.float64
    energy = torch.zeros(kp, dtype=edt, device=dev)
    energy[:k] = torch.from_numpy(np.asarray(E, dtype=np.int64 if exact else np.float64))
    delta = torch.zeros(kp, dtype=edt, device=dev)
    flips = torch.zeros(kp, dtype=torch.int32, device=dev)
    sweep = torch.zeros(kp, dtype=torch.int32, device=dev)
    thr = torch.zeros(kp, dtype=torch.int32 if exact else torch.float64, device=dev)
    active = np.ones(k, dtype=bool)
    active_dev = _upload_words(_bits_per_group(active, groups), dev)
    rng = np.random.default_rng(int(seed))
    order = rng.permutation(C)
    pos = 0
    steps = 0
    sweeps = 0
    max_updates = int(max_iterations) + int(cfg["final_sweep_iters"]) * int(cfg["final_sweep_rounds"])
    at, rt = float(cfg["absolute_tolerance"]), float(cfg["relative_tolerance"])

    def prep():
        if exact:
            call("sa_guard_class_prep", groups, scale, at, rt, ptr(energy), ptr(delta), ptr(flips), ptr(sweep),
                 ptr(thr), stream())
        else:
            call("sa_guard_class_prep_f64", groups, at, rt, ptr(energy), ptr(delta), ptr(flips), ptr(sweep),
                 ptr(thr), stream())

    for _ in range(max(max_updates, 0)):
        if not active.any():
            break
        cls = int(order[pos])
        lo, hi = col.offsets[cls], col.offsets[cls + 1]
        nodes = col.class_nodes[lo:hi]
        prep()
        if exact:
            call("sa_guard_class_step", prob.csr_ref, groups, ptr(nodes), hi - lo, ptr(active_dev), ptr(thr),
                 ptr(spins), ptr(fields), ptr(flips), ptr(delta), stream())
        else:
            call("sa_guard_class_step_f64", prob.csr_ref, ptr(fv), groups, ptr(nodes), hi - lo, ptr(active_dev),
                 ptr(thr), ptr(spins), ptr(fields), ptr(flips), ptr(delta), stream())
        steps += 1
        pos += 1
        if pos == C:
            prep()
            certified = active & (sweep[:k].cpu().numpy() == 0)
            active &= ~certified
            sweep.zero_()
            active_dev = _upload_words(_bits_per_group(active, groups), dev)
            order = rng.permutation(C)
            pos = 0
            sweeps += 1
    prep()
    E = energy[:k].cpu().numpy().copy()
    _check(prob, spins, fields, groups, k, exact, E0, E, POLICIES["independent_set"])
    summary = dict(implementation="conflict_free_coloring", color_count=C, kernel_updates=steps,
                   completed_color_sweeps=sweeps, color_sweep_certified_candidates=int((~active).sum()),
                   rejected_updates=0, accepted_monotonicity_violations=0)
    return _state(prob, spins, fields, groups, k, steps, E0, E, exact, summary)


def sequential_refine(prob: DeviceProblem, cand_bits: torch.Tensor, k: int, *, cfg: dict | None = None
                      ) -> GuardState:
    """Sequential-Oracle (sequential_oracle.refine with the policies' settings)."""
    cfg = dict(PAPER_LOCAL_SEARCH, **(cfg or {}))
    spins, fields, groups, E, exact, scale, fv = _start(prob, cand_bits, k)
    E0 = E.copy()
    dev, kp = prob.device, 64 * groups
    energy = torch.zeros(kp, dtype=torch.int64 if exact else torch.float64, device=dev)
    energy[:k] = torch.from_numpy(np.asarray(E, dtype=np.int64 if exact else np.float64))
    steps = torch.zeros(groups, dtype=torch.int64, device=dev)
    flips = torch.zeros(kp, dtype=torch.int32, device=dev)
    max_steps = int(cfg["oracle_max_steps_multiplier"]) * prob.n
    at, rt = float(cfg["absolute_tolerance"]), float(cfg["relative_tolerance"])
    if exact:
        call("sa_guard_sequential", prob.csr_ref, k, groups, ptr(spins), ptr(fields), ptr(energy), scale, at, rt,
             max_steps, ptr(steps), ptr(flips), stream())
    else:
        call("sa_guard_sequential_f64", prob.csr_ref, ptr(fv), k, groups, ptr(spins), ptr(fields), ptr(energy), at,
             rt, max_steps, ptr(steps), ptr(flips), stream())
    st_h = steps.cpu().numpy()
    if np.any(st_h < 0):
        raise GuardInvariantError(
            f"Sequential-Oracle failed to reach a 1-flip local optimum within maximum_steps={max_steps}")
    E = energy[:k].cpu().numpy().copy()
    _check(prob, spins, fields, groups, k, exact, E0, E, POLICIES["sequential_oracle"])
    iters = int(st_h.max()) if st_h.size else 0
    summary = dict(implementation="sequential_oracle", kernel_updates=iters,
                   executed_flips=int(flips[:k].sum().item()), local_optimal_candidate_count=k)
    return _state(prob, spins, fields, groups, k, iters, E0, E, exact, summary)


def refine(policy: str, prob: DeviceProblem, cand_bits: torch.Tensor, k: int, *, max_iterations: int,
           seed: int) -> GuardState:
    """The policy's refinement of k candidates given as spin bits (k x nwords)."""
    if policy not in POLICIES:
        raise KeyError(f"Unknown guard policy {policy!r}; choose {', '.join(POLICIES)}")
    seed = 0 if seed is None else int(seed)
    if policy in ("fixed_guard", "adaptive_guard"):
        return safe_refine(prob, cand_bits, k, adaptive=policy == "adaptive_guard", max_iterations=max_iterations,
                           seed=seed)
    if policy == "independent_set":
        return independent_set_refine(prob, cand_bits, k, max_iterations=max_iterations, seed=seed)
    return sequential_refine(prob, cand_bits, k)


def _problem_and_bits(J_csr, spins_list, dtype):
    from .spectral_annealing import _problem_of, _spin_bits_of
    prob = _problem_of(J_csr, dtype)
    cols = torch.stack([t if isinstance(t, torch.Tensor) else torch.from_numpy(np.asarray(t))
                        for t in spins_list]).to(device=prob.device, dtype=torch.int8).contiguous()
    return prob, _spin_bits_of(prob, cols), cols.shape[0]


def run_policy(policy, J_csr, spins_list, *, max_iterations, dtype=np.float32, seed=0, matrix_cpu=None):
    """policies.py:40-52: refine the candidates with a guarded policy; (n x k int8 device spins,
    iterations, summary). J_csr: scipy couplings or a DeviceProblem; matrix_cpu is accepted and unused
    (the colouring runs on the GPU)."""
    prob, bits, k = _problem_and_bits(J_csr, spins_list, dtype)
    st = refine(policy, prob, bits, k, max_iterations=int(max_iterations), seed=seed)
    return st.spins[:, :k], st.iterations, dict(st.summary)


def refine_candidates(policy, J_csr, spins_list, *, max_iterations, dtype=np.float32, seed=0, J_host=None):
    """policies.py:55-69: run_policy for SpecAdj and SpecSignedLap; (n x k int8 spins, iterations)."""
    refined, iterations, _ = run_policy(policy, J_csr, spins_list, max_iterations=max_iterations, dtype=dtype,
                                        seed=seed)
    return refined, iterations


__all__ = ["POLICIES", "LS_POLICIES", "PAPER_LOCAL_SEARCH", "GuardInvariantError", "GuardState", "refine",
           "run_policy", "refine_candidates", "safe_refine", "independent_set_refine", "sequential_refine"]
