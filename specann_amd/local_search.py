"""Batched deterministic local search on the GPU: the counterpart of flip_refinement and
gpu_final_sweep (models/spectral_annealing.py:983-1112).

The reference flips spins in place from many warps at once, so its result depends on thread timing
(README.md:204-208). Here a pass visits the colour classes of a greedy colouring (independent sets)
one after another; inside a class every selected node with s_i h_i < 0 flips at once, which is the
same as flipping them one at a time because no two of them are adjacent. The outcome is therefore
the sequential single-flip descent in that order, bit for bit, for every candidate, and the energy
never increases. For integer couplings the fields h = J s are kept exactly in int32 and energies in
int64. Other couplings (update_spins_kernel_float32/64, :821-919) use the FP32 normalised couplings
J / c with FP64 fields, which are exact sums when every nonzero |J / c| is at least 2^-30 (checked),
so the search is again order-independent and bit-exact; their energies -c/2 sum_i s_i h_i are summed in
FP64 in a fixed order (fp32_energies, :962-976).

The termination schedule is the reference's: main passes over a `subset_fraction` share of the
nodes until every candidate has had a pass without a flip (at most `max_iterations`), then
`num_rounds` rounds of at most `final_sweep_iters` passes over a `final_sweep_fraction` share,
each round ending at a pass without flips. Node selection and class order are counter-based hashes
of (seed, pass), so the CPU oracle reproduces them exactly.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import call
from .device import DeviceProblem, ptr, stream

ORDER_SALT = 0x0DE5C1A55C0FFEE1
FINAL_PASS_BASE = 1 << 40
_M = (1 << 64) - 1


def _splitmix64(x: int) -> int:
    z = (x + 0x9E3779B97F4A7C15) & _M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M
    return z ^ (z >> 31)


def _mix3(a: int, b: int, c: int) -> int:
    return _splitmix64(_splitmix64(_splitmix64(a & _M) ^ (b & _M)) ^ (c & _M))


def class_order(seed: int, q: int, num_classes: int) -> list[int]:
    """Order in which pass q visits the colour classes."""
    keys = [_mix3(seed ^ ORDER_SALT, q, c) for c in range(num_classes)]
    return sorted(range(num_classes), key=lambda c: (keys[c], c))


@dataclass
class Colouring:
    colours: torch.Tensor      # int32 [n]
    class_nodes: torch.Tensor  # int32 [n], nodes grouped by colour, ascending within a class
    offsets: list              # class q = class_nodes[offsets[q]:offsets[q+1]]
    rounds: int
    _offsets_dev: torch.Tensor | None = None

    @property
    def num_classes(self) -> int:
        return len(self.offsets) - 1

    def offsets_device(self) -> torch.Tensor:
        """The class offsets as an int32 device tensor (for the persistent search)."""
        if self._offsets_dev is None:
            self._offsets_dev = torch.tensor(self.offsets, dtype=torch.int32).to(self.class_nodes.device)
        return self._offsets_dev


COLOUR_BATCH = 8  # rounds queued between checks of the worklist count


def colour_graph(prob: DeviceProblem, seed: int, largest_first: bool = False) -> Colouring:
    """Greedy colouring on the GPU (descending degree, then hash, then index; largest_first: degree,
    then the smaller index, as local_search_guards/common.py:136-159) in rounds over a shrinking
    worklist of uncoloured nodes; rounds past the last one do nothing."""
    n = prob.n
    colours = torch.empty(n, dtype=torch.int32, device=prob.device)
    lists = torch.empty(3 * n, dtype=torch.int32, device=prob.device)  # two worklists, resume points
    counts = torch.empty(3, dtype=torch.int32, device=prob.device)
    call("sa_colour_start", prob.csr_ref, ptr(colours), ptr(lists), ptr(counts), stream())
    rounds = 0
    while True:
        if largest_first:
            call("sa_colour_rounds_lf", prob.csr_ref, ptr(colours), ptr(lists), ptr(counts), rounds, COLOUR_BATCH,
                 stream())
        else:
            call("sa_colour_rounds", prob.csr_ref, int(seed) & _M, ptr(colours), ptr(lists), ptr(counts), rounds,
                 COLOUR_BATCH, stream())
        rounds += COLOUR_BATCH
        if int(counts[rounds % 3].item()) == 0:
            break
        if rounds > n + COLOUR_BATCH:
            raise RuntimeError("colouring did not finish")
    num = int(colours.max().item()) + 1
    class_nodes = torch.argsort(colours, stable=True).to(torch.int32)
    counts = torch.bincount(colours, minlength=num).cpu().numpy()
    offsets = [0] + np.cumsum(counts).tolist()
    return Colouring(colours, class_nodes, offsets, rounds)


@dataclass
class LocalSearchState:
    spins: torch.Tensor        # int8 [n][kp]
    fields: torch.Tensor       # int32 (integer couplings) or float64 [n][kp]
    groups: int
    k: int
    passes: int                # main passes (the reference's local_search_iterations)
    total_passes: int
    e2_start: np.ndarray       # sum_i s_i h_i of the start candidates (int64, or float64 in units of c)
    e2_final: np.ndarray
    scale: float = 1.0         # c for the floating-point path (h in units of J / c), 1 for integers

    @property
    def exact(self) -> bool:
        return self.e2_start.dtype.kind == "i"

    def _energy(self, e2: np.ndarray) -> np.ndarray:
        if self.exact:
            return -(e2 // 2)
        return -0.5 * self.scale * e2

    @property
    def energies_start(self) -> np.ndarray:
        return self._energy(self.e2_start)

    @property
    def energies_final(self) -> np.ndarray:
        return self._energy(self.e2_final)


def candidates_to_node_major(prob: DeviceProblem, cand_bits: torch.Tensor, k: int) -> tuple[torch.Tensor, int]:
    """Candidate-major spin bits (k x nwords uint64) -> node-major bits (n x groups)."""
    groups = (k + 63) // 64
    nwords = cand_bits.shape[1]
    sb = torch.empty((prob.n, groups), dtype=torch.int64, device=prob.device)
    call("sa_bits_transpose", prob.n, k, ptr(cand_bits), nwords, groups, ptr(sb), stream())
    return sb, groups


def _float_energies(prob: DeviceProblem, spins: torch.Tensor, fields: torch.Tensor, groups: int, k: int) -> np.ndarray:
    part = torch.empty(int(_lib.load().sa_ls_energy_f64_scratch(groups)) // 8, dtype=torch.float64, device=prob.device)
    e2 = torch.empty(64 * groups, dtype=torch.float64, device=prob.device)
    call("sa_ls_energy_f64", prob.n, k, groups, ptr(spins), ptr(fields), ptr(part), ptr(e2), stream())
    return e2[:k].cpu().numpy().copy()


def init_fields(prob: DeviceProblem, cand_bits: torch.Tensor, k: int, fvals: torch.Tensor | None = None):
    """Spins, fields and the start sums e2 = sum_i s_i h_i of k candidates: (spins, fields, groups, e2) with
    e2 an int64 device tensor (integer couplings) or a float64 host array. fvals: FP32 couplings to use
    instead of the problem's (the floating-point path; each field is then summed in row order)."""
    sb, groups = candidates_to_node_major(prob, cand_bits, k)
    kp = 64 * groups
    spins = torch.empty((prob.n, kp), dtype=torch.int8, device=prob.device)
    if prob.jint is not None and fvals is None:
        fields = torch.empty((prob.n, kp), dtype=torch.int32, device=prob.device)
        e2 = torch.zeros(kp, dtype=torch.int64, device=prob.device)
        call("sa_ls_init", prob.csr_ref, k, groups, ptr(sb), ptr(spins), ptr(fields), ptr(e2), stream())
        return spins, fields, groups, e2
    fv = prob.float_couplings() if fvals is None else fvals
    fields = torch.empty((prob.n, kp), dtype=torch.float64, device=prob.device)
    call("sa_ls_init_f64", prob.csr_ref, ptr(fv), k, groups, ptr(sb), ptr(spins), ptr(fields), stream())
    return spins, fields, groups, _float_energies(prob, spins, fields, groups, k)


MAX_PERSISTENT_CLASSES = 1024  # the persistent search ranks the class-order keys in LDS
# Nodes x candidate groups up to which the one-launch search is the faster form (measured on ER graphs of
# mean degree 40: 1.45x faster at n = 1024, even at 2^14, 1.7x slower at 2^16, where each class phase is
# long enough that the grid barriers cost more than separate launches and the per-pass host check).
PERSISTENT_MAX_WORK = 1 << 14


class LocalSearchTimeout(RuntimeError):
    """The persistent search stopped at a grid barrier that did not complete in time."""


def _persistent_search(prob, col, k, groups, seed, spins, fields, fv, max_iterations, subset_fraction,
                       num_rounds, final_sweep_iters, final_sweep_fraction) -> tuple[int, int]:
    """The whole schedule in one cooperative launch (sa_ls_search); (main passes, all passes)."""
    C = col.num_classes
    ctl = torch.zeros(int(_lib.load().sa_ls_search_ctl_ints(C, groups)), dtype=torch.int32, device=prob.device)
    common = (k, groups, ptr(col.class_nodes), ptr(col.offsets_device()), C, seed, int(max_iterations),
              float(subset_fraction), int(num_rounds), int(final_sweep_iters), float(final_sweep_fraction),
              ptr(spins), ptr(fields), ptr(ctl), stream())
    if fv is None:
        call("sa_ls_search", prob.csr_ref, *common)
    else:
        call("sa_ls_search_f64", prob.csr_ref, ptr(fv), *common)
    head = ctl[:8].cpu().numpy()
    if head[7] != 0:
        raise LocalSearchTimeout("the persistent local search stopped: a grid barrier did not complete in time")
    return int(head[4]), int(head[5])


def flip_refinement(prob: DeviceProblem, cand_bits: torch.Tensor, k: int, max_iterations: int = 1000,
                    final_sweep_fraction: float = 1.0, final_sweep_iters: int = 25, num_rounds: int = 25,
                    subset_fraction: float = 0.9, seed: int = 0, colouring: Colouring | None = None,
                    init: tuple | None = None, persistent: bool | None = None) -> LocalSearchState:
    """Local search of k candidates given as spin bits (flip_refinement, :1039-1112). init: the result
    of init_fields(prob, cand_bits, k) when the caller has already computed it. persistent: the whole
    schedule in one cooperative launch (colourings of up to 1024 classes); False: one launch per colour
    class with the schedule on the host; None: the persistent form up to PERSISTENT_MAX_WORK nodes x
    candidate groups. Both forms run the same passes, bit for bit."""
    seed = int(seed) & _M
    spins, fields, groups, e2 = init if init is not None else init_fields(prob, cand_bits, k)
    exact = fields.dtype == torch.int32
    e2_start = e2[:k].cpu().numpy().copy() if exact else np.asarray(e2[:k], dtype=np.float64).copy()
    fv = None if exact else prob.float_couplings()
    col = colouring or colour_graph(prob, seed)
    C = col.num_classes
    if persistent is None:
        persistent = prob.n * groups <= PERSISTENT_MAX_WORK
    if persistent and C <= MAX_PERSISTENT_CLASSES:
        passes, total = _persistent_search(prob, col, k, groups, seed, spins, fields, fv, max_iterations,
                                           subset_fraction, num_rounds, final_sweep_iters, final_sweep_fraction)
        return _finish(prob, spins, fields, groups, k, passes, total, e2, e2_start, exact)
    flags = torch.zeros(64 * groups, dtype=torch.int32, device=prob.device)
    q = [0]

    def run_pass(sel_id: int, frac: float) -> np.ndarray:
        flags.zero_()
        for cls in class_order(seed, q[0], C):
            lo, hi = col.offsets[cls], col.offsets[cls + 1]
            if hi <= lo:
                continue
            nodes = ctypes.c_void_p(col.class_nodes.data_ptr() + 4 * lo)
            if exact:
                call("sa_ls_class_pass", prob.csr_ref, k, groups, nodes, hi - lo, seed, int(sel_id), float(frac),
                     ptr(spins), ptr(fields), ptr(flags), stream())
            else:
                call("sa_ls_class_pass_f64", prob.csr_ref, ptr(fv), k, groups, nodes, hi - lo, seed, int(sel_id),
                     float(frac), ptr(spins), ptr(fields), ptr(flags), stream())
        q[0] += 1
        return flags[:k].cpu().numpy() != 0

    active = np.ones(k, dtype=bool)
    passes = 0
    for it in range(int(max_iterations)):
        passes += 1
        active &= run_pass(it, subset_fraction)
        if not active.any():
            break
    fixed_point = False
    for r in range(int(num_rounds)):
        if fixed_point:
            break
        for _ in range(int(final_sweep_iters)):
            if not run_pass(FINAL_PASS_BASE + r, final_sweep_fraction).any():
                # a pass over every node without a flip is a fixed point of all later passes
                fixed_point = final_sweep_fraction >= 1.0
                break
    return _finish(prob, spins, fields, groups, k, passes, q[0], e2, e2_start, exact)


def _finish(prob, spins, fields, groups, k, passes, total, e2, e2_start, exact) -> LocalSearchState:
    if exact:
        e2.zero_()
        call("sa_ls_energy", prob.n, k, groups, ptr(spins), ptr(fields), ptr(e2), stream())
        e2_final = e2[:k].cpu().numpy().copy()
        return LocalSearchState(spins, fields, groups, k, passes, total, e2_start, e2_final)
    e2_final = _float_energies(prob, spins, fields, groups, k)
    return LocalSearchState(spins, fields, groups, k, passes, total, e2_start, e2_final, scale=prob.c)


def extract(prob: DeviceProblem, state: LocalSearchState, c: int) -> torch.Tensor:
    """Candidate c of the local-search buffer as a +-1 vector in the problem's dtype."""
    out = torch.empty(prob.n, dtype=prob.dtype, device=prob.device)
    call(f"sa_ls_extract_{prob.sfx}", prob.n, state.groups, int(c), ptr(state.spins), ptr(out), stream())
    return out


def energies_of_bits(prob: DeviceProblem, cand_bits: torch.Tensor, k: int) -> np.ndarray:
    """Energies -1/2 s^T J s of k candidate bit vectors: exact int64 for integer couplings, else FP64 from
    the FP32 couplings J / c (fp32_energies, models/spectral_annealing.py:962-976)."""
    if prob.jint is None:
        _, _, _, e2 = init_fields(prob, cand_bits, k)
        return -0.5 * prob.c * e2
    e2 = torch.zeros(k, dtype=torch.int64, device=prob.device)
    call("sa_energy_bits_i64", prob.csr_ref, k, ptr(cand_bits), cand_bits.shape[1], ptr(e2), stream())
    return -(e2.cpu().numpy() // 2)
