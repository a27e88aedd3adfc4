"""Batched SpecAnn sweep: many instances packed block-diagonally and swept together on one GPU.

No counterpart in the reference, which solves one instance per call (utils/core.py loops over
instances). Instance b occupies the packed rows [row_off[b], row_off[b] + n_b), each instance
starting on a multiple of 64 rows (empty padding rows fill the gaps), and is normalised by its own
max absolute degree c_b. Every Lanczos step is one set of launches for all instances: the fused
SpMV over the packed CSR, then projections, update, Ritz solves and restarts per instance (each in
its own work buffer with its own j, status and residual test; see sa_batch_* in specann.h).

Per alpha and instance the control flow is the one-instance sweep's (specann_amd.spectral_annealing,
after models/spectral_annealing.py:311-421): Gershgorin bound and rho from the instance's last
lambda_1, the diagonal (or secant) predictor gated by its step-0 residual, a fall-back start from
y_last when the gate rejects it, a random start in place of a zero predicted vector, one cold retry
of a failed solve, sign alignment, rounding and first-appearance deduplication. Instances move from one
alpha to the next independently (BatchSweep). The difference: the local search runs on the packed graph
with column c holding every instance's c-th distinct candidate (an instance with fewer repeats its last
one), so its pass count is the batch's; per-instance energies are exact int64 sums over the instance's
rows.
"""
from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp
import torch

from . import _lib
from . import local_search as LS
from ._lib import (SA_BROKEN, SA_CONVERGED, SA_MAXITER, SA_REJECTED, SA_RUNNING, SaBatch, SaState, call)
from .device import DeviceProblem, ptr, require_gpu, stream
from .eigsh import default_ncv, default_orth_tau
from .spectral_annealing import compute_shift_rho

TERMINAL = (SA_CONVERGED, SA_MAXITER, SA_BROKEN)
ALIGN = 64


def _head_dtype() -> np.dtype:
    """numpy form of the fields of SaState that lie in its first STATE_HEAD_BYTES bytes."""
    full = np.dtype(SaState)
    names = [k for k in full.names if full.fields[k][1] < _lib.STATE_HEAD_BYTES]
    return np.dtype({"names": names, "formats": [full.fields[k][0] for k in names],
                     "offsets": [full.fields[k][1] for k in names], "itemsize": _lib.STATE_HEAD_BYTES})


HEAD_DTYPE = _head_dtype()


def pack(Js: list[sp.spmatrix]) -> tuple[sp.csr_matrix, np.ndarray, np.ndarray]:
    """Block-diagonal CSR of the instances, each starting on a multiple of 64 rows; returns
    (J_packed, row_off[B+1], sizes[B])."""
    sizes = np.array([J.shape[0] for J in Js], dtype=np.int64)
    padded = (sizes + ALIGN - 1) // ALIGN * ALIGN
    row_off = np.zeros(len(Js) + 1, dtype=np.int64)
    row_off[1:] = np.cumsum(padded)
    blocks = []
    for J, p in zip(Js, padded):
        J = sp.csr_matrix(J)
        if p > J.shape[0]:
            J = sp.block_diag([J, sp.csr_matrix((int(p - J.shape[0]), int(p - J.shape[0])))], format="csr")
        blocks.append(J)
    Jp = sp.block_diag(blocks, format="csr")
    Jp.sort_indices()
    return Jp, row_off, sizes


def inst_arrays(indptr, indices, vals, jint, row_off, sizes, kc: int) -> dict:
    """The per-instance and per-row tables of the instance-staged layout, built with torch on the
    tensors' device: base / real / end (int64 per instance: first row, end of the real columns, end of the
    rows with padding), split ((passes - 1) * n int32: split[(c - 1) n + r] is the first entry of row r
    whose column lies at or past base + c * kc, so pass c reads [split[c - 1][r], split[c][r]) with
    indptr at both ends), and, when every coupling of an instance has one magnitude (signed), ent
    (column | sign << 31 per entry) and mag (the magnitude per instance, vals' type)."""
    dev = indptr.device
    ro = np.asarray(row_off, dtype=np.int64)
    B, n = len(sizes), int(ro[-1])
    passes = int(-(-int(np.max(sizes)) // kc))
    out = {"base": torch.from_numpy(ro[:-1].copy()).to(dev),
           "real": torch.from_numpy(ro[:-1] + np.asarray(sizes, dtype=np.int64)).to(dev),
           "end": torch.from_numpy(ro[1:].copy()).to(dev)}
    deg = torch.diff(indptr.to(torch.int64))
    inst_row = torch.from_numpy(np.repeat(np.arange(B, dtype=np.int64), np.diff(ro))).to(dev)
    row_e = torch.repeat_interleave(torch.arange(n, device=dev), deg)
    inst_e = inst_row[row_e]
    rel = indices.to(torch.int64) - out["base"][inst_e]    # the column's offset in its instance
    split = []
    for c in range(1, passes):
        below = torch.zeros(n, dtype=torch.int64, device=dev)
        below.index_add_(0, row_e, (rel < c * kc).to(torch.int64))
        split.append((indptr[:-1].to(torch.int64) + below).to(torch.int32))
    out["split"] = torch.cat(split) if split else None
    out["passes"] = passes
    mags = torch.abs(jint.to(torch.int64))
    mx = torch.zeros(B, dtype=torch.int64, device=dev).scatter_reduce_(0, inst_e, mags, "amax")
    out["signed"] = bool(torch.all(mags == mx[inst_e]).item())
    if out["signed"]:
        out["ent"] = torch.bitwise_or(indices, (vals < 0).to(torch.int32) << 31)
        first = indptr[torch.from_numpy(ro[:-1].copy()).to(dev)].to(torch.int64)
        out["mag"] = torch.abs(vals[first.clamp(max=max(int(indices.numel()) - 1, 0))]).contiguous()
    return out


class BatchProblem:
    """The packed couplings on the GPU plus the batch descriptor (row and workgroup tables and one
    work buffer per instance)."""

    def __init__(self, Js: list[sp.spmatrix], dtype: torch.dtype = torch.float32, rows_per_block: int = 2048):
        self.device = require_gpu()
        self.B = len(Js)
        if self.B < 1:
            raise ValueError("empty batch")
        Jp, row_off, sizes = pack(Js)
        self.sizes, self.row_off_h = sizes, row_off
        self.c_h = np.array([float(abs(sp.csr_matrix(J)).sum(axis=1).max()) for J in Js])
        if not np.all(self.c_h > 0):
            raise ValueError("an instance has no nonzero coupling")
        row_c = np.ones(Jp.shape[0])
        for b in range(self.B):
            row_c[row_off[b]:row_off[b] + sizes[b]] = self.c_h[b]
        self.prob = DeviceProblem(Jp, dtype, row_c=row_c)
        if self.prob.jint is None:
            raise NotImplementedError("the batched sweep needs integer couplings (exact energies)")
        self.n = self.prob.n
        self.dtype = dtype
        self.rpb = int(rows_per_block)
        if self.rpb % 64 or self.rpb < 64:
            raise ValueError("rows_per_block must be a positive multiple of 64")
        blk = (sizes + self.rpb - 1) // self.rpb
        if blk.max() > 2048:
            # the per-instance reduction partials hold (SA_MAX_NCV + 8) x 2048 values
            raise ValueError(f"an instance needs {blk.max()} workgroups of {self.rpb} rows (at most 2048): "
                             "raise rows_per_block")
        blk_off = np.zeros(self.B + 1, dtype=np.int64)
        blk_off[1:] = np.cumsum(blk)
        self.blk_off_h = blk_off
        self.word_off_h = row_off // 64
        dev = self.device
        self.row_off = torch.from_numpy(row_off).to(dev)
        self.row_end = torch.from_numpy(row_off[:-1] + sizes).to(dev)
        self.blk_off = torch.from_numpy(blk_off.astype(np.int32)).to(dev)
        self.word_off = torch.from_numpy(self.word_off_h.astype(np.int64)).to(dev)
        self.stride = int(_lib.load().sa_work_bytes(0))
        self.works = torch.zeros(self.B * self.stride, dtype=torch.uint8, device=dev)
        self._indptr = Jp.indptr.astype(np.int64)
        self.desc = SaBatch(B=self.B, rows_per_block=self.rpb, row_off=self.row_off.data_ptr(),
                            row_end=self.row_end.data_ptr(), blk_off=self.blk_off.data_ptr(),
                            works=self.works.data_ptr(), stride=self.stride, nblocks=int(blk_off[-1]),
                            tile_inst=0)
        self.set_tiling(self.prob.tile_nnz, self.prob.csr.lanes_per_row)
        self.tuning = None
        self._head = torch.empty((self.B, _lib.STATE_HEAD_BYTES), dtype=torch.uint8, pin_memory=True)

    def set_tiling(self, tile_nnz: int, lanes_per_row: int = 0) -> None:
        """Cut the SpMV tiles per instance (no tile crosses two, so a step skips the tiles of stopped
        instances): about tile_nnz entries each; lanes_per_row 0 picks the lanes per row from each tile's
        mean row length, 1..64 fixes them."""
        if tile_nnz not in (512, 1024, 2048, 4096) or not 0 <= int(lanes_per_row) <= 64:
            raise ValueError("tile_nnz must be 512, 1024, 2048 or 4096 and lanes_per_row in [0, 64]")
        indptr, row_off = self._indptr, self.row_off_h
        starts, inst = [], []
        for b in range(self.B):
            r0, r1 = int(row_off[b]), int(row_off[b + 1])
            e0, e1 = indptr[r0], indptr[r1]
            targets = e0 + np.arange(0, max(int(e1 - e0), 1), tile_nnz, dtype=np.int64)
            rows = r0 + np.searchsorted(indptr[r0:r1 + 1], targets, side="left")
            rows[0] = r0
            starts.append(np.minimum(rows, r1))
            inst.append(np.full(rows.size, b, dtype=np.int32))
        tile_rows = np.concatenate(starts + [np.array([self.n], dtype=np.int64)]).astype(np.int32)
        self.tile_off_h = np.concatenate([[0], np.cumsum([t.size for t in starts])]).astype(np.int64)
        self.unit_start, self.unit_count = self.tile_off_h[:-1], np.diff(self.tile_off_h)
        self.layout = {"kind": "tiles", "tile_nnz": int(tile_nnz), "lanes_per_row": int(lanes_per_row)}
        self.prob.csr.imode = 0
        dev = self.device
        self.tile_rows = torch.from_numpy(tile_rows).to(dev)
        self.tile_inst = torch.from_numpy(np.concatenate(inst)).to(dev)
        prob = self.prob
        prob.tile_nnz = int(tile_nnz)
        prob.tile_rows = self.tile_rows
        prob.ntiles = int(tile_rows.size - 1)
        prob.csr.tile_rows = self.tile_rows.data_ptr()
        prob.csr.ntiles = prob.ntiles
        prob.csr.tile_nnz = int(tile_nnz)
        prob.csr.lanes_per_row = int(lanes_per_row)
        self.desc.tile_inst = self.tile_inst.data_ptr()

    def inst_eligible(self) -> bool:
        """Whether every instance fits the instance-staged layout (at most 4 passes of staged columns)."""
        kc = int(_lib.load().sa_inst_columns_per_pass(self.prob.vals.element_size()))
        return int(self.sizes.max()) <= 4 * kc

    def set_inst_layout(self, rows: int = 2048, lanes: int = 8) -> None:
        """Switch the SpMV to the instance-staged layout (sa_csr imode 1): workgroups of `rows` rows of
        one instance, the biggest instances' first, each staging its instance's slice of the gathered
        vector in LDS and summing every row with `lanes` lanes."""
        if not 64 <= rows <= 2048 or rows % 64 or lanes not in (4, 8, 16, 32):
            raise ValueError("rows must be a multiple of 64 in [64, 2048] and lanes 4, 8, 16 or 32")
        if not self.inst_eligible():
            raise ValueError("an instance is too large for the instance-staged layout")
        prob, dev = self.prob, self.device
        kc = int(_lib.load().sa_inst_columns_per_pass(prob.vals.element_size()))
        ro = self.row_off_h
        order = np.argsort(-self.sizes, kind="stable")
        wg_inst, wg_row = [], []
        start = np.zeros(self.B, dtype=np.int64)
        count = np.zeros(self.B, dtype=np.int64)
        for b in order:
            rs = np.arange(ro[b], ro[b + 1], rows, dtype=np.int64)
            start[b] = len(wg_row)
            count[b] = rs.size
            wg_row.extend(rs.tolist())
            wg_inst.extend([int(b)] * rs.size)
        passes = int(-(-int(self.sizes.max()) // kc))
        cache = self._inst_cache = getattr(self, "_inst_cache", None) or {}
        if "base" not in cache:
            cache.update(inst_arrays(prob.indptr, prob.indices, prob.vals, prob.jint, ro, self.sizes, kc))
        cache["iinst"] = torch.tensor(wg_inst, dtype=torch.int32, device=dev)
        cache["irow"] = torch.tensor(wg_row, dtype=torch.int32, device=dev)
        csr = prob.csr
        csr.imode = 1
        csr.irows, csr.ilanes, csr.ipasses = int(rows), int(lanes), passes
        csr.isigned = int(cache["signed"])
        csr.inwg = len(wg_row)
        csr.iinst, csr.irow = cache["iinst"].data_ptr(), cache["irow"].data_ptr()
        csr.ibase, csr.ireal, csr.iend = cache["base"].data_ptr(), cache["real"].data_ptr(), cache["end"].data_ptr()
        csr.isplit = cache["split"].data_ptr() if cache["split"] is not None else None
        csr.ient = cache["ent"].data_ptr() if cache["signed"] else None
        csr.imag = cache["mag"].data_ptr() if cache["signed"] else None
        self.unit_start, self.unit_count = start, count
        self.layout = {"kind": "instance-staged", "rows": int(rows), "lanes": int(lanes), "passes": passes,
                       "signed": cache["signed"]}

    def istream_eligible(self) -> bool:
        """Whether the instance-streamed layout applies: FP32, and every instance within its chunks'
        16-bit column offsets."""
        return self.dtype == torch.float32 and int(self.sizes.max()) <= 65536

    def set_istream_layout(self, rows: int = 20416) -> bool:
        """Switch the SpMV to the instance-streamed layout (sa_csr imode 2): workgroups of at most `rows`
        rows of one instance, each streaming its rows' entries sorted by column (the column-streamed
        chunks, built with these row blocks and one panel) with FP64 row sums in LDS. False (layout
        unchanged) when it does not apply."""
        prob, dev = self.prob, self.device
        cap = int(_lib.load().sa_cstream_max_rows())
        if not self.istream_eligible() or not 64 <= rows <= cap or rows % 64:
            return False
        cache = self._istream_cache = getattr(self, "_istream_cache", None) or {}
        if cache.get("rows") != rows:
            ro = self.row_off_h
            wg_row, wg_inst, start, count = [], [], np.zeros(self.B, np.int64), np.zeros(self.B, np.int64)
            for b in range(self.B):
                rs = np.arange(ro[b], ro[b + 1], rows, dtype=np.int64)
                start[b], count[b] = len(wg_row), rs.size
                wg_row.extend(rs.tolist())
                wg_inst.extend([b] * rs.size)
            rbound = np.asarray(wg_row + [self.n], dtype=np.int64)
            bounds = np.array([0] + [self.n] * 8, dtype=np.int64)   # every column in panel 0
            arr = prob.cstream_arrays(rows, len(wg_row), torch.from_numpy(rbound.astype(np.int32)).to(dev), bounds,
                                      signed=False)
            if arr is None:
                return False
            iinst = torch.tensor(wg_inst, dtype=torch.int32, device=dev)
            mag = self._istream_sign_pack(arr, iinst)
            cache.clear()
            cache.update(arr)
            cache.update(rows=rows, start=start, count=count, nwg=len(wg_row), iinst=iinst, mag=mag,
                         irow=torch.tensor(wg_row, dtype=torch.int32, device=dev),
                         iend=torch.from_numpy(ro[1:].astype(np.int64)).to(dev))
        csr = prob.csr
        csr.imode = 2
        csr.irows, csr.inwg = int(rows), int(cache["nwg"])
        csr.iinst, csr.irow, csr.iend = (cache["iinst"].data_ptr(), cache["irow"].data_ptr(),
                                         cache["iend"].data_ptr())
        signed = cache["mag"] is not None
        csr.isigned = int(signed)
        csr.imag = cache["mag"].data_ptr() if signed else None
        csr.bchunk, csr.bent, csr.bcbase = (cache["bchunk"].data_ptr(), cache["bent"].data_ptr(),
                                            cache["bcbase"].data_ptr())
        csr.bvals = None if signed else cache["bvals"].data_ptr()
        self.unit_start, self.unit_count = cache["start"], cache["count"]
        self.layout = {"kind": "instance-streamed", "rows": int(rows), "chunks": int(cache["nchunks"]),
                       "signed": signed}
        return True

    def _istream_sign_pack(self, arr: dict, iinst: torch.Tensor) -> torch.Tensor | None:
        """When every coupling of each instance has one magnitude, fold the signs of the instance-streamed
        layout's values into its entries (bit 31) and drop the values: returns the per-instance magnitudes
        (float32, B) and updates arr in place; None (arr unchanged) otherwise. Padding slots (local row =
        sa_cstream_pad_row(), value 0) keep a clear sign bit."""
        pad = int(_lib.load().sa_cstream_pad_row())
        nch = int(arr["nchunks"])
        if nch == 0:
            return None
        bent = arr["bent"][: 256 * nch]
        vals = arr["bvals"][: 256 * nch]
        real = ((bent >> 16) & 0x7FFF) != pad
        # the instance of each chunk: chunk k lies in segment w of bchunk (8 segments per workgroup)
        seg = torch.searchsorted(arr["bchunk"][1:].to(torch.int64), torch.arange(nch, device=bent.device),
                                 right=True)
        inst_e = iinst[(seg // 8).clamp(max=iinst.numel() - 1)].to(torch.int64).repeat_interleave(256)
        a = torch.abs(vals)
        big = torch.full((self.B,), -1.0, dtype=torch.float32, device=bent.device)
        big.scatter_reduce_(0, inst_e[real], a[real], "amax")
        small = torch.full((self.B,), float("inf"), dtype=torch.float32, device=bent.device)
        small.scatter_reduce_(0, inst_e[real], a[real], "amin")
        seen = big >= 0
        if not bool(torch.all(~seen | (big == small)).item()) or not bool(torch.all(a[real] > 0).item()):
            return None
        mag = torch.where(seen, big, torch.ones_like(big)).contiguous()
        arr["bent"][: 256 * nch] = torch.where(real & (vals < 0), bent | torch.tensor(-(1 << 31), dtype=torch.int32,
                                                                                        device=bent.device), bent)
        arr["bvals"] = None
        return mag

    def autotune(self, reps: int = 4, verbose: bool = False) -> dict:
        """Pick the batched SpMV's layout by timing the packed matrix's SpMV on each: plain per-instance
        tiles (tile size, lanes per row), when every instance fits, the instance-staged layout (rows per
        workgroup, lanes per row), and in FP32 the instance-streamed layout (rows per workgroup)."""
        results = []
        for tile in (512, 1024, 2048, 4096):
            for lanes in (0, 4, 8, 16, 32, 64):
                self.set_tiling(tile, lanes)
                sec = self.prob.time_spmv(reps)
                results.append({"kind": "tiles", "tile_nnz": tile, "lanes_per_row": lanes, "seconds": sec})
                if verbose:
                    print(f"[batch autotune] tile {tile} lanes {lanes}: {sec * 1e6:.1f} us", flush=True)
        if self.inst_eligible():
            for rows in (1024, 2048):
                for lanes in (4, 8, 16):
                    self.set_inst_layout(rows, lanes)
                    sec = self.prob.time_spmv(reps)
                    results.append({"kind": "instance-staged", "rows": rows, "lanes": lanes, "seconds": sec})
                    if verbose:
                        print(f"[batch autotune] instance-staged rows {rows} lanes {lanes}: {sec * 1e6:.1f} us",
                              flush=True)
        if self.istream_eligible():
            for rows in (4096, 20416):
                if self.set_istream_layout(rows):
                    sec = self.prob.time_spmv(reps)
                    results.append({"kind": "instance-streamed", "rows": rows, "seconds": sec})
                    if verbose:
                        print(f"[batch autotune] instance-streamed rows {rows}: {sec * 1e6:.1f} us", flush=True)
        best = min(results, key=lambda r: r["seconds"])
        if best["kind"] == "tiles":
            self.set_tiling(best["tile_nnz"], best["lanes_per_row"])
        elif best["kind"] == "instance-staged":
            self.set_inst_layout(best["rows"], best["lanes"])
        else:
            self.set_istream_layout(best["rows"])
        self.tuning = {"choice": best, "tried": results}
        return self.tuning

    @property
    def ref(self):
        return ctypes.byref(self.desc)

    def heads(self) -> np.ndarray:
        """The scalar heads of every instance's state, B records of HEAD_DTYPE (one copy, one
        synchronisation of the current stream)."""
        hb = _lib.STATE_HEAD_BYTES
        view = self.works.view(self.B, self.stride)[:, :hb]
        self._head.copy_(view, non_blocking=True)
        torch.cuda.current_stream().synchronize()
        return self._head.numpy().reshape(-1).view(HEAD_DTYPE).copy()

    def rows_of(self, b: int) -> slice:
        return slice(int(self.row_off_h[b]), int(self.row_off_h[b] + self.sizes[b]))

    def words_of(self, b: int) -> slice:
        return slice(int(self.word_off_h[b]), int(self.word_off_h[b] + (self.sizes[b] + 63) // 64))


@dataclass
class InstanceResult:
    best_energy: float
    best_spins: np.ndarray | None
    best_alpha: float | None
    rows: list = field(default_factory=list)
    num_candidates: int = 0
    raw_energies: list = field(default_factory=list)
    refined_energies: list = field(default_factory=list)


@dataclass
class BatchOutcome:
    results: list
    sweep_time: float
    post_sweep_time: float
    steps: int                 # batched Lanczos steps (launch sets)
    matvecs: int               # operator applications summed over instances
    local_passes: int
    packed_candidates: object = None   # device int64 [K][nwords]: column c = every instance's c-th candidate
    ls_state: object = None


class _Buffers:
    """Device buffers of one batched sweep."""

    def __init__(self, bp: BatchProblem, ncv_arg, num_alphas: int):
        dev, dt, n = bp.device, bp.dtype, bp.n
        self.ncv = default_ncv(int(bp.sizes.min()), ncv_arg)
        self.keep = max(1, min(_lib.SA_MAX_KEEP, self.ncv // 2, self.ncv - 1))
        self.ldv = _lib.padded_ld(n, 4 if dt == torch.float32 else 8)
        self.V = torch.zeros((self.ncv + 1, self.ldv), dtype=dt, device=dev)
        self.Z, self.W, self.Y, self.s, self.x_rand = (torch.zeros(n, dtype=dt, device=dev) for _ in range(5))
        self.nwords = (n + 63) // 64
        self.bounds = torch.empty(bp.B, dtype=torch.float64, device=dev)


class BatchSweep:
    """The alpha sweep of SpectralAnnealing (diagonal, secant or prev predictor) over a batch.

    Two schedules give every instance the same arithmetic. ``lockstep=True`` moves the whole batch from
    one alpha to the next together, so every alpha waits for its slowest instance. The default moves
    each instance on as soon as its own solve stops: at each check (a host read of the states every
    ``chunk`` batched steps) the stopped instances take their Ritz vector, rounding and next start
    together (those whose predicted start the gate rejected restart from y_last), as one masked set of
    launches with
    per-instance alphas and predictor exponents, while the rest keep stepping. The batched step count
    then follows the instance with the most Lanczos steps over the whole sweep instead of the sum over
    alphas of the slowest instance per alpha.

    The default schedule also retries a failed solve once from a random vector (breakdown, a
    Ritz value above the Gershgorin bound, a zero or non-finite Ritz vector) and replaces a zero
    predicted start by a random gated one, as the one-instance sweep does
    (models/spectral_annealing.py:568-669); the lockstep schedule skips the alpha instead.
    """

    def __init__(self, random_seed: int = 0, num_alphas: int = 128, schedule: str = "sqrt",
                 schedule_interval=(0.0, 1.0), pred_mode: str = "diagonal", pred_tol: float = 1e-3,
                 shift_eps: float = 1e-3, eigsh_tol: float = 1e-6, eigsh_maxiter: int = 1000,
                 eigsh_ncv: int | None = None, max_local_iterations: int = 1000, use_local_search: bool = True,
                 chunk: int = 8, lockstep: bool = False):
        if pred_mode not in ("diagonal", "secant", "prev"):
            raise ValueError("the batched sweep supports pred_mode diagonal, secant or prev")
        self.seed = 0 if random_seed is None else int(random_seed)
        t0, t1 = (float(v) for v in schedule_interval)
        t = np.linspace(t0, t1, int(num_alphas))
        self.alphas = np.sqrt(t) if schedule == "sqrt" else t
        self.pred_mode, self.pred_tol, self.shift_eps = pred_mode, float(pred_tol), float(shift_eps)
        self.tol, self.maxiter = float(eigsh_tol), int(eigsh_maxiter)
        self.ncv_arg = eigsh_ncv
        self.max_local_iterations = int(max_local_iterations)
        self.use_local_search = bool(use_local_search)
        self.chunk = int(chunk)
        if self.chunk < 1:
            raise ValueError("chunk must be >= 1")
        self.lockstep = bool(lockstep)
        self.retries = 0           # cold retries of the last run (default schedule)
        self.fallbacks = 0         # zero predicted starts replaced by random ones (default schedule)

    def run(self, bp: BatchProblem) -> BatchOutcome:
        return self._run_lockstep(bp) if self.lockstep else self._run_async(bp)

    def _ok(self, st, theta, g, ai):
        """The one-instance sweep's acceptance of a stopped solve (:679-691): no breakdown, a finite Ritz
        value within the Gershgorin bound (arrays over the batch; ai: each instance's alpha index)."""
        return (st != SA_BROKEN) & np.isfinite(theta) & (theta <= g * (1.0 + max(0.5 * self.shift_eps, 1e-5)))

    def _row(self, bp, b, ai, hs, g, rho, dsum, ok, pred_res):
        theta = float(hs['theta'][b])
        return {'alpha': float(self.alphas[ai]), 'alpha_index': int(ai), 'eigsh_iterations': int(hs['steps'][b]),
                'eigsh_converged': bool(hs['status'][b] == SA_CONVERGED and ok), 'lambda1': theta,
                'rho': float(rho), 'gershgorin': float(g),
                'energy_bound': float(-0.5 * theta * dsum * bp.c_h[b]) if ok else float('nan'),
                'pred_residual': float(pred_res),
                'pred_accepted': (None if np.isnan(pred_res) else bool(not pred_res > self.pred_tol)),
                'residual_estimate': float(hs['res'][b]), 'lanczos_restarts': int(hs['cycles'][b]),
                'failed': not bool(ok), 'retry': False, 'candidate': None}

    # ---- every instance at the same alpha ------------------------------------------------------
    def _run_lockstep(self, bp: BatchProblem) -> BatchOutcome:
        t_start = time.time()
        prob, B, sfx = bp.prob, bp.B, bp.prob.sfx
        dev, dt, n = bp.device, bp.dtype, bp.n
        buf = _Buffers(bp, self.ncv_arg, len(self.alphas))
        ncv, keep, ldv, V = buf.ncv, buf.keep, buf.ldv, buf.V
        Z, W, Y, s, x_rand, bounds, nwords = buf.Z, buf.W, buf.Y, buf.s, buf.x_rand, buf.bounds, buf.nwords
        ybuf = [torch.zeros(n, dtype=dt, device=dev) for _ in range(3)]
        A = len(self.alphas)
        cand_bits = torch.zeros((A, nwords), dtype=torch.int64, device=dev)
        R = bp.ref

        def step_all():
            call(f"sa_batch_lanczos_step_{sfx}", prob.csr_ref, R, ptr(s), ptr(V), ldv, ptr(Z), ptr(W), stream())
            call(f"sa_batch_lanczos_thick_restart_{sfx}", R, ptr(s), ptr(V), ldv, ptr(Z), stream())

        def start(x, **kw):
            call(f"sa_batch_lanczos_start_{sfx}", R, ptr(s), ptr(x), ptr(kw.get("d")), float(kw.get("pexp", 0.0)),
                 float(kw.get("cx", 1.0)), ptr(kw.get("y")), float(kw.get("cy", 0.0)), ptr(V), ldv, ptr(Z),
                 int(kw.get("gate", 0)), int(kw.get("only_if", -1)), stream())

        def begin(pred_tol):
            call("sa_batch_lanczos_begin", R, ncv, self.maxiter, self.tol, ptr(bounds), float(pred_tol), keep,
                 default_orth_tau(bp.dtype), stream())

        def run_until_done(first_check: int):
            """Queue steps until every instance stopped; restart REJECTED ones from y_last. Returns heads
            and each instance's first_beta (the predictor residual, read at the first check)."""
            check = first_check
            first = None
            steps = 0
            limit = self.maxiter + 8 * ncv + 64
            while True:
                for _ in range(check):
                    step_all()
                    steps += 1
                hs = bp.heads()
                if first is None:
                    first = hs['first_beta'].copy()
                stats = hs['status']
                if np.any(stats == SA_REJECTED):
                    start(y_last, gate=0, only_if=SA_REJECTED)
                    hs = bp.heads()
                    stats = hs['status']
                if np.all(np.isin(stats, TERMINAL)):
                    return hs, first, steps
                if steps > limit:
                    raise RuntimeError(f"batched Lanczos did not stop after {steps} steps")
                check = self.chunk

        results = [InstanceResult(float("inf"), None, None) for _ in range(B)]
        lam_prev = np.full(B, np.nan)
        y_last = y_prev = None
        alpha_prev = None
        total_steps = total_matvecs = 0
        by_hash = [dict() for _ in range(B)]
        cand_alpha = [[] for _ in range(B)]       # per instance: alpha slot of each distinct candidate
        cand_of_row = [[] for _ in range(B)]
        for ai, alpha in enumerate(self.alphas):
            call(f"sa_batch_alpha_scaling_{sfx}", R, ptr(prob.d), float(alpha), 1e-12, ptr(s), stream())
            hs = bp.heads()
            dsum = hs['r'][:, 1].copy()
            g = hs['r'][:, 2].copy()
            rho = np.array([compute_shift_rho(None if np.isnan(lam_prev[b]) else lam_prev[b], g[b], self.shift_eps)
                            for b in range(B)])
            bounds.copy_(torch.from_numpy(rho + g))
            cold = ai == 0 or y_last is None
            if cold:
                begin(np.inf)
                prob.fill_normal(self.seed, ai * 4, x_rand)
                start(x_rand)
                hs, first, steps = run_until_done(self.chunk)
                pred_res = np.full(B, np.nan)
            else:
                if self.pred_mode == "prev":
                    begin(np.inf)
                    start(y_last)
                    hs, first, steps = run_until_done(self.chunk)
                    pred_res = np.full(B, np.nan)
                else:
                    begin(self.pred_tol)
                    if self.pred_mode == "diagonal" or y_prev is None:
                        start(y_last, d=prob.d, pexp=0.5 * (alpha - alpha_prev), gate=1)
                    else:
                        start(y_last, cx=2.0, y=y_prev, cy=-1.0, gate=1)
                    hs, first, steps = run_until_done(1)
                    pred_res = first
            total_steps += steps
            stats, theta = hs['status'], hs['theta']
            total_matvecs += int(hs['steps'].sum())
            ok = self._ok(stats, theta, g, np.full(B, ai))
            call(f"sa_batch_lanczos_ritz_{sfx}", R, ptr(V), ldv, ptr(y_last), ptr(Y), -1, stream())
            free = next(b for b in ybuf if b is not y_last and b is not y_prev)
            call(f"sa_batch_round_{sfx}", R, ptr(Y), ptr(free), ptr(cand_bits[ai]), int(y_last is not None), stream())
            if y_last is not None and not ok.all():
                for b in np.nonzero(~ok)[0]:          # a failed instance keeps its previous vector
                    free[bp.rows_of(b)].copy_(y_last[bp.rows_of(b)])
            bits_h = cand_bits[ai].cpu().numpy()
            for b in range(B):
                row = self._row(bp, b, ai, hs, g[b], rho[b], dsum[b], ok[b], pred_res[b])
                if ok[b]:
                    key = bits_h[bp.words_of(b)].tobytes()
                    c = by_hash[b].get(key)
                    if c is None:
                        c = len(cand_alpha[b])
                        by_hash[b][key] = c
                        cand_alpha[b].append(ai)
                    row['candidate'] = c
                    lam_prev[b] = theta[b]
                results[b].rows.append(row)
                cand_of_row[b].append(row['candidate'])
            y_prev, y_last = y_last, free
            alpha_prev = alpha
        torch.cuda.synchronize()
        return self._post_sweep(bp, results, cand_alpha, cand_of_row, cand_bits, time.time() - t_start,
                                total_steps, total_matvecs)

    # ---- every instance at its own alpha ---------------------------------------------------------
    def _run_async(self, bp: BatchProblem) -> BatchOutcome:
        t_start = time.time()
        prob, B, sfx = bp.prob, bp.B, bp.prob.sfx
        dev, dt, n = bp.device, bp.dtype, bp.n
        buf = _Buffers(bp, self.ncv_arg, len(self.alphas))
        ncv, keep, ldv, V = buf.ncv, buf.keep, buf.ldv, buf.V
        Z, W, Y, s, x_rand, nwords = buf.Z, buf.W, buf.Y, buf.s, buf.x_rand, buf.nwords
        Ylast, Yprev, Ynew = (torch.zeros(n, dtype=dt, device=dev) for _ in range(3))
        bits = torch.zeros(nwords, dtype=torch.int64, device=dev)
        bits_h = torch.empty(nwords, dtype=torch.int64, pin_memory=True)
        alphas, A = self.alphas, len(self.alphas)
        cand_h = np.zeros((A, nwords), dtype=np.int64)
        orth_tau = default_orth_tau(bp.dtype)
        R = bp.ref
        keep_alive = []                                   # device arrays of the launches queued this check

        def dev_array(values, dtype):
            t = torch.from_numpy(np.ascontiguousarray(values, dtype=dtype)).to(dev)
            keep_alive.append(t)
            return t.data_ptr()

        def masked(sel):
            d = SaBatch.from_buffer_copy(bp.desc)
            d.inst_mask = dev_array(sel, np.uint8)
            return ctypes.byref(d)

        ai = np.zeros(B, dtype=np.int64)          # alpha index each instance is solving (A: finished)
        gated = np.zeros(B, dtype=bool)           # a gated start whose first_beta is still to be read
        predicted = np.zeros(B, dtype=bool)       # the current solve started from a predictor
        retried = np.zeros(B, dtype=bool)
        fallback = np.zeros(B, dtype=bool)
        pred_res = np.full(B, np.nan)
        lam_prev = np.full(B, np.nan)
        dsum, g, rho = np.zeros(B), np.zeros(B), np.zeros(B)
        started = np.zeros(B, dtype=np.int64)     # batched step count at the current solve's start
        matvecs = np.zeros(B, dtype=np.int64)
        results = [InstanceResult(float("inf"), None, None) for _ in range(B)]
        by_hash = [dict() for _ in range(B)]
        cand_alpha = [[] for _ in range(B)]
        cand_of_row = [[] for _ in range(B)]
        self.retries = self.fallbacks = 0
        steps = 0
        limit = self.maxiter + 8 * ncv + 64

        step_desc = SaBatch.from_buffer_copy(bp.desc)
        step_ref = ctypes.byref(step_desc)
        tile_list = None

        def step_all():
            call(f"sa_batch_lanczos_step_{sfx}", prob.csr_ref, step_ref, ptr(s), ptr(V), ldv, ptr(Z), ptr(W), stream())
            call(f"sa_batch_lanczos_thick_restart_{sfx}", R, ptr(s), ptr(V), ldv, ptr(Z), stream())

        def begin(sel, pred_tol):
            call("sa_batch_lanczos_begin", masked(sel), ncv, self.maxiter, self.tol, dev_array(rho + g, np.float64),
                 float(pred_tol), keep, orth_tau, stream())
            started[sel] = steps

        def start(sel, x, cx=1.0, y=None, cy=0.0, gate=0):
            call(f"sa_batch_lanczos_start_{sfx}", masked(sel), ptr(s), ptr(x), None, 0.0, float(cx), ptr(y),
                 float(cy), ptr(V), ldv, ptr(Z), int(gate), -1, stream())

        def scale(sel):
            """s at each selected instance's alpha, then its degree sum, Gershgorin bound and shift."""
            call(f"sa_batch_alpha_scaling_each_{sfx}", masked(sel), ptr(prob.d),
                 dev_array(alphas[np.minimum(ai, A - 1)], np.float64), 1e-12, ptr(s), stream())
            hs = bp.heads()
            keep_alive.clear()
            dsum[sel] = hs['r'][sel, 1]
            g[sel] = hs['r'][sel, 2]
            for b in np.nonzero(sel)[0]:
                rho[b] = compute_shift_rho(None if np.isnan(lam_prev[b]) else lam_prev[b], g[b], self.shift_eps)

        def random_start(sel, offset, gate):
            """Random start vectors (stream alpha_index * 4 + offset, as the one-instance sweep draws)."""
            for a in np.unique(ai[sel]):
                grp = sel & (ai == a)
                prob.fill_normal(self.seed, int(a) * 4 + offset, x_rand)
                begin(grp, self.pred_tol if gate else np.inf)
                start(grp, x_rand, gate=gate)
            gated[sel] = bool(gate)

        def predicted_start(sel):
            if self.pred_mode == "prev":
                begin(sel, np.inf)
                start(sel, Ylast)
                return
            begin(sel, self.pred_tol)
            diag = sel & ((ai == 1) if self.pred_mode == "secant" else np.ones(B, dtype=bool))
            if diag.any():
                pexp = 0.5 * (alphas[np.minimum(ai, A - 1)] - alphas[np.maximum(ai - 1, 0)])
                call(f"sa_batch_lanczos_start_each_{sfx}", masked(diag), ptr(s), ptr(Ylast), ptr(prob.d),
                     dev_array(pexp, np.float64), ptr(V), ldv, ptr(Z), 1, stream())
            sec = sel & ~diag
            if sec.any():
                start(sec, Ylast, cx=2.0, y=Yprev, cy=-1.0, gate=1)
            gated[sel] = True
            predicted[sel] = True

        every = np.ones(B, dtype=bool)
        scale(every)
        random_start(every, 0, 0)
        interval = self.chunk
        sweeping = B
        while np.any(ai < A):
            if int(np.count_nonzero(ai < A)) != sweeping:
                # the SpMV launches only the tiles of the instances still sweeping
                sweeping = int(np.count_nonzero(ai < A))
                us, uc = bp.unit_start, bp.unit_count
                tiles = np.concatenate([np.arange(us[b], us[b] + uc[b], dtype=np.int32) for b in np.nonzero(ai < A)[0]])
                tile_list = torch.from_numpy(tiles).to(dev)
                step_desc.tile_list = tile_list.data_ptr()
                step_desc.ntile_list = int(tiles.size)
            for _ in range(interval):
                step_all()
                steps += 1
            hs = bp.heads()
            keep_alive.clear()
            st = hs['status']
            live = ai < A
            rd = gated & live
            pred_res[rd] = hs['first_beta'][rd]       # kept in the state until the next begin
            gated[rd] = False
            over = live & (st == SA_RUNNING) & (steps - started > limit)
            if over.any():
                raise RuntimeError(f"batched Lanczos: instance {int(np.argmax(over))} did not stop after "
                                   f"{limit} steps")
            if np.any(live & (st == SA_REJECTED)):
                call(f"sa_batch_lanczos_start_{sfx}", R, ptr(s), ptr(Ylast), None, 0.0, 1.0, None, 0.0, ptr(V), ldv,
                     ptr(Z), 0, SA_REJECTED, stream())
            term = live & np.isin(st, TERMINAL)
            if term.any():
                theta = hs['theta']
                zero = term & (st == SA_BROKEN) & predicted & ~fallback & ~(hs['r'][:, 3] >= 1e-12)
                term &= ~zero
                ok = term & self._ok(st, theta, g, ai)
                vec_ok = np.zeros(B, dtype=bool)
                if ok.any():
                    m_ok = masked(ok)
                    call(f"sa_batch_lanczos_ritz_{sfx}", m_ok, ptr(V), ldv, ptr(Ylast), ptr(Y), -1, stream())
                    call(f"sa_batch_round_{sfx}", m_ok, ptr(Y), ptr(Ynew), ptr(bits), 1, stream())
                    bits_h.copy_(bits, non_blocking=True)
                    h2 = bp.heads()
                    keep_alive.clear()
                    nrm, dot = h2['r'][:, 4], h2['r'][:, 5]
                    vec_ok = ok & (nrm > 0) & np.isfinite(nrm) & np.isfinite(dot)
                    words = bits_h.numpy()
                retry = term & ~vec_ok & ~retried
                moved = term & ~retry
                matvecs += np.where(term, hs['steps'], 0)
                for b in np.nonzero(moved)[0]:
                    row = self._row(bp, b, ai[b], hs, g[b], rho[b], dsum[b], vec_ok[b], pred_res[b])
                    row['retry'] = bool(retried[b])
                    if vec_ok[b]:
                        ws = bp.words_of(b)
                        key = words[ws].tobytes()
                        c = by_hash[b].get(key)
                        if c is None:
                            c = len(cand_alpha[b])
                            by_hash[b][key] = c
                            cand_alpha[b].append(int(ai[b]))
                            cand_h[ai[b], ws] = words[ws]
                        row['candidate'] = c
                        lam_prev[b] = theta[b]
                    results[b].rows.append(row)
                    cand_of_row[b].append(row['candidate'])
                if moved.any():                          # y_prev <- y_last; y_last <- the new vector
                    call(f"sa_batch_copy_{sfx}", masked(moved), ptr(Ylast), ptr(Yprev), stream())
                if vec_ok.any():
                    call(f"sa_batch_copy_{sfx}", masked(vec_ok), ptr(Ynew), ptr(Ylast), stream())
                if zero.any():
                    self.fallbacks += int(zero.sum())
                    fallback |= zero
                    predicted[zero] = False
                    random_start(zero, 1, 1)
                if retry.any():
                    self.retries += int(retry.sum())
                    retried |= retry
                    predicted[retry] = False
                    random_start(retry, 2, 0)
                ai[moved] += 1
                retried[moved] = fallback[moved] = predicted[moved] = False
                pred_res[moved] = np.nan
                nxt = moved & (ai < A)
                if nxt.any():
                    scale(nxt)
                    predicted_start(nxt)
            interval = self.chunk
        torch.cuda.synchronize()
        sweep_time = time.time() - t_start
        cand_bits = torch.from_numpy(cand_h).to(dev)
        return self._post_sweep(bp, results, cand_alpha, cand_of_row, cand_bits, sweep_time, steps,
                                int(matvecs.sum()))

    def _post_sweep(self, bp, results, cand_alpha, cand_of_row, cand_bits, sweep_time, total_steps, total_matvecs):
        prob, B, dev, R = bp.prob, bp.B, bp.device, bp.ref
        nwords = cand_bits.shape[1]
        # packed candidates, exact energies, local search, selection
        t_post = time.time()
        K = max(len(c) for c in cand_alpha)
        passes = 0
        if K == 0:
            return BatchOutcome(results, sweep_time, 0.0, total_steps, total_matvecs, 0)
        src = np.zeros((K, B), dtype=np.int32)
        for b in range(B):
            ca = cand_alpha[b] or [0]
            for c in range(K):
                src[c, b] = ca[min(c, len(ca) - 1)]
        src_d = torch.from_numpy(src).to(dev)
        packed = torch.zeros((K, nwords), dtype=torch.int64, device=dev)
        call("sa_bits_gather", nwords, K, B, ptr(bp.word_off), ptr(src_d), ptr(cand_bits), ptr(packed), stream())
        # exact start energies per instance from the local search's initial spins and fields (one pass
        # over the couplings for 64 candidates at a time; the same int64 sums as the candidate-bit form)
        init = LS.init_fields(prob, packed, K)
        e2 = torch.zeros((B, K), dtype=torch.int64, device=dev)
        call("sa_batch_ls_energy", R, K, init[2], ptr(init[0]), ptr(init[1]), ptr(e2), stream())
        raw = -(e2.cpu().numpy() // 2)
        refined = raw.copy()
        state = None
        if self.use_local_search:
            state = LS.flip_refinement(prob, packed, K, max_iterations=self.max_local_iterations, seed=self.seed,
                                       init=init)
            passes = state.passes
            e2.zero_()
            call("sa_batch_ls_energy", R, K, state.groups, ptr(state.spins), ptr(state.fields), ptr(e2), stream())
            refined = -(e2.cpu().numpy() // 2)
        spins_all = {}
        for b in range(B):
            res = results[b]
            kb = len(cand_alpha[b])
            res.num_candidates = kb
            res.raw_energies = [int(v) for v in raw[b, :kb]]
            res.refined_energies = [int(v) for v in refined[b, :kb]]
            best_e, best_src, best_alpha = float("inf"), None, None
            for r, c in zip(res.rows, cand_of_row[b]):
                if c is None:
                    continue
                e = float(raw[b, c])
                r['raw_energy'] = e
                r['energy'] = e
                if e < best_e:
                    best_e, best_src, best_alpha = e, ('original', c), r['alpha']
            if self.use_local_search:
                for r, c in zip(res.rows, cand_of_row[b]):
                    if c is None:
                        continue
                    if refined[b, c] > raw[b, c]:
                        src_e = (('original', c), float(raw[b, c]))
                    else:
                        src_e = (('refined', c), float(refined[b, c]))
                    r['energy'] = src_e[1]
                    r['local_search_iterations'] = passes
                    if src_e[1] < best_e:
                        best_e, best_src, best_alpha = src_e[1], src_e[0], r['alpha']
            res.best_energy, res.best_alpha = best_e, best_alpha
            if best_src is not None:
                kind, c = best_src
                if kind == 'refined':
                    if c not in spins_all:
                        spins_all[c] = LS.extract(prob, state, c).cpu().numpy()
                    res.best_spins = spins_all[c][bp.rows_of(b)].astype(np.int8)
                else:
                    words = cand_bits[cand_alpha[b][c]].cpu().numpy()[bp.words_of(b)]
                    bitsv = np.unpackbits(words.view(np.uint8), bitorder="little")[: bp.sizes[b]]
                    res.best_spins = np.where(bitsv > 0, 1, -1).astype(np.int8)
        torch.cuda.synchronize()
        return BatchOutcome(results, sweep_time, time.time() - t_post, total_steps, total_matvecs, passes, packed,
                            state)
