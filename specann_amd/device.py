"""Device-resident problem data and the plumbing between torch tensors and libspecann.

torch provides device memory and the current HIP stream; every computation on the data is a
libspecann call. ``DeviceProblem`` is the reference's set-up of ``solve`` (models/spectral_annealing.py
:270-288): upload J, absolute degrees, c = max degree, J/c and d = deg/c, all on the GPU.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import scipy.sparse as sp
import torch

from . import _lib
from ._lib import SaCsr, SaState, SpecAnnLibraryError, call

TILE_CHOICES = (512, 1024, 2048, 4096)
# LDS split of the blocked SpMV's workgroup (one per CU), by element size: the most rows whose sums it
# holds and the largest column block
BLOCK_ROWS = {4: 26240, 8: 13312}
BLOCK_COLS = {4: 14336, 8: 6976}
BLOCK_MAX_TILES = 368       # column blocks per panel the blocked SpMV can walk
CUS_PER_XCD = 32
PBLOCK_MAX_ROW_BLOCKS = 3840   # row blocks of the two-pass layout (pass one keeps two int32 each in LDS)
INT_FIELD_LIMIT = float(1 << 29)   # largest absolute degree the exact int32-field local search takes


def exact_int_fields(all_integer: bool, c: float) -> bool:
    """Whether the exact local search (int32 fields) may take couplings that are all integers and whose
    largest absolute degree is c."""
    return bool(all_integer) and c < INT_FIELD_LIMIT


FLOAT_FIELD_MIN = 2.0 ** -30     # smallest |J / c| (FP32) whose FP64 field sums stay exact


def float_fields_exact(min_coupling: float) -> bool:
    """Whether FP64 sums of FP32 couplings J / c with this smallest nonzero magnitude are exact: every
    value is then a multiple of 2^-53 (at least 2^-30 times 2^-23), and a field's partial sums stay
    within the largest absolute degree over c, i.e. 1 (+ rounding), below 2^53 such units."""
    return min_coupling >= FLOAT_FIELD_MIN


def require_gpu() -> torch.device:
    """The HIP device the solver runs on; raises when there is none (there is no CPU path)."""
    if not torch.cuda.is_available():
        raise SpecAnnLibraryError("no HIP device is visible: libspecann runs on the GPU only")
    _lib.load()
    return torch.device("cuda", torch.cuda.current_device())


def stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t: torch.Tensor | None) -> ctypes.c_void_p | None:
    return None if t is None else ctypes.c_void_p(t.data_ptr())


STAGE_BYTES = 64 << 20      # pinned staging buffer size of the chunked host-to-device upload


def upload(a: np.ndarray, device: torch.device) -> torch.Tensor:
    """Copy a host array to the device. Arrays above two staging buffers go through two pinned
    buffers in turn: the host fills one while the DMA engine drains the other, so the copy runs at
    pinned-memory speed without pinning (or registering) the caller's array."""
    a = np.ascontiguousarray(a)
    src = torch.from_numpy(a)
    if a.nbytes <= 2 * STAGE_BYTES or device.type != "cuda":
        return src.to(device)
    flat = src.view(-1)
    out = torch.empty(flat.numel(), dtype=flat.dtype, device=device)
    per = STAGE_BYTES // flat.element_size()
    stages = [torch.empty(per, dtype=flat.dtype, pin_memory=True) for _ in range(2)]
    done = [None, None]
    s = torch.cuda.current_stream(device)
    for i, lo in enumerate(range(0, flat.numel(), per)):
        hi = min(lo + per, flat.numel())
        b = i & 1
        if done[b] is not None:
            done[b].synchronize()            # the copy that last read this buffer has finished
        stages[b][: hi - lo].copy_(flat[lo:hi])
        out[lo:hi].copy_(stages[b][: hi - lo], non_blocking=True)
        done[b] = torch.cuda.Event()
        done[b].record(s)
    s.synchronize()
    return out.view(src.shape)


def suffix(dtype: torch.dtype) -> str:
    if dtype == torch.float32:
        return "f32"
    if dtype == torch.float64:
        return "f64"
    raise ValueError(f"unsupported dtype {dtype}")


class Work:
    """The work buffer of sa_work_bytes: the sa_state block plus reduction partials."""

    def __init__(self, device: torch.device):
        nbytes = int(_lib.load().sa_work_bytes(0))
        self.buf = torch.zeros(nbytes, dtype=torch.uint8, device=device)
        self._host = torch.empty(_lib.STATE_HEAD_BYTES, dtype=torch.uint8, pin_memory=True)

    @property
    def ptr(self) -> ctypes.c_void_p:
        return ctypes.c_void_p(self.buf.data_ptr())

    def head(self) -> SaState:
        """Synchronously read the scalar head of the state (status, counters, results)."""
        self._host.copy_(self.buf[: _lib.STATE_HEAD_BYTES], non_blocking=True)
        torch.cuda.current_stream().synchronize()
        st = SaState()
        ctypes.memmove(ctypes.addressof(st), self._host.data_ptr(), _lib.STATE_HEAD_BYTES)
        return st

    def full(self) -> SaState:
        raw = self.buf[: ctypes.sizeof(SaState)].cpu().numpy().tobytes()
        return SaState.from_buffer_copy(raw)


def default_tile(nnz: int, n: int) -> int:
    """Stored entries per SpMV tile: about 50 rows per 256-thread workgroup for short rows."""
    mean = nnz / max(n, 1)
    t = 2048 if mean >= 24 else 1024 if mean >= 8 else 512
    return t


class DeviceProblem:
    """Normalised couplings on the GPU.

    Attributes: n, nnz, c (the normalisation factor, max absolute degree), d (normalised degrees,
    device, dtype), vals (J/c, device, dtype), jint (integer J or None), csr (SaCsr for the C ABI).
    """

    def __init__(self, J: sp.spmatrix, dtype: torch.dtype = torch.float32, tile_nnz: int | None = None,
                 lanes_per_row: int = 0, device: torch.device | None = None, row_c: np.ndarray | None = None):
        """row_c (host, one factor per row): normalise row r and its entries by row_c[r] instead of one
        global max degree (a block-diagonal batch, each instance by its own max degree); c is then 1."""
        device = device or require_gpu()
        J = sp.csr_matrix(J)
        if not J.has_sorted_indices:
            J = J.copy()
            J.sort_indices()
        if J.shape[0] != J.shape[1]:
            raise ValueError("J must be square")
        n, nnz = J.shape[0], J.nnz
        if n < 2:
            raise ValueError("need at least 2 spins")
        if nnz >= 2**31:
            raise ValueError("more than 2^31 stored couplings: int32 CSR indices cannot address them")
        self.n, self.nnz, self.dtype, self.device = n, nnz, dtype, device
        sfx = suffix(dtype)
        s = stream()
        self.indptr = upload(np.asarray(J.indptr, dtype=np.int32), device)
        self.indices = upload(np.asarray(J.indices, dtype=np.int32), device)
        raw = upload(np.asarray(J.data, dtype=np.float64), device)
        self.vals = torch.empty(nnz, dtype=dtype, device=device)
        self.jint = torch.empty(nnz, dtype=torch.int32, device=device)
        nonint = torch.zeros(1, dtype=torch.int32, device=device)
        self.work = Work(device)
        if nnz:
            call(f"sa_prepare_values_{sfx}", nnz, ptr(raw), ptr(self.vals), ptr(self.jint), ptr(nonint), s)
        del raw
        self.tile_nnz = int(tile_nnz or default_tile(nnz, n))
        if self.tile_nnz not in TILE_CHOICES:
            raise ValueError(f"tile_nnz must be one of {TILE_CHOICES}")
        self.ntiles = max(1, math.ceil(nnz / self.tile_nnz))
        self.tile_rows = torch.empty(self.ntiles + 1, dtype=torch.int32, device=device)
        call("sa_csr_tiles", n, ptr(self.indptr), self.tile_nnz, ptr(self.tile_rows), self.ntiles, s)
        self.csr = SaCsr(n=n, nnz=nnz, indptr=self.indptr.data_ptr(), indices=self.indices.data_ptr(),
                         vals=self.vals.data_ptr(), jint=self.jint.data_ptr(), tile_rows=self.tile_rows.data_ptr(),
                         ntiles=self.ntiles, tile_nnz=self.tile_nnz, lanes_per_row=int(lanes_per_row))
        self.deg = torch.empty(n, dtype=dtype, device=device)
        call(f"sa_abs_degree_{sfx}", ctypes.byref(self.csr), ptr(self.deg), self.work.ptr, s)
        head = self.work.head()
        self.c = float(head.r[0])
        # The exact local search keeps int32 fields h_i = sum_j J_ij s_j (|h_i| <= the absolute degree,
        # at most c) and adds 2 J_ij s_j per flip: both fit int32 while c < 2^30. The margin to 2^29
        # covers the rounding of c in FP32; larger integer couplings take the floating-point path.
        self.integer = exact_int_fields(nonint.item() == 0, self.c)
        if not self.integer:
            self.jint = None
            self.csr.jint = None
        if not (self.c > 0.0 and math.isfinite(self.c)):
            raise ValueError("J has no nonzero coupling")
        self.d = torch.empty(n, dtype=dtype, device=device)
        if row_c is not None:
            rc = torch.from_numpy(np.ascontiguousarray(row_c, dtype=np.float64)).to(device=device, dtype=dtype)
            call(f"sa_normalize_rows_{sfx}", ctypes.byref(self.csr), ptr(self.vals), ptr(self.deg), ptr(self.d),
                 ptr(rc), s)
            self.c = 1.0
        else:
            call(f"sa_normalize_{sfx}", nnz, ptr(self.vals), n, ptr(self.deg), ptr(self.d), self.c, s)
        self.sfx = sfx
        self.row_counts = np.diff(J.indptr).astype(np.int64)   # entries per row = per column (symmetric J)
        # one magnitude for every normalised coupling (J in {-1, +1}, J = -1, ...): the stored |J/c|
        absj = np.abs(J.data)
        nz = absj[absj > 0]
        # smallest nonzero |J / c| in FP32: the floating-point local search needs it >= 2^-30
        self.min_coupling = float(np.float32(nz.min() / self.c)) if nz.size and row_c is None else 0.0
        self._fvals = None
        self.uniform_magnitude = None
        if nnz and row_c is None and absj.min() == absj.max():
            self.uniform_magnitude = float(torch.abs(self.vals[:1]).item())
        self.npanels = 0
        self.tuning = None

    # -- couplings of the floating-point local search ---------------------------------------------
    def float_couplings(self) -> torch.Tensor:
        """J / c in FP32 (the couplings of the floating-point local search and energies). Raises when
        their dynamic range would make the FP64 fields inexact (order-dependent)."""
        if not float_fields_exact(self.min_coupling):
            raise NotImplementedError(
                f"the floating-point local search needs every nonzero |J| >= 2^-30 of the largest absolute "
                f"degree (smallest here {self.min_coupling:.3g}); its FP64 fields are exact only then")
        if self._fvals is None:
            self._fvals = self.vals if self.dtype == torch.float32 else self.vals.to(torch.float32)
        return self._fvals

    # -- XCD panel layout ----------------------------------------------------------------------
    def panel_bounds(self, npanels: int) -> np.ndarray:
        """Column boundaries of npanels panels holding about nnz / npanels entries each."""
        cum = np.cumsum(self.row_counts)
        targets = (np.arange(1, npanels) * (self.nnz / npanels))
        inner = np.searchsorted(cum, targets, side="right") + 1
        b = np.concatenate([[0], np.clip(inner, 0, self.n), [self.n]]).astype(np.int64)
        return np.maximum.accumulate(b)

    def auto_panels(self, budget_bytes: float = 2.5 * 2**20) -> int:
        """Smallest multiple of 8 panels (at most 64) whose widest panel's slice of the gathered
        vector fits the per-XCD L2 budget; 0 if even 64 do not."""
        esz = 4 if self.dtype == torch.float32 else 8
        for P in (8, 16, 32, 64):
            if np.max(np.diff(self.panel_bounds(P))) * esz <= budget_bytes:
                return P
        return 0

    def build_panels(self, npanels: int, signed: bool | None = None) -> None:
        """Regroup the couplings by column panel (sa_panel_count / scan / sa_panel_scatter). When every
        coupling has the same magnitude (signed None: detect it), the panel entries carry their sign in
        the index and the panel values are dropped (sa_panel_sign_fold)."""
        self._clear_blocks()
        if npanels == 0:
            self.npanels = 0
            self.csr.npanels = 0
            return
        if npanels % 8 or not 8 <= npanels <= 64:
            raise ValueError("npanels must be a multiple of 8 in [8, 64]")
        n, dev, s = self.n, self.device, stream()
        bounds = self.panel_bounds(npanels)
        self.pbound = torch.from_numpy(bounds.astype(np.int32)).to(dev)
        counts = torch.zeros((npanels, n), dtype=torch.int32, device=dev)
        call("sa_panel_count", self.csr_ref, npanels, ptr(self.pbound), ptr(counts), s)
        csum = torch.cumsum(counts.to(torch.int64), dim=1)
        totals = csum[:, -1]
        starts = torch.cumsum(totals, dim=0) - totals
        pind = torch.empty((npanels, n + 1), dtype=torch.int64, device=dev)
        pind[:, 0] = starts
        pind[:, 1:] = csum + starts[:, None]
        self.pindptr = pind.to(torch.int32)
        del counts, csum, pind
        self.pindices = torch.empty(self.nnz, dtype=torch.int32, device=dev)
        self.pvals = torch.empty(self.nnz, dtype=self.dtype, device=dev)
        call(f"sa_panel_scatter_{self.sfx}", self.csr_ref, npanels, ptr(self.pbound), ptr(self.pindptr),
             ptr(self.pindices), ptr(self.pvals), s)
        psigned, pmag = 0, 0.0
        if signed is None:
            signed = self.uniform_magnitude is not None
        if signed and self.uniform_magnitude is not None and self.nnz > 0:
            mismatch = torch.zeros(1, dtype=torch.int32, device=dev)
            pmag = self.uniform_magnitude
            call(f"sa_panel_sign_fold_{self.sfx}", self.nnz, ptr(self.pindices), ptr(self.pvals), pmag,
                 ptr(mismatch), s)
            if int(mismatch.item()) != 0:
                raise RuntimeError("panel values do not share one magnitude")
            psigned = 1
            del self.pvals
            self.pvals = None
        totals = totals.cpu().numpy()
        ntiles = [max(1, math.ceil(int(t) / self.tile_nnz)) for t in totals]
        off = np.zeros(npanels + 1, dtype=np.int64)
        off[1:] = np.cumsum(np.array(ntiles) + 1)
        self.ptile_rows = torch.empty(int(off[-1]), dtype=torch.int32, device=dev)
        for p in range(npanels):
            call("sa_panel_tiles", n, ctypes.c_void_p(self.pindptr.data_ptr() + 4 * p * (n + 1)), self.tile_nnz,
                 ctypes.c_void_p(self.ptile_rows.data_ptr() + 4 * int(off[p])), ntiles[p], s)
        self.ptile_off = torch.from_numpy(off).to(dev)
        # Panel rows padded like the Lanczos basis (row starts off a common HBM channel offset).
        self.pld = _lib.padded_ld(n, self.dtype.itemsize)
        self.partial = torch.empty((npanels, self.pld), dtype=self.dtype, device=dev)
        self.npanels = npanels
        c = self.csr
        c.npanels = npanels
        c.pindptr = self.pindptr.data_ptr()
        c.pindices = self.pindices.data_ptr()
        c.pvals = self.pvals.data_ptr() if self.pvals is not None else None
        c.psigned = psigned
        c.pmag = pmag
        c.ptile_rows = self.ptile_rows.data_ptr()
        c.ptile_off = self.ptile_off.data_ptr()
        c.pmax_tiles = int(max(ntiles))
        c.partial = self.partial.data_ptr()
        c.pld = self.pld

    # -- LDS-blocked layout ----------------------------------------------------------------------
    def block_geometry(self) -> tuple[int, int, int]:
        """(brows, bcols, bnr) of the LDS-blocked layout: the fewest row blocks, in whole rounds of one
        workgroup per CU of an XCD, whose row sums fit the workgroup's LDS; the column block is the
        largest the LDS split holds."""
        es = self.dtype.itemsize
        bnr = CUS_PER_XCD
        while True:
            rb = -(-(-(-self.n // bnr)) // 64) * 64
            if rb <= BLOCK_ROWS[es]:
                break
            bnr += CUS_PER_XCD
        return rb, BLOCK_COLS[es], bnr

    def row_block_imbalance(self, brows: int) -> float:
        """Largest over mean entry count of the uniform row blocks of brows rows."""
        ip = self.indptr.cpu().numpy().astype(np.int64)
        edges = np.minimum(np.arange(0, self.n + brows, brows), self.n)
        per = np.diff(ip[np.unique(edges)])
        return float(per.max() / max(per.mean(), 1.0))

    def balanced_row_bounds(self, nblocks: int, cap: int, ip: np.ndarray | None = None) -> np.ndarray:
        """Row-block boundaries (multiples of 64, the last = n) holding about nnz / nblocks entries each
        and at most cap rows: rows with many entries (the hubs of a scale-free graph) get short blocks,
        so every workgroup of the blocked SpMV sums about as many entries. ip: indptr on the host."""
        n = self.n
        if ip is None:
            ip = self.indptr.cpu().numpy().astype(np.int64)
        gstart = np.arange(0, n, 64)
        target = max(1.0, ip[-1] / max(1, nblocks))
        bid = ip[gstart] // target
        starts = gstart[np.flatnonzero(np.r_[True, bid[1:] != bid[:-1]])]
        edges = np.r_[starts, n]
        out = [0]
        for a, b in zip(edges[:-1].tolist(), edges[1:].tolist()):
            k = -(-(b - a) // cap)
            step = -(-(-(-(b - a) // k)) // 64) * 64
            out += [a + i * step for i in range(1, k)] + [b]
        return np.asarray(out, dtype=np.int64)

    def whole_round_row_bounds(self, nblocks: int, cap: int, round_of: int = CUS_PER_XCD) -> np.ndarray:
        """balanced_row_bounds for about nblocks blocks of at most cap rows, in a block count that is a
        whole number of rounds of round_of workgroups (default one per CU of an XCD) if a nearby target
        gives one."""
        n = self.n
        ip = self.indptr.cpu().numpy().astype(np.int64)
        bounds_r = self.balanced_row_bounds(nblocks, cap, ip)
        for extra in range(1, 2 * round_of):
            if (len(bounds_r) - 1) % round_of == 0:
                break
            cand = self.balanced_row_bounds(nblocks + extra, cap, ip)
            if (len(cand) - 1) % round_of == 0:
                bounds_r = cand
                break
        if bounds_r[-1] != n or np.any(np.diff(bounds_r) <= 0) or np.any(np.diff(bounds_r) > cap) or \
                np.any(bounds_r[:-1] % 64):
            raise RuntimeError("balanced row blocks out of shape")
        return bounds_r

    def capped_panel_bounds(self, wmax: int) -> np.ndarray | None:
        """Eight column panels (boundaries multiples of 64, the last = n), none wider than wmax columns,
        whose largest entry share is as small as a greedy cover allows: a binary search on that share,
        each panel growing from the left until it holds the share or reaches wmax. The couplings are
        symmetric, so the entries of columns [0, c) number indptr[c]. None if no cover exists."""
        n = self.n
        ip = self.indptr.cpu().numpy().astype(np.int64)
        grid = np.r_[np.arange(0, n, 64), n]
        cum = ip[grid]

        def cover(share: float):
            b, out = 0, [0]
            for _ in range(8):
                i = int(np.searchsorted(grid, out[-1]))
                j = int(np.searchsorted(cum, cum[i] + share, side="right")) - 1  # most columns within share
                j = max(j, i + 1)
                end = min(int(grid[min(j, len(grid) - 1)]), out[-1] + wmax // 64 * 64, n)
                out.append(end)
                if end >= n:
                    break
            return out if out[-1] >= n else None

        lo, hi = ip[-1] / 8.0, float(ip[-1])
        best = cover(hi)
        if best is None:
            return None
        for _ in range(40):
            mid = 0.5 * (lo + hi)
            c = cover(mid)
            if c is None:
                lo = mid
            else:
                hi, best = mid, c
        while len(best) < 9:  # fewer than eight panels sufficed: the rest are empty
            best.append(n)
        return np.asarray(best, dtype=np.int64)

    def build_blocks(self, signed: bool | None = None, brows: int | None = None, bcols: int | None = None,
                     balanced: bool = False) -> bool:
        """Regroup the couplings into the LDS-blocked layout (sa_block_plan / sa_block_fill) on the eight
        entry-balanced panels; False (layout unchanged) if a panel needs more column blocks than the
        kernel walks.
        brows / bcols override the geometry of block_geometry (multiples of 64 that fit LDS).
        balanced: row blocks of variable height (at most brows rows, by default the most the LDS
        holds) balanced by entries, for degree distributions whose uniform row blocks are not."""
        n, dev, s = self.n, self.device, stream()
        rb, cb, bnr = self.block_geometry()
        es = self.dtype.itemsize
        if brows is not None or bcols is not None:
            rb = int(brows or rb)
            cb = int(bcols or cb)
            bnr = -(-n // rb)
        rbound = None
        if balanced:
            cap = int(brows or BLOCK_ROWS[es])
            bounds_r = self.whole_round_row_bounds(bnr, cap)
            rb, bnr = cap, len(bounds_r) - 1
            rbound = torch.from_numpy(bounds_r.astype(np.int32)).to(dev)
        if not (64 <= rb <= BLOCK_ROWS[es] and rb % 64 == 0 and cb % 64 == 0 and 64 <= cb <= BLOCK_COLS[es]):
            raise ValueError(f"blocked layout geometry {rb} x {cb} does not fit the workgroup's LDS split")
        bounds = self.panel_bounds(8)
        bounds[1:-1] = (bounds[1:-1] // 64) * 64
        bounds = np.maximum.accumulate(bounds)
        wmax = BLOCK_MAX_TILES * cb - 64
        if np.max(np.diff(bounds)) > wmax and 8 * wmax >= n:
            # entry-balanced panels too wide for the kernel's walk (a scale-free graph's sparse columns)
            capped = self.capped_panel_bounds(wmax)
            if capped is None:
                return False
            bounds = capped
        bnc = max(1, int(-(-int(np.max(np.diff(bounds))) // cb)))
        if bnc > BLOCK_MAX_TILES:
            return False
        ntiles = 8 * bnr * bnc
        pbound = torch.from_numpy(bounds.astype(np.int32)).to(dev)
        runs = torch.zeros(128 * ntiles, dtype=torch.int32, device=dev)
        bchunk = torch.empty(ntiles + 1, dtype=torch.int32, device=dev)
        flag = torch.zeros(1, dtype=torch.int32, device=dev)
        old_rbound = self.csr.rbound
        self.csr.rbound = rbound.data_ptr() if rbound is not None else None  # the plan and fill read it
        call("sa_block_plan", self.csr_ref, ptr(pbound), rb, cb, bnr, bnc, ptr(runs), ptr(bchunk), ptr(flag), s)
        if int(flag.item()) != 0:
            self.csr.rbound = old_rbound
            return False
        nent = 64 * int(bchunk[-1].item())
        if signed is None:
            signed = self.uniform_magnitude is not None
        signed = bool(signed and self.uniform_magnitude is not None)
        mag = self.uniform_magnitude if signed else 0.0
        bent = torch.empty(max(nent, 64), dtype=torch.int32, device=dev)
        bvals = None if signed else torch.empty(max(nent, 64), dtype=self.dtype, device=dev)
        call(f"sa_block_fill_{self.sfx}", self.csr_ref, ptr(pbound), rb, cb, bnr, bnc, ptr(runs), ptr(bchunk), nent,
             ptr(bent), ptr(bvals), mag, ptr(flag), s)
        if int(flag.item()) != 0:
            self.csr.rbound = old_rbound  # the previous layout, if any, stays in force
            raise RuntimeError(f"LDS-blocked layout fill failed (flag {int(flag.item())})")
        del runs
        self.build_panels(0)
        self.pbound, self.bchunk, self.bent, self.bvals, self.rbound = pbound, bchunk, bent, bvals, rbound
        self.pld = _lib.padded_ld(n, self.dtype.itemsize)
        self.partial = torch.empty((8, self.pld), dtype=self.dtype, device=dev)
        self.npanels = 8
        self.block = {"brows": rb, "bcols": cb, "bnr": bnr, "bnc": bnc, "entries": nent, "balanced": rbound is not None}
        c = self.csr
        c.npanels, c.partial, c.pld = 8, self.partial.data_ptr(), self.pld
        c.psigned, c.pmag = (1, mag) if signed else (0, 0.0)
        c.bmode, c.brows, c.bcols, c.bnr, c.bnc = 1, rb, cb, bnr, bnc
        c.pbound, c.bchunk, c.bent = pbound.data_ptr(), bchunk.data_ptr(), bent.data_ptr()
        c.bvals = bvals.data_ptr() if bvals is not None else None
        c.rbound = rbound.data_ptr() if rbound is not None else None  # cleared by build_panels(0) above
        return True

    # -- column-streamed layout -------------------------------------------------------------------
    def cstream_geometry(self) -> tuple[int, int]:
        """(brows, bnr) of the column-streamed layout: the fewest row blocks, in whole rounds of one
        workgroup per CU of an XCD, whose FP64 row sums fit the workgroup's LDS."""
        cap = int(_lib.load().sa_cstream_max_rows())
        bnr = CUS_PER_XCD
        while True:
            rb = -(-(-(-self.n // bnr)) // 64) * 64
            if rb <= cap:
                return rb, bnr
            bnr += CUS_PER_XCD

    def build_cstream(self, brows: int | None = None, balanced: bool = False, signed: bool | None = None) -> bool:
        """Regroup the couplings into the column-streamed layout (FP32 only): per row block and column
        panel, the entries sorted by column in 256-entry chunks (sa_cstream_plan, a radix sort on the
        GPU, then sa_cstream_fill_f32). False (layout unchanged) for FP64 problems or when a chunk would
        span more columns than its 16-bit offsets reach.
        brows: rows per row block (at most sa_cstream_max_rows); balanced: row blocks of variable
        height holding about equal entry counts (hubs); signed: sign-packed entries (None: whenever
        every coupling has one magnitude; False: values stored beside the entries)."""
        if self.dtype != torch.float32:
            return False
        n, dev = self.n, self.device
        cap = int(_lib.load().sa_cstream_max_rows())
        rb, bnr = self.cstream_geometry()
        if brows is not None:
            rb = int(brows)
            if not (64 <= rb <= cap):
                raise ValueError(f"column-streamed row blocks must hold 64..{cap} rows")
            bnr = -(-n // rb)
        rbound = None
        if balanced:
            bounds_r = self.whole_round_row_bounds(bnr, rb)
            bnr = len(bounds_r) - 1
            rbound = torch.from_numpy(bounds_r.astype(np.int32)).to(dev)
        bounds = self.panel_bounds(8)
        bounds[1:-1] = (bounds[1:-1] // 64) * 64
        bounds = np.maximum.accumulate(bounds)
        if signed is None:
            signed = self.uniform_magnitude is not None
        arr = self.cstream_arrays(rb, bnr, rbound, bounds, bool(signed))
        if arr is None:
            return False
        self.build_panels(0)
        pbound, bchunk, bent, bvals, bcbase, mag = (arr[k] for k in ("pbound", "bchunk", "bent", "bvals", "bcbase",
                                                                    "mag"))
        self.pbound, self.bchunk, self.bent, self.bvals, self.rbound, self.bcbase = \
            pbound, bchunk, bent, bvals, rbound, bcbase
        self.pld = _lib.padded_ld(n, 4)
        self.partial = torch.empty((8, self.pld), dtype=self.dtype, device=dev)
        self.npanels = 8
        self.block = {"layout": "cstream", "brows": rb, "bnr": bnr, "chunks": arr["nchunks"],
                      "balanced": rbound is not None}
        c = self.csr
        c.npanels, c.partial, c.pld = 8, self.partial.data_ptr(), self.pld
        c.psigned, c.pmag = (1, mag) if mag > 0.0 else (0, 0.0)
        c.bmode, c.brows, c.bcols, c.bnr, c.bnc = 2, rb, 0, bnr, 0
        c.pbound, c.bchunk, c.bent = pbound.data_ptr(), bchunk.data_ptr(), bent.data_ptr()
        c.bvals = bvals.data_ptr() if bvals is not None else None
        c.rbound = rbound.data_ptr() if rbound is not None else None
        c.bcbase = bcbase.data_ptr()
        return True

    def rstream_geometry(self) -> tuple[int, int]:
        """(brows, bnr) of the row-owned column-streamed layout: whole rounds of one workgroup per CU
        (8 XCDs x 32), each row block as tall as the rounds allow (denser columns per gather) and at
        most what the workgroup's LDS holds in FP64 row sums."""
        cap = int(_lib.load().sa_cstream_max_rows())
        cus = 8 * CUS_PER_XCD
        rounds = max(1, -(-self.n // (cus * cap)))
        rb = -(-(-(-self.n // (cus * rounds))) // 64) * 64
        return rb, -(-self.n // rb)

    def build_rstream(self, brows: int | None = None, balanced: bool = False, signed: bool | None = None) -> bool:
        """Regroup the couplings into the row-owned column-streamed layout (FP32 only, bmode 4): the
        column-streamed arrays built with one panel, so that each workgroup sums every entry of its
        row block (sorted by column, 256-entry chunks) and writes its rows of y with no panel partials.
        False (layout unchanged) for FP64 problems or when a chunk would span more columns than its
        16-bit offsets reach. brows: rows per row block (at most sa_cstream_max_rows); balanced: row
        blocks of variable height holding about equal entry counts, in whole rounds of one workgroup
        per CU; signed: as for build_cstream."""
        if self.dtype != torch.float32:
            return False
        n, dev = self.n, self.device
        cap = int(_lib.load().sa_cstream_max_rows())
        rb, bnr = self.rstream_geometry()
        if brows is not None:
            rb = int(brows)
            if not (64 <= rb <= cap):
                raise ValueError(f"column-streamed row blocks must hold 64..{cap} rows")
            bnr = -(-n // rb)
        rbound = None
        if balanced:
            bounds_r = self.whole_round_row_bounds(bnr, rb, 8 * CUS_PER_XCD)
            bnr = len(bounds_r) - 1
            rbound = torch.from_numpy(bounds_r.astype(np.int32)).to(dev)
        bounds = np.array([0] + [n] * 8, dtype=np.int64)  # one panel: every column
        if signed is None:
            signed = self.uniform_magnitude is not None
        arr = self.cstream_arrays(rb, bnr, rbound, bounds, bool(signed))
        if arr is None:
            return False
        self.build_panels(0)
        pbound, bchunk, bent, bvals, bcbase, mag = (arr[k] for k in ("pbound", "bchunk", "bent", "bvals", "bcbase",
                                                                    "mag"))
        self.pbound, self.bchunk, self.bent, self.bvals, self.rbound, self.bcbase = \
            pbound, bchunk, bent, bvals, rbound, bcbase
        self.block = {"layout": "rstream", "brows": rb, "bnr": bnr, "chunks": arr["nchunks"],
                      "balanced": rbound is not None}
        c = self.csr
        c.psigned, c.pmag = (1, mag) if mag > 0.0 else (0, 0.0)
        c.bmode, c.brows, c.bcols, c.bnr, c.bnc = 4, rb, 0, bnr, 0
        c.pbound, c.bchunk, c.bent = pbound.data_ptr(), bchunk.data_ptr(), bent.data_ptr()
        c.bvals = bvals.data_ptr() if bvals is not None else None
        c.rbound = rbound.data_ptr() if rbound is not None else None
        c.bcbase = bcbase.data_ptr()
        return True

    def cstream_arrays(self, rb: int, bnr: int, rbound: torch.Tensor | None, bounds: np.ndarray,
                       signed: bool) -> dict | None:
        """The column-streamed layout's arrays for row blocks of at most rb rows (bnr of them, or the
        bnr + 1 boundaries rbound) and the 8 column panels bounds (9 boundaries): sa_cstream_plan (a radix
        sort on the GPU) then sa_cstream_fill_f32. None when the layout cannot hold this matrix (a chunk
        wider than its 16-bit column offsets, a row block taller than rb)."""
        n, nnz, dev, s = self.n, self.nnz, self.device, stream()
        pbound = torch.from_numpy(np.asarray(bounds).astype(np.int32)).to(dev)
        nwg = 8 * bnr
        keys = torch.empty(max(nnz, 1), dtype=torch.int64, device=dev)
        idx = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)
        keys_out, idx_out = torch.empty_like(keys), torch.empty_like(idx)
        tmp_bytes = ctypes.c_size_t()
        call("sa_cstream_sort_bytes", nnz, ctypes.byref(tmp_bytes))
        tmp = torch.empty(max(int(tmp_bytes.value), 16), dtype=torch.uint8, device=dev)
        start = torch.empty(nwg + 1, dtype=torch.int32, device=dev)
        flag = torch.zeros(1, dtype=torch.int32, device=dev)
        old_rbound = self.csr.rbound
        self.csr.rbound = rbound.data_ptr() if rbound is not None else None  # the plan reads it
        try:
            call("sa_cstream_plan", self.csr_ref, ptr(pbound), rb, bnr, ptr(keys), ptr(idx), ptr(keys_out),
                 ptr(idx_out), ptr(tmp), tmp.numel(), ptr(start), ptr(flag), s)
        finally:
            self.csr.rbound = old_rbound
        del keys, idx, tmp
        if int(flag.item()) != 0:
            return None
        counts = (start[1:] - start[:-1]).to(torch.int64)
        chunks = (counts + 255) // 256
        bchunk = torch.zeros(nwg + 1, dtype=torch.int64, device=dev)
        bchunk[1:] = torch.cumsum(chunks, 0)
        nchunks = int(bchunk[-1].item())
        if nchunks >= 2**31:
            return None
        bchunk = bchunk.to(torch.int32)
        signed = bool(signed and self.uniform_magnitude is not None)
        mag = self.uniform_magnitude if signed else 0.0
        bent = torch.empty(max(256 * nchunks, 256), dtype=torch.int32, device=dev)
        bvals = None if signed else torch.empty(max(256 * nchunks, 256), dtype=torch.float32, device=dev)
        bcbase = torch.zeros(max(nchunks, 1), dtype=torch.int32, device=dev)
        call("sa_cstream_fill_f32", self.csr_ref, ptr(pbound), rb, bnr, ptr(keys_out), ptr(idx_out), ptr(start),
             ptr(bchunk), nchunks, ptr(bent), ptr(bvals), ptr(bcbase), mag, ptr(flag), s)
        del keys_out, idx_out, start
        if int(flag.item()) != 0:
            return None
        return {"pbound": pbound, "bchunk": bchunk, "bent": bent, "bvals": bvals, "bcbase": bcbase,
                "nchunks": nchunks, "mag": mag}

    # -- two-pass layout ----------------------------------------------------------------------------
    def build_pblock(self, brows: int | None = None, balanced: bool = False, signed: bool | None = None) -> bool:
        """Regroup the couplings into the two-pass layout (FP32 only; for vectors far larger than an
        XCD's L2): entries sorted by (column block, row block, column, row) with sa_pblock_plan (a radix
        sort on the GPU) and sa_pblock_fill_f32, plus the products buffer between the passes. False
        (layout unchanged) for FP64 problems. brows, balanced, signed: as for build_cstream."""
        if self.dtype != torch.float32:
            return False
        n, nnz, dev, s = self.n, self.nnz, self.device, stream()
        lib = _lib.load()
        cap = int(lib.sa_cstream_max_rows())
        cols = int(lib.sa_pblock_cols())
        rb, bnr = self.cstream_geometry()
        if brows is not None:
            rb = int(brows)
            if not (64 <= rb <= cap):
                raise ValueError(f"two-pass row blocks must hold 64..{cap} rows")
            bnr = -(-n // rb)
        rbound = None
        if balanced:
            bounds_r = self.whole_round_row_bounds(bnr, rb)
            bnr = len(bounds_r) - 1
            rbound = torch.from_numpy(bounds_r.astype(np.int32)).to(dev)
        if bnr > PBLOCK_MAX_ROW_BLOCKS:
            return False
        # column blocks of about equal entry counts (twice as many as uniform blocks of the widest size
        # the LDS holds), so that a hub's columns do not make one pass-one workgroup the long pole
        ip = self.indptr.cpu().numpy().astype(np.int64)
        cb_np = self.balanced_row_bounds(2 * -(-n // cols), cols, ip)
        if cb_np[-1] != n or np.any(np.diff(cb_np) <= 0) or np.any(np.diff(cb_np) > cols) or np.any(cb_np[:-1] % 64):
            raise RuntimeError("two-pass column blocks out of shape")
        nc = len(cb_np) - 1
        gran = np.searchsorted(cb_np, np.arange(0, n, 64, dtype=np.int64), side="right") - 1
        cbound = torch.from_numpy(cb_np.astype(np.int32)).to(dev)
        cgran = torch.from_numpy(gran.astype(np.int32)).to(dev)
        keys = torch.empty(max(nnz, 1), dtype=torch.int64, device=dev)
        idx = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)
        keys_out, idx_out = torch.empty_like(keys), torch.empty_like(idx)
        tmp_bytes = ctypes.c_size_t()
        call("sa_cstream_sort_bytes", nnz, ctypes.byref(tmp_bytes))
        tmp = torch.empty(max(int(tmp_bytes.value), 16), dtype=torch.uint8, device=dev)
        toff = torch.empty(nc * bnr + 1, dtype=torch.int32, device=dev)
        flag = torch.zeros(1, dtype=torch.int32, device=dev)
        old_rbound = self.csr.rbound
        self.csr.rbound = rbound.data_ptr() if rbound is not None else None  # the plan reads it
        try:
            call("sa_pblock_plan", self.csr_ref, rb, bnr, ptr(cbound), nc, ptr(cgran), ptr(keys), ptr(idx),
                 ptr(keys_out), ptr(idx_out), ptr(tmp), tmp.numel(), ptr(toff), ptr(flag), s)
        finally:
            self.csr.rbound = old_rbound
        del keys, idx, tmp, cgran
        if int(flag.item()) != 0:
            return False
        if signed is None:
            signed = self.uniform_magnitude is not None
        signed = bool(signed and self.uniform_magnitude is not None)
        mag = self.uniform_magnitude if signed else 0.0
        # tile (c, r)'s place in the products array, ordered by row block then column block
        sizes = (toff[1:] - toff[:-1]).to(torch.int64).view(nc, bnr).t().reshape(-1)
        rtoff = torch.zeros(nc * bnr + 1, dtype=torch.int64, device=dev)
        rtoff[1:] = torch.cumsum(sizes, 0)
        rtoff = rtoff.to(torch.int32)
        pbent = torch.zeros(nnz + 16, dtype=torch.int16, device=dev)
        pbrow = torch.empty(max(nnz, 1), dtype=torch.int16, device=dev)
        bvals = None if signed else torch.zeros(nnz + 16, dtype=torch.float32, device=dev)
        call("sa_pblock_fill_f32", self.csr_ref, rb, bnr, nc, ptr(keys_out), ptr(idx_out), ptr(toff), ptr(rtoff),
             ptr(pbent), ptr(pbrow), ptr(bvals), mag, ptr(flag), s)
        del keys_out, idx_out
        if int(flag.item()) != 0:
            return False
        contrib = torch.empty(nnz + 16, dtype=torch.float32, device=dev)
        self.build_panels(0)
        self.pbent, self.pbrow, self.pbtoff, self.pbrtoff, self.pbcontrib = pbent, pbrow, toff, rtoff, contrib
        self.bvals, self.rbound, self.pbound = bvals, rbound, cbound
        self.block = {"layout": "pblock", "brows": rb, "bnr": bnr, "column_blocks": nc,
                      "balanced": rbound is not None}
        c = self.csr
        c.psigned, c.pmag = (1, mag) if signed else (0, 0.0)
        c.bmode, c.brows, c.bcols, c.bnr, c.bnc = 3, rb, 0, bnr, 0
        c.bvals = bvals.data_ptr() if bvals is not None else None
        c.rbound = rbound.data_ptr() if rbound is not None else None
        c.pbent, c.pbrow, c.pbtoff, c.pbcontrib, c.pbnc = (pbent.data_ptr(), pbrow.data_ptr(), toff.data_ptr(),
                                                           contrib.data_ptr(), nc)
        c.pbound, c.pbrtoff = cbound.data_ptr(), rtoff.data_ptr()
        return True

    def _clear_blocks(self) -> None:
        c = self.csr
        c.bmode, c.brows, c.bcols, c.bnr, c.bnc = 0, 0, 0, 0, 0
        c.pbound = c.bchunk = c.bent = c.bvals = c.rbound = c.bcbase = None
        c.pbent = c.pbrow = c.pbtoff = c.pbrtoff = c.pbcontrib = None
        c.pbnc = 0
        for name in ("bchunk", "bent", "bvals", "block", "rbound", "bcbase", "pbent", "pbrow", "pbtoff", "pbrtoff",
                     "pbcontrib"):
            if hasattr(self, name):
                delattr(self, name)

    def time_spmv(self, reps: int = 5) -> float:
        """Average seconds of one fused SpMV application (the Lanczos step's launch, plus the panel
        combine when panels are on), measured with HIP events on the current stream."""
        s = torch.empty(self.n, dtype=self.dtype, device=self.device)
        self.alpha_scaling(0.5, 1e-12, s)
        g = self.fill_normal(11, 0, torch.empty(self.n, dtype=self.dtype, device=self.device))
        y = torch.empty_like(g)
        fn = f"sa_spmv_gather_{self.sfx}"
        call(fn, self.csr_ref, ptr(s), ptr(g), 0.0, None, 1.0, ptr(y), stream())
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            call(fn, self.csr_ref, ptr(s), ptr(g), 0.0, None, 1.0, ptr(y), stream())
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3 / reps

    def autotune(self, reps: int = 4, verbose: bool = False, blocks: bool = True) -> dict:
        """Pick the SpMV layout (plain tiles, XCD panels, LDS blocks, column-streamed chunks per panel
        or per row block, or the two-pass layout) and tile size by timing them on this matrix; small
        problems keep the plain layout."""
        configs = [(0, t) for t in (1024, 2048, 4096)]
        P = self.auto_panels() if self.n >= (1 << 17) else 0
        if P:
            configs += [(P, t) for t in (1024, 2048)]
        if self.n >= (1 << 20):
            # large vectors: more (narrower) panels trade partial-sum traffic for L2-resident gathers,
            # and skewed degree distributions make the entry-balanced panels' widths uneven, so the
            # rule above is only a starting point: time the other panel counts too
            configs += [(p, 1024) for p in (8, 16, 32, 64) if p != P]
        if self.n >= (1 << 16) and blocks:
            # the LDS-blocked and (FP32) column-streamed layouts (their tile size does not matter; the
            # plain tiles stay for other uses), with row blocks balanced by entries too where the
            # uniform ones are not (hubs)
            configs.append((-1, self.tile_nnz))
            if self.row_block_imbalance(self.block_geometry()[0]) > 1.25:
                configs.append((-2, self.tile_nnz))
            if self.dtype == torch.float32:
                unbalanced = self.row_block_imbalance(self.cstream_geometry()[0]) > 1.25
                configs.append((-3, self.tile_nnz))
                if unbalanced:
                    configs.append((-4, self.tile_nnz))
                # the row-owned form: no panel partials, one workgroup per CU per round
                configs.append((-7, self.tile_nnz))
                if self.row_block_imbalance(self.rstream_geometry()[0]) > 1.25:
                    configs.append((-8, self.tile_nnz))
                if self.n >= (1 << 21):
                    # the two-pass layout, for vectors far larger than an XCD's L2
                    configs.append((-5, self.tile_nnz))
                    if unbalanced:
                        configs.append((-6, self.tile_nnz))
        best, results = None, []
        for panels, tile in configs:
            self.set_tiling(tile, self.csr.lanes_per_row)
            if panels in (-1, -2):
                if not self.build_blocks(balanced=panels == -2):
                    continue
            elif panels in (-3, -4):
                if not self.build_cstream(balanced=panels == -4):
                    continue
            elif panels in (-5, -6):
                if not self.build_pblock(balanced=panels == -6):
                    continue
            elif panels in (-7, -8):
                if not self.build_rstream(balanced=panels == -8):
                    continue
            else:
                self.build_panels(panels) if panels else self.build_panels(0)
            sec = self.time_spmv(reps)
            layout = {-1: "blocked", -2: "blocked", -3: "cstream", -4: "cstream", -5: "pblock", -6: "pblock",
                      -7: "rstream", -8: "rstream"}.get(panels, "panels" if panels else "csr")
            results.append({"layout": layout, "npanels": 8 if panels in (-1, -2, -3, -4) else max(panels, 0),
                            "tile_nnz": tile, "seconds": sec})
            if panels < 0:
                results[-1].update({k: v for k, v in self.block.items()
                                    if k not in ("entries", "layout", "chunks", "column_blocks")})
            if verbose:
                print(f"[autotune] {layout} npanels={results[-1]['npanels']} tile_nnz={tile}: {sec * 1e6:.1f} us",
                      flush=True)
            if best is None or sec < best["seconds"]:
                best = results[-1]
        self.set_tiling(best["tile_nnz"], self.csr.lanes_per_row)
        if best["layout"] == "blocked":
            if self.csr.bmode != 1 or self.block.get("balanced") != best.get("balanced"):
                self.build_blocks(balanced=bool(best.get("balanced")))
        elif best["layout"] == "cstream":
            if self.csr.bmode != 2 or self.block.get("balanced") != best.get("balanced"):
                if not self.build_cstream(balanced=bool(best.get("balanced"))):
                    raise RuntimeError("column-streamed layout could not be rebuilt")
        elif best["layout"] == "pblock":
            if self.csr.bmode != 3 or self.block.get("balanced") != best.get("balanced"):
                if not self.build_pblock(balanced=bool(best.get("balanced"))):
                    raise RuntimeError("two-pass layout could not be rebuilt")
        elif best["layout"] == "rstream":
            if self.csr.bmode != 4 or self.block.get("balanced") != best.get("balanced"):
                if not self.build_rstream(balanced=bool(best.get("balanced"))):
                    raise RuntimeError("row-owned column-streamed layout could not be rebuilt")
        elif best["npanels"] != self.npanels or best["npanels"] or self.csr.bmode:
            self.build_panels(best["npanels"])
        self._free_unused_panels()
        self.tuning = {"choice": best, "tried": results}
        return self.tuning

    def _free_unused_panels(self) -> None:
        if self.npanels == 0:
            for name in ("pbound", "pindptr", "pindices", "pvals", "ptile_rows", "ptile_off", "partial"):
                if hasattr(self, name):
                    delattr(self, name)
            torch.cuda.empty_cache()

    def set_tiling(self, tile_nnz: int, lanes_per_row: int = 0) -> None:
        """Re-cut the SpMV tiles (used by the autotuner)."""
        if tile_nnz not in TILE_CHOICES:
            raise ValueError(f"tile_nnz must be one of {TILE_CHOICES}")
        self.tile_nnz = tile_nnz
        self.ntiles = max(1, math.ceil(self.nnz / tile_nnz))
        self.tile_rows = torch.empty(self.ntiles + 1, dtype=torch.int32, device=self.device)
        call("sa_csr_tiles", self.n, ptr(self.indptr), tile_nnz, ptr(self.tile_rows), self.ntiles, stream())
        self.csr.tile_rows = self.tile_rows.data_ptr()
        self.csr.ntiles = self.ntiles
        self.csr.tile_nnz = tile_nnz
        self.csr.lanes_per_row = int(lanes_per_row)

    @property
    def csr_ref(self):
        return ctypes.byref(self.csr)

    def alpha_scaling(self, alpha: float, clamp: float, out: torch.Tensor) -> tuple[float, float]:
        """s = max(d^alpha, clamp)^(-1/2) into out; returns (sum d^alpha, max d^(1-alpha))."""
        call(f"sa_alpha_scaling_{self.sfx}", self.n, ptr(self.d), float(alpha), float(clamp), ptr(out),
             self.work.ptr, stream())
        head = self.work.head()
        return float(head.r[1]), float(head.r[2])

    def matvec(self, s: torch.Tensor | None, x: torch.Tensor, a: float, b: float, out: torch.Tensor,
               scratch: torch.Tensor) -> torch.Tensor:
        """out = a x + b s (.) (J/c) (s (.) x)."""
        call(f"sa_spmv_scaled_{self.sfx}", self.csr_ref, ptr(s), ptr(x), float(a), float(b), ptr(scratch), ptr(out),
             stream())
        return out

    def norm_dot(self, x: torch.Tensor, y: torch.Tensor | None = None, a: float = 1.0, b: float = 0.0
                 ) -> tuple[float, float]:
        """(||a x + b y||, x . y) in FP64."""
        call(f"sa_norm_dot_{self.sfx}", self.work.ptr, self.n, ptr(x), ptr(y), float(a), float(b), 6, stream())
        head = self.work.head()
        return float(head.r[6]), float(head.r[7])

    def fill_normal(self, seed: int, stream_id: int, out: torch.Tensor) -> torch.Tensor:
        call(f"sa_fill_normal_{self.sfx}", self.n, int(seed) & (2**64 - 1), int(stream_id) & (2**64 - 1), ptr(out),
             stream())
        return out

    def combine3(self, a: float, x: torch.Tensor, b: float, y: torch.Tensor | None, c: float,
                 z: torch.Tensor | None, out: torch.Tensor) -> torch.Tensor:
        """out = a x + b y + c z."""
        call(f"sa_combine3_{self.sfx}", self.n, float(a), ptr(x), float(b), ptr(y), float(c), ptr(z), ptr(out),
             stream())
        return out

    def scale_copy(self, x: torch.Tensor, scale: float, out: torch.Tensor) -> torch.Tensor:
        call(f"sa_scale_copy_{self.sfx}", self.n, ptr(x), float(scale), ptr(out), stream())
        return out
