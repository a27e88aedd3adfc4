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
                 lanes_per_row: int = 0, device: torch.device | None = None):
        device = device or require_gpu()
        J = sp.csr_matrix(J)
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
        self.indptr = torch.from_numpy(np.ascontiguousarray(J.indptr, dtype=np.int32)).to(device)
        self.indices = torch.from_numpy(np.ascontiguousarray(J.indices, dtype=np.int32)).to(device)
        raw = torch.from_numpy(np.ascontiguousarray(J.data, dtype=np.float64)).to(device)
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
        self.integer = bool(nonint.item() == 0)
        if not self.integer:
            self.jint = None
            self.csr.jint = None
        if not (self.c > 0.0 and math.isfinite(self.c)):
            raise ValueError("J has no nonzero coupling")
        self.d = torch.empty(n, dtype=dtype, device=device)
        call(f"sa_normalize_{sfx}", nnz, ptr(self.vals), n, ptr(self.deg), ptr(self.d), self.c, s)
        self.sfx = sfx

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

    def scale_copy(self, x: torch.Tensor, scale: float, out: torch.Tensor) -> torch.Tensor:
        call(f"sa_scale_copy_{self.sfx}", self.n, ptr(x), float(scale), ptr(out), stream())
        return out
