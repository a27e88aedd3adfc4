"""The reference's predictor callables (utils/predictors.py:6-56) on device tensors, computed by
libspecann (sa_diag_scale, sa_norm_dot, sa_combine3, sa_scale_copy, sa_fill_normal).

The sweep itself does not call these: it fuses both predictors into the Lanczos start
(sa_lanczos_start), so that the predicted vector is built and its residual gate taken in the same
pass. These standalone forms keep the reference's inner boundary for callers that use it directly.
"""
from __future__ import annotations

import numpy as np
import torch

from ._lib import call
from .device import Work, ptr, require_gpu, stream, suffix

_works: dict = {}


def _as_device(x, dtype: torch.dtype) -> torch.Tensor:
    dev = require_gpu()
    t = x if isinstance(x, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(x))
    return t.to(device=dev, dtype=dtype).contiguous()


def _torch_dtype(dtype) -> torch.dtype:
    if isinstance(dtype, torch.dtype):
        return dtype
    return torch.float64 if np.dtype(dtype) == np.float64 else torch.float32


def _norm_dot(x: torch.Tensor, y: torch.Tensor | None = None, a: float = 1.0, b: float = 0.0) -> tuple[float, float]:
    """(||a x + b y||, x . y) in FP64 on the device."""
    work = _works.get(x.device)
    if work is None:
        work = _works[x.device] = Work(x.device)
    call(f"sa_norm_dot_{suffix(x.dtype)}", work.ptr, x.numel(), ptr(x), ptr(y), float(a), float(b), 6, stream())
    head = work.head()
    return float(head.r[6]), float(head.r[7])


def _scaled(x: torch.Tensor, scale: float) -> torch.Tensor:
    out = torch.empty_like(x)
    call(f"sa_scale_copy_{suffix(x.dtype)}", x.numel(), ptr(x), float(scale), ptr(out), stream())
    return out


def diagonal_predictor(y_last, d_vec, delta_alpha: float, dtype, seed: int = 0, stream_id: int = 0) -> torch.Tensor:
    """D^(delta_alpha/2) y_last normalised, D = diag(max(d, 1e-12)) (utils/predictors.py:6-26); a vector
    whose norm is below 1e-12 is replaced by a normalised normal vector of the stream (seed, stream_id)."""
    dt = _torch_dtype(dtype)
    y, d = _as_device(y_last, dt), _as_device(d_vec, dt)
    out = torch.empty_like(y)
    call(f"sa_diag_scale_{suffix(dt)}", y.numel(), ptr(y), ptr(d), 0.5 * float(delta_alpha), ptr(out), stream())
    nrm, _ = _norm_dot(out)
    if not nrm >= 1e-12:
        call(f"sa_fill_normal_{suffix(dt)}", y.numel(), int(seed) & (2**64 - 1), int(stream_id) & (2**64 - 1),
             ptr(out), stream())
        nrm, _ = _norm_dot(out)
    return _scaled(out, 1.0 / nrm)


def secant_predictor(y_last, y_prev, dtype, gamma: float = 1.0) -> torch.Tensor:
    """y_last + gamma (y_last - y_prev) normalised, y_prev sign-aligned with y_last
    (utils/predictors.py:29-56); y_last itself when y_prev is None or either norm is below 1e-12."""
    dt = _torch_dtype(dtype) if dtype is not None else (y_last.dtype if isinstance(y_last, torch.Tensor)
                                                        else _torch_dtype(np.asarray(y_last).dtype))
    yl = _as_device(y_last, dt)
    if y_prev is None:
        return yl
    yp = _as_device(y_prev, dt)
    _, dot = _norm_dot(yl, yp)
    sgn = -1.0 if dot < 0 else 1.0
    snorm, _ = _norm_dot(yl, yp, 1.0, -sgn)
    if snorm < 1e-12:
        return yl
    out = torch.empty_like(yl)
    call(f"sa_combine3_{suffix(dt)}", yl.numel(), 1.0 + float(gamma), ptr(yl), -float(gamma) * sgn, ptr(yp), 0.0,
         None, ptr(out), stream())
    nrm, _ = _norm_dot(out)
    if nrm < 1e-12:
        return yl
    return _scaled(out, 1.0 / nrm)
