"""The allocation-free SpMV entry point (sa_spmv_scaled_static_f32) against the allocating one
(sa_spmv_scaled_f32) through the C ABI: bit-identical outputs on two calls in a row, the second on a
smaller problem, for the plain, XCD-panel, LDS-blocked, column-streamed (per panel and per row block)
and two-pass layouts; sizes beyond its static scratch are refused with a message."""
import ctypes

import numpy as np
import pytest
import torch

from specann_amd import _lib, graphs
from specann_amd._lib import SaCsr, call

MAX_N = 1 << 23        # 32 MiB of FP32 scratch


def _problem(gpu, n, deg, seed, layout):
    from specann_amd.device import DeviceProblem
    J = graphs.erdos_renyi(n, deg / (n - 1), seed=seed, kind="pm1")
    prob = DeviceProblem(J, torch.float32)
    if layout == "panels":
        prob.build_panels(8)
    elif layout == "blocked":
        assert prob.build_blocks()
    elif layout == "cstream":
        assert prob.build_cstream()
    elif layout == "pblock":
        assert prob.build_pblock()
    elif layout == "rstream":
        assert prob.build_rstream()
    s = torch.empty(prob.n, dtype=torch.float32, device=gpu)
    prob.alpha_scaling(0.4, 1e-12, s)
    x = torch.from_numpy(np.random.default_rng(seed).standard_normal(prob.n).astype(np.float32)).to(gpu)
    return prob, s, x


def _reference(prob, s, x, a, b):
    from specann_amd.device import ptr, stream
    z = torch.empty_like(x)
    y = torch.empty_like(x)
    call("sa_spmv_scaled_f32", prob.csr_ref, ptr(s), ptr(x), a, b, ptr(z), ptr(y), stream())
    return y


def _static(prob, s, x, a, b):
    from specann_amd.device import ptr, stream
    y = torch.full_like(x, float("nan"))
    call("sa_spmv_scaled_static_f32", prob.csr_ref, ptr(s), ptr(x), a, b, ptr(y), stream())
    return y


@pytest.mark.gpu
@pytest.mark.parametrize("layout", ["csr", "panels", "blocked", "cstream", "rstream", "pblock"])
def test_static_scratch_spmv_equals_allocating_entry_point(gpu, layout):
    big_n = 1 << 17
    print(f"no-alloc SpMV, layout {layout}: n = {big_n} then 5000", flush=True)
    big = _problem(gpu, big_n, 30, 3, layout)
    small = _problem(gpu, 5000, 12, 4, "csr")
    a, b = 1.001, -1.0
    want_big = _reference(*big, a, b)
    want_small = _reference(*small, 0.0, 1.0)
    got_big = _static(*big, a, b)                 # first call on this device: looks up the scratch
    got_small = _static(*small, 0.0, 1.0)         # second call, smaller input, same scratch
    torch.cuda.synchronize()
    assert torch.equal(got_big, want_big)
    assert torch.equal(got_small, want_small)
    assert torch.equal(_static(*big, a, b), want_big)   # the scratch's previous contents do not matter


def test_static_scratch_refuses_large_n():
    """n beyond the 32 MiB scratch is refused before any device work (no GPU needed)."""
    lib = _lib.load()
    A = SaCsr(n=MAX_N + 1, nnz=1, indptr=1, indices=1, vals=1, tile_rows=1, ntiles=1, tile_nnz=1024)
    rc = lib.sa_spmv_scaled_static_f32(ctypes.byref(A), None, ctypes.c_void_p(16), 0.0, 1.0,
                                       ctypes.c_void_p(16), None)
    assert rc == _lib.SA_E_ARG
    assert "static scratch" in _lib.last_error()
