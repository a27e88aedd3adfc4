"""Kernel-level parity of libspecann against the CPU oracle (GPU)."""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from oracle import specann_oracle as O
from specann_amd import graphs
from specann_amd.device import DeviceProblem, ptr, stream
from specann_amd import _lib
from specann_amd._lib import call

pytestmark = pytest.mark.gpu

F32 = np.finfo(np.float32).eps
F64 = np.finfo(np.float64).eps


def star(n):
    rows = np.zeros(n - 1, dtype=np.int64)
    cols = np.arange(1, n, dtype=np.int64)
    vals = np.where(np.arange(n - 1) % 3 == 0, 1.0, -1.0)
    return graphs.symmetric_csr(n, rows, cols, vals)


CASES = {
    "er_2k_pm1": lambda: graphs.erdos_renyi(2000, 40 / 1999, seed=1, kind="pm1"),
    "ba_20k_neg": lambda: graphs.barabasi_albert(20000, 3, seed=2, kind="neg"),
    "star_9k": lambda: star(9000),
    "er_300_sparse": lambda: graphs.erdos_renyi(300, 2 / 299, seed=3, kind="pm1"),
}


def _dt(torch_dtype):
    return np.float32 if torch_dtype == torch.float32 else np.float64


@pytest.mark.parametrize("tdtype", [torch.float32, torch.float64])
def test_sign_packed_panels_equal_plain_panels(gpu, tdtype):
    """Couplings of one magnitude: the sign-packed panel entries give bit-identical SpMV results."""
    J = graphs.erdos_renyi(5000, 30 / 4999, seed=11, kind="pm1")
    outs = []
    for signed in (False, True):
        prob = DeviceProblem(J, tdtype, tile_nnz=1024)
        assert prob.uniform_magnitude is not None
        prob.build_panels(8, signed=signed)
        assert prob.csr.psigned == int(signed)
        s = torch.empty(prob.n, dtype=tdtype, device=gpu)
        prob.alpha_scaling(0.6, 1e-12, s)
        x = torch.from_numpy(O.normal_vector(prob.n, 2, 0)).to(gpu, tdtype)
        y = torch.empty_like(x)
        prob.matvec(s, x, 0.25, -1.0, y, torch.empty_like(x))
        outs.append(y.cpu().numpy())
    np.testing.assert_array_equal(outs[0], outs[1])


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("tdtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("tile,panels", [(512, 0), (2048, 0), (1024, 8), (2048, 16)])
def test_scaled_spmv_matches_oracle(gpu, case, tdtype, tile, panels):
    J = CASES[case]()
    print(f"spmv case={case} dtype={tdtype} tile={tile} panels={panels} n={J.shape[0]} nnz={J.nnz}", flush=True)
    prob = DeviceProblem(J, tdtype, tile_nnz=tile)
    prob.build_panels(panels)
    dt = _dt(tdtype)
    Jn, d, c = O.normalize(J, dt)
    assert prob.c == pytest.approx(c, rel=0, abs=0)
    np.testing.assert_array_equal(prob.d.cpu().numpy(), d)
    alpha = 0.37
    s_ref, dsum, g = O.alpha_scaling(d, alpha)
    s = torch.empty(prob.n, dtype=tdtype, device=gpu)
    dsum_g, g_g = prob.alpha_scaling(alpha, 1e-12, s)
    np.testing.assert_allclose(s.cpu().numpy(), s_ref, rtol=4 * np.finfo(dt).eps)
    assert dsum_g == pytest.approx(dsum, rel=1e-5 if dt == np.float32 else 1e-12)
    assert g_g == pytest.approx(g, rel=1e-6 if dt == np.float32 else 1e-14)
    x = O.normal_vector(prob.n, 5, 0).astype(dt)
    xt = torch.from_numpy(x).to(gpu)
    y = torch.empty_like(xt)
    z = torch.empty_like(xt)
    rho = 1.001
    prob.matvec(s, xt, rho, -1.0, y, z)
    s_dev = s.cpu().numpy()
    ref = O.s_matvec(Jn, s_dev, x, rho)
    # error bound: a few ulps of the magnitudes involved (FP64 row sums of exact products)
    mag = rho * np.abs(x) + s_dev.astype(np.float64) * (abs(Jn).astype(np.float64) @ np.abs(s_dev * x))
    eps = np.finfo(dt).eps
    err = np.abs(y.cpu().numpy().astype(np.float64) - ref)
    # with panels every per-panel partial sum is rounded to T once more
    assert np.all(err <= (6 + panels) * eps * mag + 1e-300), f"max excess {np.max(err / (mag + 1e-300))}"


def _max_run(J, bounds, brows, bcols):
    """Longest run (a row's entries in one tile) of the LDS-blocked layout, on the host."""
    J = sp.csr_matrix(J)
    rows = np.repeat(np.arange(J.shape[0]), np.diff(J.indptr))
    p = np.minimum(np.searchsorted(bounds, J.indices, side="right") - 1, 7)
    t = ((p * (J.shape[0] // brows + 1) + rows // brows) * (J.shape[0] // bcols + 2)
         + (J.indices - bounds[p]) // bcols)
    key = rows.astype(np.int64) * (t.max() + 1) + t
    _, counts = np.unique(key, return_counts=True)
    return int(counts.max())


@pytest.mark.parametrize("case", sorted(CASES) + ["er_100k", "er_100003"])
@pytest.mark.parametrize("tdtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("geom", [None, (1024, 2048), (8192, None)])
@pytest.mark.parametrize("signed", [True, False])
def test_blocked_spmv_matches_oracle(gpu, case, tdtype, geom, signed):
    """The LDS-blocked layout: same operator as the oracle within the FP summation bound and bitwise
    reproducible, hub rows (runs of more than 64 entries in one tile, cut into pieces) included."""
    if case == "er_100k":
        J = graphs.erdos_renyi(100000, 40 / 99999, seed=12, kind="pm1")
    elif case == "er_100003":  # n % 4 == 3: the last column block ends inside a 16-byte load
        J = graphs.erdos_renyi(100003, 30 / 100002, seed=13, kind="pm1")
    else:
        J = CASES[case]()
    _check_blocked(gpu, J, case, tdtype, geom, signed, balanced=False)


@pytest.mark.parametrize("case", ["ba_20k_neg", "star_9k", "er_100003", "ba_200k"])
@pytest.mark.parametrize("tdtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("geom", [None, (1024, 2048)])
@pytest.mark.parametrize("signed", [True, False])
def test_blocked_balanced_spmv_matches_oracle(gpu, case, tdtype, geom, signed):
    """Row blocks of variable height balanced by entries (hubs in short blocks): the same operator
    within the same bound, bitwise reproducible, every row covered once."""
    if case == "er_100003":
        J = graphs.erdos_renyi(100003, 30 / 100002, seed=13, kind="pm1")
    elif case == "ba_200k":
        J = graphs.barabasi_albert(200000, 20, seed=5, kind="neg")
    else:
        J = CASES[case]()
    _check_blocked(gpu, J, case, tdtype, geom, signed, balanced=True)


@pytest.mark.parametrize("tdtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("balanced", [False, True])
def test_blocked_capped_panels_match_oracle(gpu, tdtype, balanced):
    """Column panels capped at the kernel's walk (368 column blocks): with 64-column blocks the
    entry-balanced panels of a scale-free graph are too wide, and the capped split must still cover
    every column once."""
    J = graphs.barabasi_albert(150000, 10, seed=8, kind="neg")
    _check_blocked(gpu, J, "ba_150k", tdtype, (1024, 64), True, balanced)


def _check_blocked(gpu, J, case, tdtype, geom, signed, balanced):
    prob = DeviceProblem(J, tdtype)
    brows, bcols = geom if geom else (None, None)
    print(f"blocked spmv case={case} dtype={tdtype} geom={brows, bcols} signed={signed} balanced={balanced} "
          f"n={prob.n}", flush=True)
    assert prob.build_blocks(signed=signed, brows=brows, bcols=bcols, balanced=balanced)
    pb = prob.pbound.cpu().numpy()
    assert pb[0] == 0 and pb[-1] == prob.n and np.all(np.diff(pb) >= 0)
    assert np.all(np.diff(pb) <= prob.csr.bnc * prob.csr.bcols) and prob.csr.bnc <= 368
    if balanced:
        rbound = prob.rbound.cpu().numpy()
        heights = np.diff(rbound)
        assert rbound[0] == 0 and rbound[-1] == prob.n and np.all(heights > 0)
        assert np.all(heights <= prob.csr.brows) and prob.csr.bnr == len(heights)
        print(f"  {len(heights)} row blocks, heights {heights.min()}..{heights.max()}", flush=True)
    rb, cb, _ = prob.block_geometry()
    bounds = prob.panel_bounds(8)
    bounds[1:-1] = (bounds[1:-1] // 64) * 64
    print(f"  longest run {_max_run(J, np.maximum.accumulate(bounds), brows or rb, bcols or cb)}", flush=True)
    assert prob.csr.bmode == 1 and prob.csr.psigned == int(signed and prob.uniform_magnitude is not None)
    dt = _dt(tdtype)
    Jn, d, c = O.normalize(J, dt)
    s = torch.empty(prob.n, dtype=tdtype, device=gpu)
    prob.alpha_scaling(0.41, 1e-12, s)
    x = O.normal_vector(prob.n, 7, 0).astype(dt)
    xt = torch.from_numpy(x).to(gpu)
    y = torch.empty_like(xt)
    z = torch.empty_like(xt)
    rho = 0.9
    prob.matvec(s, xt, rho, -1.0, y, z)
    y1 = y.cpu().numpy()
    prob.matvec(s, xt, rho, -1.0, y, z)
    np.testing.assert_array_equal(y.cpu().numpy(), y1)
    s_dev = s.cpu().numpy()
    ref = O.s_matvec(Jn, s_dev, x, rho)
    mag = rho * np.abs(x) + s_dev.astype(np.float64) * (abs(Jn).astype(np.float64) @ np.abs(s_dev * x))
    # row sums in T within a panel: at most (entries of the row) roundings, then the panel combine
    m = int(np.max(np.diff(sp.csr_matrix(J).indptr)))
    err = np.abs(y1.astype(np.float64) - ref)
    assert np.all(err <= (10 + m) * np.finfo(dt).eps * mag + 1e-300), f"max excess {np.max(err / (mag + 1e-300))}"


def _valued(J):
    """J with couplings of several magnitudes (the column-streamed layout then stores values)."""
    J = sp.csr_matrix(J, dtype=np.float64, copy=True)
    rows = np.repeat(np.arange(J.shape[0]), np.diff(J.indptr))
    J.data = J.data * (1.0 + ((rows + J.indices) % 3))  # symmetric in (row, column)
    return J


@pytest.mark.parametrize("case", sorted(CASES) + ["er_100k", "er_100003", "ba_200k", "er_100k_valued"])
@pytest.mark.parametrize("brows", [None, 1024])
@pytest.mark.parametrize("balanced", [False, True])
def test_cstream_spmv_matches_oracle(gpu, case, brows, balanced):
    """The column-streamed layout (FP32): the same operator as the oracle within the FP summation
    bound, every entry placed once, bitwise repeatable across applications (FP64 row sums of FP32
    addends), for sign-packed and valued entries, hubs and ragged sizes."""
    if case.startswith("er_100k"):
        J = graphs.erdos_renyi(100000, 40 / 99999, seed=12, kind="pm1")
    elif case == "er_100003":
        J = graphs.erdos_renyi(100003, 30 / 100002, seed=13, kind="pm1")
    elif case == "ba_200k":
        J = graphs.barabasi_albert(200000, 20, seed=5, kind="neg")
    else:
        J = CASES[case]()
    if case.endswith("valued"):
        J = _valued(J)
    prob = DeviceProblem(J, torch.float32)
    print(f"cstream spmv case={case} brows={brows} balanced={balanced} n={prob.n} nnz={prob.nnz}", flush=True)
    assert prob.build_cstream(brows=brows, balanced=balanced)
    c = prob.csr
    assert c.bmode == 2 and c.npanels == 8 and c.psigned == int(prob.uniform_magnitude is not None)
    # every stored coupling sits in exactly one chunk slot; the rest are padding
    bent = prob.bent.cpu().numpy().view(np.uint32)
    pad = (bent >> 16) & 0x7FFF == _lib.load().sa_cstream_pad_row()
    assert int((~pad).sum()) == prob.nnz
    bch = prob.bchunk.cpu().numpy()
    assert bch[0] == 0 and np.all(np.diff(bch) >= 0) and 256 * bch[-1] == bent.size
    if balanced:
        rbound = prob.rbound.cpu().numpy()
        assert rbound[0] == 0 and rbound[-1] == prob.n and np.all(np.diff(rbound) > 0)
        assert np.all(np.diff(rbound) <= c.brows) and len(rbound) == c.bnr + 1
    Jn, d, _ = O.normalize(J, np.float32)
    s = torch.empty(prob.n, dtype=torch.float32, device=gpu)
    prob.alpha_scaling(0.41, 1e-12, s)
    x = O.normal_vector(prob.n, 7, 0).astype(np.float32)
    xt = torch.from_numpy(x).to(gpu)
    y = torch.empty_like(xt)
    z = torch.empty_like(xt)
    rho = 0.9
    prob.matvec(s, xt, rho, -1.0, y, z)
    y1 = y.cpu().numpy()
    for _ in range(3):
        prob.matvec(s, xt, rho, -1.0, y, z)
        np.testing.assert_array_equal(y.cpu().numpy(), y1)
    s_dev = s.cpu().numpy()
    ref = O.s_matvec(Jn, s_dev, x, rho)
    mag = rho * np.abs(x) + s_dev.astype(np.float64) * (abs(Jn).astype(np.float64) @ np.abs(s_dev * x))
    # exact FP64 row sums of FP32 addends (each valued product rounded once), one rounding per panel
    # partial, the panel combine and the epilogue
    err = np.abs(y1.astype(np.float64) - ref)
    assert np.all(err <= 16 * F32 * mag + 1e-300), f"max excess {np.max(err / (mag + 1e-300))}"


@pytest.mark.parametrize("case", sorted(CASES) + ["er_100k", "er_100003", "ba_200k", "er_100k_valued"])
@pytest.mark.parametrize("brows", [None, 1024])
@pytest.mark.parametrize("balanced", [False, True])
def test_rstream_spmv_matches_oracle(gpu, case, brows, balanced):
    """The row-owned column-streamed layout (FP32): one workgroup per row block over every column,
    final rows written with no panel partials; the same operator as the oracle within the FP
    summation bound (one rounding of an exact FP64 row sum, then the epilogue), every entry placed
    once, bitwise repeatable, for sign-packed and valued entries, hubs and ragged sizes."""
    if case.startswith("er_100k"):
        J = graphs.erdos_renyi(100000, 40 / 99999, seed=12, kind="pm1")
    elif case == "er_100003":
        J = graphs.erdos_renyi(100003, 30 / 100002, seed=13, kind="pm1")
    elif case == "ba_200k":
        J = graphs.barabasi_albert(200000, 20, seed=5, kind="neg")
    else:
        J = CASES[case]()
    if case.endswith("valued"):
        J = _valued(J)
    prob = DeviceProblem(J, torch.float32)
    print(f"rstream spmv case={case} brows={brows} balanced={balanced} n={prob.n} nnz={prob.nnz}", flush=True)
    assert prob.build_rstream(brows=brows, balanced=balanced)
    c = prob.csr
    assert c.bmode == 4 and c.npanels == 0 and c.psigned == int(prob.uniform_magnitude is not None)
    assert c.brows * c.bnr >= prob.n or balanced
    bent = prob.bent.cpu().numpy().view(np.uint32)
    pad = (bent >> 16) & 0x7FFF == _lib.load().sa_cstream_pad_row()
    assert int((~pad).sum()) == prob.nnz
    bch = prob.bchunk.cpu().numpy().astype(np.int64)
    assert bch.size == 8 * c.bnr + 1 and bch[0] == 0 and np.all(np.diff(bch) >= 0) and 256 * bch[-1] == bent.size
    seg = bch[:-1].reshape(c.bnr, 8)
    assert np.all(seg[:, 1:] == np.r_[seg[1:, 0], bch[-1]][:, None])  # one panel: segments 1..7 empty
    if balanced:
        rbound = prob.rbound.cpu().numpy()
        assert rbound[0] == 0 and rbound[-1] == prob.n and np.all(np.diff(rbound) > 0)
        assert np.all(np.diff(rbound) <= c.brows) and len(rbound) == c.bnr + 1
    Jn, d, _ = O.normalize(J, np.float32)
    s = torch.empty(prob.n, dtype=torch.float32, device=gpu)
    prob.alpha_scaling(0.41, 1e-12, s)
    x = O.normal_vector(prob.n, 7, 0).astype(np.float32)
    xt = torch.from_numpy(x).to(gpu)
    y = torch.empty_like(xt)
    z = torch.empty_like(xt)
    rho = 0.9
    prob.matvec(s, xt, rho, -1.0, y, z)
    y1 = y.cpu().numpy()
    for _ in range(3):
        prob.matvec(s, xt, rho, -1.0, y, z)
        np.testing.assert_array_equal(y.cpu().numpy(), y1)
    s_dev = s.cpu().numpy()
    ref = O.s_matvec(Jn, s_dev, x, rho)
    mag = rho * np.abs(x) + s_dev.astype(np.float64) * (abs(Jn).astype(np.float64) @ np.abs(s_dev * x))
    err = np.abs(y1.astype(np.float64) - ref)
    assert np.all(err <= 8 * F32 * mag + 1e-300), f"max excess {np.max(err / (mag + 1e-300))}"


@pytest.mark.parametrize("case", ["er_2k_pm1", "star_9k", "er_100k", "er_100003", "ba_200k", "ba_200k_valued"])
@pytest.mark.parametrize("brows", [None, 1024])
@pytest.mark.parametrize("balanced", [False, True])
def test_pblock_spmv_matches_oracle(gpu, case, brows, balanced):
    """The two-pass layout (FP32): products per column block, then FP64 row sums per row block; the
    same operator as the oracle within the FP summation bound, every entry in its tile once, bitwise
    repeatable, several column blocks (n > 32768), hubs, ragged sizes and valued couplings."""
    if case == "er_100k":
        J = graphs.erdos_renyi(100000, 40 / 99999, seed=12, kind="pm1")
    elif case == "er_100003":
        J = graphs.erdos_renyi(100003, 30 / 100002, seed=13, kind="pm1")
    elif case.startswith("ba_200k"):
        J = graphs.barabasi_albert(200000, 20, seed=5, kind="neg")
    else:
        J = CASES[case]()
    if case.endswith("valued"):
        J = _valued(J)
    prob = DeviceProblem(J, torch.float32)
    print(f"pblock spmv case={case} brows={brows} balanced={balanced} n={prob.n} nnz={prob.nnz}", flush=True)
    assert prob.build_pblock(brows=brows, balanced=balanced)
    c = prob.csr
    cols = _lib.load().sa_pblock_cols()
    cbound = prob.pbound.cpu().numpy().astype(np.int64)
    assert c.bmode == 3 and c.npanels == 0 and cbound.size == c.pbnc + 1
    assert cbound[0] == 0 and cbound[-1] == prob.n and np.all(np.diff(cbound) > 0) and np.all(np.diff(cbound) <= cols)
    toff = prob.pbtoff.cpu().numpy()
    assert toff[0] == 0 and toff[-1] == prob.nnz and np.all(np.diff(toff) >= 0) and toff.size == c.pbnc * c.bnr + 1
    # the stored local columns (stored order) and rows (products order) reproduce the matrix, tile by tile
    rtoff = prob.pbrtoff.cpu().numpy().astype(np.int64)
    sizes = np.diff(toff).reshape(c.pbnc, c.bnr)
    assert rtoff[0] == 0 and rtoff[-1] == prob.nnz and np.array_equal(np.diff(rtoff), sizes.T.reshape(-1))
    ent = prob.pbent.cpu().numpy().view(np.uint16)[: prob.nnz].astype(np.int64)
    t = np.repeat(np.arange(toff.size - 1), np.diff(toff))
    cb, rb = t // c.bnr, t % c.bnr
    rank = np.arange(prob.nnz) - toff[t]
    rows_l = prob.pbrow.cpu().numpy().view(np.uint16).astype(np.int64)[rtoff[rb * c.pbnc + cb] + rank]
    r0 = prob.rbound.cpu().numpy()[rb] if balanced else rb * c.brows
    got = sp.csr_matrix((np.ones(prob.nnz), (r0 + rows_l, cbound[cb] + (ent & 0x7FFF))), shape=J.shape)
    want = sp.csr_matrix(J, copy=True)
    want.data[:] = 1.0
    assert (got != want).nnz == 0
    Jn, d, _ = O.normalize(J, np.float32)
    s = torch.empty(prob.n, dtype=torch.float32, device=gpu)
    prob.alpha_scaling(0.41, 1e-12, s)
    x = O.normal_vector(prob.n, 7, 0).astype(np.float32)
    xt = torch.from_numpy(x).to(gpu)
    y = torch.empty_like(xt)
    z = torch.empty_like(xt)
    rho = 0.9
    prob.matvec(s, xt, rho, -1.0, y, z)
    y1 = y.cpu().numpy()
    for _ in range(3):
        prob.matvec(s, xt, rho, -1.0, y, z)
        np.testing.assert_array_equal(y.cpu().numpy(), y1)
    s_dev = s.cpu().numpy()
    ref = O.s_matvec(Jn, s_dev, x, rho)
    mag = rho * np.abs(x) + s_dev.astype(np.float64) * (abs(Jn).astype(np.float64) @ np.abs(s_dev * x))
    err = np.abs(y1.astype(np.float64) - ref)
    assert np.all(err <= 8 * F32 * mag + 1e-300), f"max excess {np.max(err / (mag + 1e-300))}"


def test_cstream_signed_and_valued_agree(gpu):
    """Sign-packed and valued entries of one matrix give the same products within rounding."""
    J = graphs.erdos_renyi(60000, 30 / 59999, seed=21, kind="pm1")
    outs = []
    for signed in (True, False):
        prob = DeviceProblem(J, torch.float32)
        assert prob.build_cstream(signed=signed)
        assert prob.csr.psigned == int(signed)
        s = torch.empty(prob.n, dtype=torch.float32, device=gpu)
        prob.alpha_scaling(0.6, 1e-12, s)
        xt = torch.from_numpy(O.normal_vector(prob.n, 2, 0).astype(np.float32)).to(gpu)
        y = torch.empty_like(xt)
        prob.matvec(s, xt, 0.25, -1.0, y, torch.empty_like(xt))
        outs.append(y.cpu().numpy().astype(np.float64))
    np.testing.assert_allclose(outs[0], outs[1], rtol=0, atol=8 * F32 * np.abs(outs[0]).max())


def test_cstream_refuses_wide_chunks_and_fp64(gpu):
    """A chunk whose columns span more than its 16-bit offsets reach (a sparse matching cut into
    short row blocks) makes the build refuse and leave the layout as it was; FP64 problems refuse."""
    n = 1 << 21
    perm = np.random.default_rng(3).permutation(n)
    a, b = perm[: n // 2], perm[n // 2:]
    J = graphs.symmetric_csr(n, np.minimum(a, b), np.maximum(a, b), np.ones(n // 2))
    prob = DeviceProblem(J, torch.float32)
    prob.build_panels(8)
    assert not prob.build_cstream(brows=1024)
    assert prob.csr.bmode == 0 and prob.csr.npanels == 8
    assert not prob.build_rstream(brows=64)
    assert prob.csr.bmode == 0 and prob.csr.npanels == 8
    assert not DeviceProblem(CASES["er_2k_pm1"](), torch.float64).build_cstream()
    assert not DeviceProblem(CASES["er_2k_pm1"](), torch.float64).build_rstream()
    assert not DeviceProblem(CASES["er_2k_pm1"](), torch.float64).build_pblock()


@pytest.mark.parametrize("tdtype", [torch.float32, torch.float64])
def test_normal_and_rounding_bits(gpu, tdtype):
    n = 10007
    prob = DeviceProblem(graphs.erdos_renyi(n, 8 / (n - 1), seed=4, kind="neg"), tdtype)
    v = torch.empty(n, dtype=tdtype, device=gpu)
    prob.fill_normal(11, 3, v)
    ref = O.normal_vector(n, 11, 3)
    np.testing.assert_allclose(v.cpu().numpy().astype(np.float64), ref, rtol=2 * np.finfo(_dt(tdtype)).eps,
                               atol=1e-6 if tdtype == torch.float32 else 1e-14)
    nwords = (n + 63) // 64
    bits = torch.empty(nwords, dtype=torch.int64, device=gpu)
    ylast = torch.empty_like(v)
    call(f"sa_round_{prob.sfx}", prob.work.ptr, n, ptr(v), -0.5, ptr(ylast), ptr(bits), stream())
    h = prob.work.head().hash
    yl = ylast.cpu().numpy()
    np.testing.assert_array_equal(yl, (-0.5 * v.cpu().numpy().astype(np.float64)).astype(_dt(tdtype)))
    spins = O.round_spins(yl)
    ref_bits = O.spin_bits(spins)
    np.testing.assert_array_equal(bits.cpu().numpy().view(np.uint64), ref_bits)
    assert h == O.bits_signature(ref_bits)
    call("sa_bits_compare", prob.work.ptr, nwords, ptr(bits), ptr(bits), stream())
    assert prob.work.head().i64[0] == 0


@pytest.mark.parametrize("case", ["er_2k_pm1", "ba_20k_neg"])
def test_energy_bits_exact(gpu, case):
    J = CASES[case]()
    prob = DeviceProblem(J, torch.float32)
    Jint = O.integer_couplings(J)
    k = 5
    cands = np.stack([O.round_spins(O.normal_vector(prob.n, 7, i)) for i in range(k)])
    bits = torch.from_numpy(np.stack([O.spin_bits(c) for c in cands]).view(np.int64)).to(gpu)
    e2 = torch.zeros(k, dtype=torch.int64, device=gpu)
    call("sa_energy_bits_i64", prob.csr_ref, k, ptr(bits), bits.shape[1], ptr(e2), stream())
    got = -(e2.cpu().numpy() // 2)
    ref = np.array([O.energy_int64(Jint, c) for c in cands])
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("case", ["er_2k_pm1", "ba_20k_neg", "star_9k", "er_100k"])
def test_colouring_equals_sequential_greedy(gpu, case):
    """The worklist rounds give the sequential greedy colouring (er_100k: dozens of rounds, several
    batches of queued rounds, and rounds that find the worklist empty)."""
    from specann_amd.local_search import colour_graph
    J = graphs.erdos_renyi(100000, 40 / 99999, seed=12, kind="pm1") if case == "er_100k" else CASES[case]()
    prob = DeviceProblem(J, torch.float32)
    col = colour_graph(prob, seed=42)
    ref = O.greedy_colouring(J.indptr, J.indices, 42)
    np.testing.assert_array_equal(col.colours.cpu().numpy(), ref)


@pytest.mark.parametrize("case,k", [("er_2k_pm1", 3), ("ba_20k_neg", 70), ("er_300_sparse", 1)])
def test_local_search_bit_exact(gpu, case, k):
    from specann_amd.local_search import flip_refinement
    J = CASES[case]()
    print(f"ls case={case} k={k}", flush=True)
    prob = DeviceProblem(J, torch.float32)
    Jint = O.integer_couplings(J)
    cands = np.stack([O.round_spins(O.normal_vector(prob.n, 9, i)) for i in range(k)])
    bits = torch.from_numpy(np.stack([O.spin_bits(c) for c in cands]).view(np.int64)).to(gpu)
    st = flip_refinement(prob, bits, k, max_iterations=1000, seed=123)
    ref = O.local_search(Jint, cands, 123, max_iterations=1000)
    got = st.spins[:, :k].cpu().numpy().T
    np.testing.assert_array_equal(got, ref.spins)
    assert st.passes == ref.passes
    np.testing.assert_array_equal(st.e2_start, ref.e2_start)
    np.testing.assert_array_equal(st.e2_final, ref.e2_final)
    # monotone descent and exact final energies
    assert np.all(st.e2_final >= st.e2_start)
    for c in range(k):
        assert -st.e2_final[c] // 2 == O.energy_int64(Jint, ref.spins[c])


@pytest.mark.parametrize("case", ["er_2k_pm1", "ba_20k_neg"])
@pytest.mark.parametrize("tdtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("panels", [0, 8, -1, -3, -5, -7])
def test_lanczos_leading_eigenpair(gpu, case, tdtype, panels):
    """panels -1: the LDS-blocked layout (where it holds the matrix); -3, -5, -7: the column-streamed,
    two-pass and row-owned column-streamed layouts (FP32 only)."""
    from specann_amd.eigsh import LanczosSolver
    from specann_amd._lib import SA_CONVERGED
    if panels < -1 and tdtype != torch.float32:
        pytest.skip("the streamed layouts are FP32 only")
    J = CASES[case]()
    prob = DeviceProblem(J, tdtype)
    if panels == -1:
        assert prob.build_blocks(brows=1024, bcols=1024)
    elif panels == -3:
        assert prob.build_cstream(brows=1024)
    elif panels == -5:
        assert prob.build_pblock(brows=1024)
    elif panels == -7:
        assert prob.build_rstream()
    else:
        prob.build_panels(panels)
    dt = _dt(tdtype)
    Jn, d, c = O.normalize(J, dt)
    alpha = 0.5
    s = torch.empty(prob.n, dtype=tdtype, device=gpu)
    prob.alpha_scaling(alpha, 1e-12, s)
    sv = s.cpu().numpy()
    lam_ref, y_ref = O.leading_eigpair(Jn, sv)
    tol = 1e-6 if tdtype == torch.float32 else 1e-10
    solver = LanczosSolver(prob, ncv=None)
    x0 = prob.fill_normal(1, 0, torch.empty(prob.n, dtype=tdtype, device=gpu))
    solver.begin(tol, 5000, bound=3.0)
    solver.start(s, x0)
    res = solver.run(s, 5000)
    print(f"lanczos case={case} dtype={tdtype} steps={res.steps} status={res.status} res={res.res:.3e}")
    assert res.status == SA_CONVERGED
    assert abs(res.theta - lam_ref) <= 10 * tol
    nrm, _ = solver.final_vector(None)
    y = solver.Y.cpu().numpy().astype(np.float64) / nrm
    true_res = np.linalg.norm(O.n_matvec(Jn, sv, y) - res.theta * y)
    assert true_res <= (1e-5 if tdtype == torch.float32 else 1e-9)


@pytest.mark.parametrize("tdtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("panels", [8, -1])
def test_fused_combine_matches_separate_combine(gpu, tdtype, panels):
    """The projections kernel forms W from the panel partials itself when they are 16-byte aligned;
    with a partial buffer that is only element-aligned the step runs the separate combine kernel.
    Both give bitwise the same basis, projections and scales over several steps."""
    from specann_amd.eigsh import LanczosSolver
    J = CASES["er_2k_pm1"]()
    prob = DeviceProblem(J, tdtype)
    if panels < 0:
        assert prob.build_blocks(brows=1024, bcols=1024)
    else:
        prob.build_panels(panels)
    s = torch.empty(prob.n, dtype=tdtype, device=gpu)
    prob.alpha_scaling(0.6, 1e-12, s)
    outs = []
    for misaligned in (False, True):
        saved = prob.csr.partial
        buf = torch.empty(8 * prob.pld + 4, dtype=tdtype, device=gpu)
        if misaligned:
            prob.csr.partial = buf.data_ptr() + buf.element_size()
        try:
            solver = LanczosSolver(prob, ncv=12, keep=5, graphs=False)
            x0 = prob.fill_normal(7, 0, torch.empty(prob.n, dtype=tdtype, device=gpu))
            solver.begin(1e-14, 50, bound=3.0)
            solver.start(s, x0)
            for _ in range(6):  # below ncv: no restart
                solver._step(s)
            torch.cuda.synchronize()
            st = prob.work.full()
            outs.append((solver.V[:7].clone(), list(st.h[:7]), list(st.sigma[:8]), st.j))
        finally:
            prob.csr.partial = saved
    assert outs[0][3] == outs[1][3] == 6
    assert torch.equal(outs[0][0], outs[1][0])
    assert outs[0][1] == outs[1][1] and outs[0][2] == outs[1][2]


@pytest.mark.parametrize("tdtype", [torch.float32, torch.float64])
def test_selective_reorthogonalisation_matches_full(gpu, tdtype):
    """Selective reorthogonalisation (old rows subtracted only when their projection exceeds
    sqrt(eps)/2 of beta) converges to the same eigenpair as the full one, with the true residual of
    its Ritz vector within the tolerance, and takes the selective path on most steps; with and
    without the half-width shadow basis (projections on rows older than j - 1 read from it), whose
    rows hold sigma_m sqrt(n) v_m rounded to the shadow type."""
    from specann_amd.eigsh import LanczosSolver, default_orth_tau
    from specann_amd._lib import SA_CONVERGED
    J = graphs.erdos_renyi(20000, 30 / 19999, seed=21, kind="pm1")
    prob = DeviceProblem(J, tdtype)
    dt = _dt(tdtype)
    Jn, d, c = O.normalize(J, dt)
    s = torch.empty(prob.n, dtype=tdtype, device=gpu)
    prob.alpha_scaling(0.45, 1e-12, s)
    sv = s.cpu().numpy()
    lam_ref, _ = O.leading_eigpair(Jn, sv)
    tol = 1e-6 if tdtype == torch.float32 else 1e-10
    out = {}
    for tau, shadow in ((0.0, False), (None, False), (None, True)):
        solver = LanczosSolver(prob, ncv=33, orth_tau=tau, shadow=shadow)
        x0 = prob.fill_normal(4, 0, torch.empty(prob.n, dtype=tdtype, device=gpu))
        solver.begin(tol, 4000, bound=3.0)
        solver.start(s, x0)
        res = solver.run(s, 4000)
        st = prob.work.full()
        counters = list(st.counters[:2])
        if shadow:
            # every row below the current step holds its shadow (written by its step's projections or
            # by the thick restart that rotated it)
            assert st.j > st.kept
            for m in range(st.j):
                want = solver.V[m].double() * (st.sigma[m] * np.sqrt(prob.n))
                torch.testing.assert_close(solver.Vh[m, : prob.n].double(), want,
                                           rtol=1e-3 if tdtype == torch.float32 else 1e-7, atol=1e-4)
        nrm, _ = solver.final_vector(None)
        y = solver.Y.cpu().numpy().astype(np.float64) / nrm
        true_res = np.linalg.norm(O.n_matvec(Jn, sv, y) - res.theta * y)
        out[tau, shadow] = (res, counters, true_res)
        print(f"reorth tau={tau} shadow={shadow} dtype={tdtype} steps={res.steps} theta={res.theta!r} "
              f"counters={counters} true residual {true_res:.2e}")
        assert res.status == SA_CONVERGED
        assert abs(res.theta - lam_ref) <= 10 * tol
        assert true_res <= (1e-5 if tdtype == torch.float32 else 1e-9)
    full, sel = out[0.0, False], out[None, False]
    sh = out[None, True]
    assert abs(sh[0].theta - sel[0].theta) <= 10 * tol
    assert sh[1][1] > sh[1][0]
    assert full[1][0] >= full[0].steps - 2 * (full[0].cycles + 1)   # tau = 0: old rows every step
    assert sel[1][1] > sel[1][0]                                      # mostly the two-row update
    assert default_orth_tau(tdtype) == pytest.approx(0.5 * np.sqrt(np.finfo(dt).eps))


def test_autotune_keeps_results(gpu):
    """Every tiling the autotuner may pick gives the same operator (within FP32 rounding)."""
    J = graphs.erdos_renyi(200000, 20 / 199999, seed=6, kind="pm1")
    prob = DeviceProblem(J, torch.float32)
    s = torch.empty(prob.n, dtype=torch.float32, device=gpu)
    prob.alpha_scaling(0.7, 1e-12, s)
    x = prob.fill_normal(3, 1, torch.empty_like(s))
    ref = torch.empty_like(s)
    z = torch.empty_like(s)
    prob.matvec(s, x, 0.0, 1.0, ref, z)
    tuning = prob.autotune(reps=2, verbose=True)
    assert tuning["choice"] in tuning["tried"]
    assert any(t["npanels"] for t in tuning["tried"])
    y = torch.empty_like(s)
    prob.matvec(s, x, 0.0, 1.0, y, z)
    torch.testing.assert_close(y, ref, rtol=2e-6, atol=2e-7)


@pytest.mark.parametrize("ncv,keep", [(6, 1), (8, 4), (12, 5), (33, 16)])
@pytest.mark.parametrize("tdtype", [torch.float32, torch.float64])
def test_lanczos_thick_restart_converges(gpu, ncv, keep, tdtype):
    """Many restarts with every keep count reach the dense leading eigenpair."""
    from specann_amd.eigsh import LanczosSolver
    from specann_amd._lib import SA_CONVERGED
    J = CASES["er_2k_pm1"]()
    prob = DeviceProblem(J, tdtype)
    dt = _dt(tdtype)
    Jn, d, c = O.normalize(J, dt)
    s = torch.empty(prob.n, dtype=tdtype, device=gpu)
    prob.alpha_scaling(0.8, 1e-12, s)
    sv = s.cpu().numpy()
    lam_ref, y_ref = O.leading_eigpair(Jn, sv)
    tol = 1e-6 if tdtype == torch.float32 else 1e-9
    solver = LanczosSolver(prob, ncv=ncv, keep=keep)
    x0 = prob.fill_normal(2, 0, torch.empty(prob.n, dtype=tdtype, device=gpu))
    solver.begin(tol, 20000, bound=3.0)
    solver.start(s, x0)
    res = solver.run(s, 20000)
    print(f"thick ncv={ncv} keep={keep} dtype={tdtype}: steps={res.steps} restarts={res.cycles} res={res.res:.2e}")
    assert res.status == SA_CONVERGED
    assert abs(res.theta - lam_ref) <= 10 * tol
    nrm, _ = solver.final_vector(None)
    y = solver.Y.cpu().numpy().astype(np.float64) / nrm
    assert np.linalg.norm(O.n_matvec(Jn, sv, y) - res.theta * y) <= (2e-5 if tdtype == torch.float32 else 1e-8)
    assert abs(abs(y @ y_ref) - 1) < 1e-3
