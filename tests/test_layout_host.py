"""Host-side planning of the LDS-blocked and column-streamed SpMV layouts (no GPU): row blocks balanced
by entries, column panels capped at the kernel's walk, whole rounds of workgroups."""
from types import SimpleNamespace

import numpy as np
import torch

from specann_amd import graphs
from specann_amd.device import BLOCK_MAX_TILES, DeviceProblem


def _host(J):
    return SimpleNamespace(n=J.shape[0], indptr=torch.from_numpy(J.indptr.astype(np.int32)))


def _check_row_bounds(b, n, cap):
    h = np.diff(b)
    assert b[0] == 0 and b[-1] == n
    assert np.all(h > 0) and np.all(h <= cap)
    assert np.all(b[:-1] % 64 == 0)


def test_balanced_row_blocks_spread_the_hubs():
    """Barabasi-Albert numbers its hubs first: uniform row blocks are several times unbalanced,
    entry-balanced ones are close to even, every row is covered once and no block exceeds the cap."""
    J = graphs.barabasi_albert(200000, 20, seed=5, kind="neg")
    h = _host(J)
    ip = J.indptr.astype(np.int64)
    uniform = DeviceProblem.row_block_imbalance(h, 3136)
    b = DeviceProblem.balanced_row_bounds(h, 64, 26240)
    _check_row_bounds(b, h.n, 26240)
    per = np.diff(ip[b])
    assert uniform > 4.0
    assert per.max() / per.mean() < 1.3


def test_balanced_row_blocks_respect_a_small_cap():
    J = graphs.barabasi_albert(20000, 4, seed=1, kind="neg")
    h = _host(J)
    b = DeviceProblem.balanced_row_bounds(h, 16, 1024)
    _check_row_bounds(b, h.n, 1024)


def test_uniform_degrees_give_even_blocks():
    """Erdos-Renyi degrees are nearly uniform: the balanced blocks come out nearly equal in height."""
    J = graphs.erdos_renyi(100000, 40 / 99999, seed=12, kind="pm1")
    h = _host(J)
    b = DeviceProblem.balanced_row_bounds(h, 8, 26240)
    _check_row_bounds(b, h.n, 26240)
    heights = np.diff(b)
    assert heights.max() - heights.min() <= 256
    assert DeviceProblem.row_block_imbalance(h, 26240) < 1.25


def test_capped_panels_cover_within_the_walk():
    """Column panels no wider than the cap, covering every column once, with the largest entry share
    no worse than clamping the entry-balanced split to the cap boundary by boundary."""
    J = graphs.barabasi_albert(150000, 10, seed=8, kind="neg")
    h = _host(J)
    n = h.n
    ip = J.indptr.astype(np.int64)
    wmax = BLOCK_MAX_TILES * 64 - 64
    b = DeviceProblem.capped_panel_bounds(h, wmax)
    assert b is not None and len(b) == 9
    w = np.diff(b)
    assert b[0] == 0 and b[-1] == n and np.all(w >= 0) and np.all(w <= wmax)
    assert np.all(b[:-1][w > 0] % 64 == 0)
    share = np.diff(ip[b]).max() / ip[-1]
    # the clamp: entry-balanced boundaries pulled into the feasible window one by one (64 columns of
    # margin, so that rounding the boundaries down keeps the last panel within the cap)
    wc = wmax - 64
    t = ip[-1] * np.arange(1, 8) / 8
    c = np.r_[0, (np.searchsorted(ip, t) // 64) * 64, n]
    for k in range(1, 8):
        c[k] = min(max(c[k], n - (8 - k) * wc), c[k - 1] + wc) // 64 * 64
    assert np.all(np.diff(c) <= wmax)
    assert share <= np.diff(ip[c]).max() / ip[-1] + 1e-12
    assert share < 0.2


def test_capped_panels_refuse_an_impossible_cover():
    J = graphs.erdos_renyi(5000, 10 / 4999, seed=3, kind="pm1")
    h = _host(J)
    assert DeviceProblem.capped_panel_bounds(h, 64 * 8) is None  # eight panels of 512 columns < 5000


def test_row_owned_geometry_fills_whole_rounds():
    """The row-owned column-streamed layout: one round of 256 workgroups (one per CU) while the rows
    fit, the tallest row blocks the rounds allow, never more rows than the LDS row sums hold."""
    from specann_amd import _lib
    cap = int(_lib.load().sa_cstream_max_rows())
    assert DeviceProblem.rstream_geometry(SimpleNamespace(n=1 << 22)) == (16384, 256)
    assert DeviceProblem.rstream_geometry(SimpleNamespace(n=1 << 20)) == (4096, 256)
    for n in (1000, 65536, 100003, 3 << 20, 1 << 23, 1 << 25):
        rb, bnr = DeviceProblem.rstream_geometry(SimpleNamespace(n=n))
        rounds = -(-n // (256 * cap))
        assert rb % 64 == 0 and 64 <= rb <= cap and rb * bnr >= n > rb * (bnr - 1)
        assert bnr <= 256 * rounds and (rounds == 1 or bnr > 256 * (rounds - 1))


def test_row_owned_balanced_blocks_come_in_whole_rounds():
    """Entry-balanced row blocks for the row-owned layout: a multiple of 256 blocks where a nearby
    count gives one, every row covered once, none taller than the cap, the hubs spread."""
    J = graphs.barabasi_albert(400000, 20, seed=5, kind="neg")
    h = _host(J)
    h.balanced_row_bounds = lambda nb, cap, ip=None: DeviceProblem.balanced_row_bounds(h, nb, cap, ip)
    b = DeviceProblem.whole_round_row_bounds(h, 256, 1600, 256)
    _check_row_bounds(b, h.n, 1600)
    assert (len(b) - 1) % 256 == 0
    per = np.diff(J.indptr.astype(np.int64)[b])
    uniform = DeviceProblem.row_block_imbalance(h, -(-h.n // (len(b) - 1)))
    assert per.max() / per.mean() < 0.5 * uniform  # 64-row granularity bounds the hub blocks
