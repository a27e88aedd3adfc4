"""Host-side planning of the LDS-blocked SpMV layout (no GPU): row blocks balanced by entries and
column panels capped at the kernel's walk."""
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
