"""Seeded synthetic Ising instances written straight to symmetric CSR.

These produce the inputs the solver is measured and tested on: Erdos-Renyi G(n, p) by geometric
skipping over the upper-triangle pair index, and Barabasi-Albert preferential attachment by the
edge-list copying model of Batagelj & Brandes (2005), vectorised with pointer jumping. Both are
linear in the number of edges, so graphs with 10^8-10^9 stored couplings can be built with numpy.

Every graph is returned as a ``scipy.sparse.csr_matrix`` with float64 data (the type the reference's
loader hands to ``solve``, utils/data_loader.py:129,173), sorted int32 column indices, no self-loops,
no duplicate edges. Couplings are drawn per undirected edge, so the matrix is exactly symmetric.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

__all__ = ["erdos_renyi", "barabasi_albert", "couplings", "symmetric_csr", "packed_batch", "erdos_renyi_native",
           "barabasi_albert_native"]


def couplings(m: int, kind: str, seed: int) -> np.ndarray:
    """Coupling values of m undirected edges: "neg" (all -1, the MaxCut form) or "pm1" (uniform +-1)."""
    if kind == "neg":
        return np.full(m, -1.0)
    if kind == "pm1":
        rng = np.random.default_rng(seed)
        return rng.integers(0, 2, size=m).astype(np.float64) * 2.0 - 1.0
    raise ValueError(f"unknown coupling kind {kind!r}")


def symmetric_csr(n: int, rows: np.ndarray, cols: np.ndarray, vals: np.ndarray) -> sp.csr_matrix:
    """J = U + U^T for an upper-triangle edge list (rows < cols, no duplicates)."""
    key = rows.astype(np.int64) * n + cols
    if key.size > 1 and not np.all(key[1:] > key[:-1]):
        order = np.argsort(key, kind="stable")
        rows, cols, vals = rows[order], cols[order], vals[order]
    del key
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(rows, minlength=n), out=indptr[1:])
    upper = sp.csr_matrix((vals, cols.astype(np.int32), indptr), shape=(n, n))
    lower = upper.T.tocsr()
    # Every stored column of row i in lower is < i and every one in upper is > i, so the per-row
    # concatenation [lower | upper] is already sorted.
    lo_len = np.diff(lower.indptr)
    up_len = np.diff(upper.indptr)
    indptr_j = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lo_len + up_len, out=indptr_j[1:])
    nnz = int(indptr_j[-1])
    idx_dtype = np.int32 if nnz < 2**31 else np.int64
    indices = np.empty(nnz, dtype=np.int32)
    data = np.empty(nnz, dtype=np.float64)
    # positions of the lower block of each row, then of the upper block
    row_of_lo = np.repeat(np.arange(n, dtype=np.int64), lo_len)
    pos_lo = indptr_j[:-1][row_of_lo] + (np.arange(lower.nnz, dtype=np.int64) - lower.indptr[:-1][row_of_lo])
    del row_of_lo
    indices[pos_lo] = lower.indices
    data[pos_lo] = lower.data
    del pos_lo
    row_of_up = np.repeat(np.arange(n, dtype=np.int64), up_len)
    pos_up = (indptr_j[:-1][row_of_up] + lo_len[row_of_up]
              + (np.arange(upper.nnz, dtype=np.int64) - upper.indptr[:-1][row_of_up]))
    del row_of_up
    indices[pos_up] = upper.indices
    data[pos_up] = upper.data
    del pos_up
    J = sp.csr_matrix((data, indices, indptr_j.astype(idx_dtype)), shape=(n, n))
    J.has_sorted_indices = True
    return J


def _decode_pairs(idx: np.ndarray, n: int):
    """Upper-triangle linear index -> (i, j), i < j, row-major over (0,1),(0,2),...,(1,2),..."""
    # offset of row i: i*(n-1) - i*(i-1)/2
    b = 2.0 * n - 1.0
    i = np.floor((b - np.sqrt(b * b - 8.0 * idx.astype(np.float64))) / 2.0).astype(np.int64)
    i = np.clip(i, 0, n - 2)

    def off(r):
        return r * (n - 1) - (r * (r - 1)) // 2

    # integer fix-up of the floating-point estimate
    for _ in range(3):
        too_big = off(i) > idx
        i[too_big] -= 1
        too_small = off(i + 1) <= idx
        i[too_small] += 1
    j = idx - off(i) + i + 1
    return i, j


def erdos_renyi(n: int, p: float, seed: int = 42, kind: str = "neg", coupling_seed: int | None = None
                ) -> sp.csr_matrix:
    """G(n, p) with couplings of the given kind; edges sampled by geometric gaps (linear time)."""
    if not 0.0 < p < 1.0:
        raise ValueError("p must be in (0, 1)")
    rng = np.random.default_rng(seed)
    total = n * (n - 1) // 2
    expect = p * total
    chunks, pos = [], -1
    while True:
        want = int(max(1024, (total - pos) * p * 1.05 + 6.0 * np.sqrt(expect) + 16))
        gaps = rng.geometric(p, size=want).astype(np.int64)
        idx = pos + np.cumsum(gaps)
        keep = idx < total
        chunks.append(idx[keep])
        if not keep.all():
            break
        pos = int(idx[-1])
    idx = np.concatenate(chunks)
    rows, cols = _decode_pairs(idx, n)
    del idx
    vals = couplings(rows.size, kind, seed if coupling_seed is None else coupling_seed)
    return symmetric_csr(n, rows, cols, vals)


def barabasi_albert(n: int, m: int, seed: int = 42, kind: str = "neg", coupling_seed: int | None = None
                    ) -> sp.csr_matrix:
    """Preferential attachment, m edges per new node (Batagelj-Brandes copy model).

    Slot 2k holds the new node k // m, slot 2k+1 copies a uniformly chosen earlier slot r_k in
    [0, 2k]; a copied odd slot is resolved by following its own r until an even slot is reached.
    Self-loops and repeated edges are dropped, so the mean degree is slightly below 2m.
    """
    if m < 1 or n <= m:
        raise ValueError("need 1 <= m < n")
    rng = np.random.default_rng(seed)
    E = n * m
    k = np.arange(E, dtype=np.int64)
    r = (rng.random(E) * (2 * k + 1)).astype(np.int64)
    np.minimum(r, 2 * k, out=r)
    cur = r.copy()
    odd = np.flatnonzero(cur & 1)
    while odd.size:
        cur[odd] = r[(cur[odd] - 1) >> 1]
        odd = odd[(cur[odd] & 1) == 1]
    src = k // m
    dst = (cur >> 1) // m
    del cur, r, k
    keep = src != dst
    a, b = np.minimum(src[keep], dst[keep]), np.maximum(src[keep], dst[keep])
    key = np.unique(a * n + b)
    rows, cols = key // n, key % n
    del key, a, b
    vals = couplings(rows.size, kind, seed if coupling_seed is None else coupling_seed)
    J = symmetric_csr(n, rows, cols, vals)
    deg = np.diff(J.indptr)
    if (deg == 0).any():
        keep_nodes = np.flatnonzero(deg)
        J = J[keep_nodes][:, keep_nodes].tocsr()
        J.sort_indices()
    return J


def packed_batch(graphs):
    """Block-diagonal CSR of several instances plus the row offset of each block."""
    offsets = np.zeros(len(graphs) + 1, dtype=np.int64)
    offsets[1:] = np.cumsum([g.shape[0] for g in graphs])
    J = sp.block_diag(graphs, format="csr")
    J.sort_indices()
    return J, offsets


# ---------------------------------------------------------------------------------------------
# Native generators (libspecann host code, multi-threaded): the large workloads
# ---------------------------------------------------------------------------------------------
_KINDS = {"neg": 0, "pm1": 1}


def _native_graph(build, *args, threads: int | None = None) -> sp.csr_matrix:
    import ctypes
    import os
    from . import _lib
    lib = _lib.load()
    if threads is None:
        threads = len(os.sched_getaffinity(0))
    handle = ctypes.c_void_p()
    rc = build(*args, int(threads), ctypes.byref(handle))
    if rc != 0:
        raise RuntimeError(f"graph generator failed ({rc}): {_lib.last_error()}")
    try:
        n = int(lib.sa_graph_n(handle))
        nnz = int(lib.sa_graph_nnz(handle))
        indptr = np.empty(n + 1, dtype=np.int64)
        indices = np.empty(nnz, dtype=np.int32)
        data = np.empty(nnz, dtype=np.float64)
        lib.sa_graph_export(handle, indptr.ctypes.data, indices.ctypes.data, data.ctypes.data)
    finally:
        lib.sa_graph_free(handle)
    J = sp.csr_matrix((data, indices, indptr.astype(np.int32) if nnz < 2**31 else indptr), shape=(n, n),
                      copy=False)
    J.has_sorted_indices = True
    return J


def erdos_renyi_native(n: int, p: float, seed: int = 0, kind: str = "neg", threads: int | None = None):
    """G(n, p) from libspecann's parallel generator (sa_graph_erdos_renyi): counter-based draws, so the
    graph depends on (n, p, seed, kind) only. Isolated nodes are dropped (n may shrink for small p)."""
    from . import _lib
    return _native_graph(_lib.load().sa_graph_erdos_renyi, int(n), float(p), int(seed) & (2**64 - 1),
                         _KINDS[kind], threads=threads)


def barabasi_albert_native(n: int, m: int, seed: int = 0, kind: str = "neg", threads: int | None = None):
    """Barabasi-Albert preferential attachment (m edges per node) from libspecann's parallel generator
    (sa_graph_barabasi_albert); self-loops dropped, repeated edges merged, isolated nodes dropped."""
    from . import _lib
    return _native_graph(_lib.load().sa_graph_barabasi_albert, int(n), int(m), int(seed) & (2**64 - 1),
                         _KINDS[kind], threads=threads)
