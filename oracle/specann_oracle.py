"""CPU restatement of the SpecAnn alpha-sweep (test infrastructure; see oracle/__init__.py).

Each function cites the reference lines it restates (paths relative to the Vgueca/SpecAnn tree).
Integer arithmetic is exact (int64); floating point follows the reference's dtype where it matters
and FP64 otherwise.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
TWO53 = 9007199254740992.0
ORDER_SALT = 0x0DE5C1A55C0FFEE1
FINAL_PASS_BASE = 1 << 40
COLOUR_SALT = 0xC0102
NORMAL_SALT = 0x5A17C0DE00000000

# ---------------------------------------------------------------------------------------------
# Counter-based hashing (SplitMix64; Steele, Lea & Flood 2014 finaliser constants)
# ---------------------------------------------------------------------------------------------


def splitmix64(x):
    """Vectorised SplitMix64 finaliser on uint64 arrays (wrapping arithmetic)."""
    z = np.asarray(x, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def mix3(a, b, c):
    a = np.asarray(a, dtype=np.uint64)
    b = np.asarray(np.uint64(int(b) & 0xFFFFFFFFFFFFFFFF) if np.ndim(b) == 0 else b, dtype=np.uint64)
    c = np.asarray(c, dtype=np.uint64)
    return splitmix64(splitmix64(splitmix64(a) ^ b) ^ c)


def u64(v: int) -> np.uint64:
    return np.uint64(int(v) & 0xFFFFFFFFFFFFFFFF)


def normal_vector(n: int, seed: int, stream_id: int) -> np.ndarray:
    """Box-Muller standard normals of the stream (seed, stream_id, i), FP64."""
    i = np.arange(n, dtype=np.uint64)
    h1 = mix3(u64(seed), u64(NORMAL_SALT ^ (int(stream_id) & 0xFFFFFFFFFFFFFFFF)), i)
    h2 = splitmix64(h1)
    u1 = ((h1 >> np.uint64(11)).astype(np.float64) + 1.0) / TWO53
    u2 = (h2 >> np.uint64(11)).astype(np.float64) / TWO53
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)


# ---------------------------------------------------------------------------------------------
# Normalisation, schedule, scaling, operators
# ---------------------------------------------------------------------------------------------


def normalize(J: sp.spmatrix, dtype=np.float32):
    """(J/c, d = deg/c, c): deg = |J| 1 in dtype, c = max deg (models/spectral_annealing.py:270-288)."""
    J = sp.csr_matrix(J)
    Jd = J.astype(dtype)
    deg = np.asarray(abs(J).sum(axis=1), dtype=np.float64).ravel().astype(dtype)
    c = deg.max()
    Jn = sp.csr_matrix(((Jd.data / c).astype(dtype), Jd.indices, Jd.indptr), shape=J.shape)
    d = (deg / c).astype(dtype)
    return Jn, d, float(c)


def alpha_values(num_alphas=128, interval=(0.0, 1.0), schedule="sqrt"):
    """t = linspace(start, end, K); alpha = sqrt(t) or t (models/spectral_annealing.py:549-555)."""
    t = np.linspace(float(interval[0]), float(interval[1]), int(num_alphas))
    return [float(x) for x in (np.sqrt(t) if schedule == "sqrt" else t)]


def alpha_scaling(d: np.ndarray, alpha: float, clamp: float = 1e-12):
    """s = 1/sqrt(max(d^alpha, clamp)) in d's dtype, sum d^alpha, g = max d^(1-alpha)
    (models/spectral_annealing.py:316, :34, :326, :579)."""
    dt = d.dtype.type
    dv = np.power(d, dt(alpha))
    s = (dt(1) / np.sqrt(np.maximum(dv, dt(clamp)))).astype(d.dtype)
    g = float(np.max(np.power(d.astype(np.float64), 1.0 - alpha)))
    return s, float(dv.astype(np.float64).sum()), g


def n_matvec(Jn: sp.spmatrix, s: np.ndarray, x: np.ndarray) -> np.ndarray:
    """N x = s (.) (J/c) (s (.) x) in FP64 (NAlphaOperator._matvec, :38-41)."""
    s64 = s.astype(np.float64)
    return s64 * (Jn.astype(np.float64) @ (s64 * x.astype(np.float64)))


def s_matvec(Jn, s, x, rho: float) -> np.ndarray:
    """S x = rho x - N x (SRhoOperator._matvec, :52-53)."""
    return rho * x.astype(np.float64) - n_matvec(Jn, s, x)


def shift_rho(lambda1_prev, g: float, eps: float = 1e-3) -> float:
    """compute_shift_rho (models/spectral_annealing.py:56-67)."""
    if lambda1_prev is None or not np.isfinite(lambda1_prev) or lambda1_prev <= 0:
        base = g
    else:
        base = max(lambda1_prev, g)
    rho = (1.0 + eps) * base
    if rho <= 0:
        rho = (1.0 + eps) * max(g, 1.0)
    return rho


def check_eigenpair(lambda1: float, y: np.ndarray, g: float, eps: float) -> bool:
    """_check_eigenpair (models/spectral_annealing.py:679-691): finite and lambda1 <= g (1 + max(eps/2, 1e-5))."""
    return bool(np.isfinite(lambda1) and np.isfinite(y).all() and lambda1 <= g * (1.0 + max(0.5 * eps, 1e-5)))


def dense_n(Jn, s) -> np.ndarray:
    s64 = s.astype(np.float64)
    return (s64[:, None] * Jn.astype(np.float64).toarray()) * s64[None, :]


def leading_eigpair(Jn, s, tol: float = 1e-13, v0: np.ndarray | None = None):
    """FP64 leading eigenpair (lambda_max, unit y) of N = diag(s) J diag(s): dense eigh for n <= 3000,
    ARPACK (which='LA') otherwise."""
    n = Jn.shape[0]
    if n <= 3000:
        w, V = np.linalg.eigh(dense_n(Jn, s))
        return float(w[-1]), V[:, -1]
    s64 = s.astype(np.float64)
    J64 = Jn.astype(np.float64)
    op = spla.LinearOperator((n, n), matvec=lambda x: s64 * (J64 @ (s64 * x)), dtype=np.float64)
    w, V = spla.eigsh(op, k=1, which="LA", tol=tol, v0=v0, ncv=min(64, n - 1), maxiter=100000)
    return float(w[0]), V[:, 0]


# ---------------------------------------------------------------------------------------------
# Predictors (utils/predictors.py) and the tangent of the eigenvector path
# ---------------------------------------------------------------------------------------------


def diagonal_predictor(y_last: np.ndarray, d: np.ndarray, delta_alpha: float):
    """D^{delta/2} y_last normalised; None if the norm is < 1e-12 (utils/predictors.py:6-26)."""
    dt = d.dtype.type
    w = np.power(np.maximum(d, dt(1e-12)), dt(delta_alpha / 2.0))
    y = (w * y_last.astype(d.dtype)).astype(np.float64)
    nrm = np.linalg.norm(y)
    if nrm < 1e-12:
        return None
    return y / nrm


def secant_predictor(y_last: np.ndarray, y_prev, gamma: float = 1.0) -> np.ndarray:
    """y_last + gamma (y_last - y_prev) after sign alignment, normalised (utils/predictors.py:29-56)."""
    if y_prev is None:
        return y_last
    yl, yp = y_last.astype(np.float64), y_prev.astype(np.float64)
    if yl @ yp < 0:
        yp = -yp
    s = yl - yp
    if np.linalg.norm(s) < 1e-12:
        return y_last
    y = yl + gamma * s
    nrm = np.linalg.norm(y)
    if nrm < 1e-12:
        return y_last
    return y / nrm


def tangent(Jn, d: np.ndarray, alpha: float, clamp: float = 1e-12):
    """dy/dalpha of the unit leading eigenvector of N(alpha) (dense, small n).

    N' = -1/2 (L N + N L), L = diag(ln max(d^alpha, clamp)^(1/alpha)) = diag(ln d) for d > 0;
    y' = -(N - lambda I)^+ N' y, the minimum-norm solution, orthogonal to y.
    """
    s, _, _ = alpha_scaling(d.astype(np.float64), alpha, clamp)
    N = dense_n(Jn, s)
    w, V = np.linalg.eigh(N)
    lam, y = w[-1], V[:, -1]
    L = np.log(np.maximum(d.astype(np.float64), 1e-300))
    Np = -0.5 * (L[:, None] * N + N * L[None, :])
    M = N - lam * np.eye(N.shape[0])
    yp = -np.linalg.pinv(M, rcond=1e-12) @ (Np @ y)
    yp -= (y @ yp) * y
    return lam, y, yp


def minres_deflated(matvec, y: np.ndarray, r: np.ndarray, shift: float, tol: float, maxit: int):
    """MINRES (Paige & Saunders 1975; the recurrence of scipy.sparse.linalg.minres without a
    preconditioner) on P (N - shift I) P x = P r, P = I - y y^T with y unit, in FP64, in the order
    the build's sa_minres kernels use: the deflation is applied to each product,
    A v = (N - shift I) v - y (y . (N - shift I) v). Returns (x, iterations, phibar, beta1); x is
    projected onto y's complement at the end. Parity for the build's eigenvector tangent and its
    adjoint (no counterpart in the reference)."""
    y = np.asarray(y, dtype=np.float64)
    r = np.asarray(r, dtype=np.float64)
    b = r - (y @ r) * y
    beta1 = float(np.linalg.norm(b))
    x = np.zeros_like(b)
    if beta1 == 0.0 or maxit <= 0:
        return x, 0, beta1, beta1
    r1 = np.zeros_like(b)
    r2 = b.copy()
    w = np.zeros_like(b)
    w1 = np.zeros_like(b)
    w2 = np.zeros_like(b)
    beta, oldb = beta1, 0.0
    dbar = epsln = 0.0
    cs, sn = -1.0, 0.0
    phibar = beta1
    itn = 0
    while itn < maxit:
        q = np.asarray(matvec(r2), dtype=np.float64) - shift * r2
        g = float(y @ q)
        av = (q - g * y) / beta
        v = r2 / beta
        cr1 = beta / oldb if itn >= 1 else 0.0
        yv = av - cr1 * r1
        alfa = float(v @ yv)
        rn = yv - (alfa / beta) * r2
        beta_new = float(np.linalg.norm(rn))
        oldeps = epsln
        delta = cs * dbar + sn * alfa
        gbar = sn * dbar - cs * alfa
        epsln = sn * beta_new
        dbar = -cs * beta_new
        gamma = max(float(np.hypot(gbar, beta_new)), np.finfo(np.float64).eps)
        cs, sn = gbar / gamma, beta_new / gamma
        phi = cs * phibar
        phibar = sn * phibar
        w1, w2 = w2, w
        w = (v - oldeps * w1 - delta * w2) / gamma
        x = x + phi * w
        r1, r2 = r2, rn
        oldb, beta = beta, beta_new
        itn += 1
        if phibar <= tol * beta1 or beta_new == 0.0:
            break
    x = x - (y @ x) * y
    return x, itn, float(phibar), beta1


def tangent_minres(Jn, d: np.ndarray, alpha: float, y: np.ndarray, lam: float, steps: int, tol: float = 0.0,
                   clamp: float = 1e-12):
    """The tangent dy/dalpha = P u + 2 lambda z with z from `steps` iterations of minres_deflated
    (u = 1/2 ln max(d, clamp) (.) y), FP64."""
    s, _, _ = alpha_scaling(d.astype(np.float64), alpha, clamp)
    Jd = Jn.astype(np.float64)
    y = np.asarray(y, dtype=np.float64)
    u = 0.5 * np.log(np.maximum(d.astype(np.float64), clamp)) * y
    z, itn, phibar, beta1 = minres_deflated(lambda v: s * (Jd @ (s * v)), y, u, lam, tol, steps)
    pu = u - (y @ u) * y
    return pu + 2.0 * lam * z, z, itn


# ---------------------------------------------------------------------------------------------
# Rounding, signatures, exact energies
# ---------------------------------------------------------------------------------------------


def round_spins(x: np.ndarray) -> np.ndarray:
    """+1 where x >= 0, else -1 (models/spectral_annealing.py:345)."""
    return np.where(x >= 0, 1, -1).astype(np.int8)


def spin_bits(s: np.ndarray) -> np.ndarray:
    """uint64 words: bit l of word w is 1 iff s[64 w + l] > 0 (node order, little-endian in the word)."""
    n = s.size
    nw = (n + 63) // 64
    b = np.zeros(nw * 64, dtype=np.uint8)
    b[:n] = (s > 0)
    packed = np.packbits(b.reshape(nw, 64), axis=1, bitorder="little")
    return packed.view(np.uint64).ravel()


def bits_signature(words: np.ndarray) -> int:
    """Order-aware signature sum_w splitmix64(word_w ^ splitmix64(w)) mod 2^64."""
    w = np.arange(words.size, dtype=np.uint64)
    with np.errstate(over="ignore"):
        v = splitmix64(words.astype(np.uint64) ^ splitmix64(w))
        return int(np.sum(v, dtype=np.uint64))


def integer_couplings(J: sp.spmatrix) -> sp.csr_matrix:
    J = sp.csr_matrix(J)
    data = J.data
    if not np.all(data == np.rint(data)):
        raise ValueError("couplings are not integers")
    return sp.csr_matrix((data.astype(np.int64), J.indices, J.indptr), shape=J.shape)


def energy_int64(Jint: sp.csr_matrix, s: np.ndarray) -> int:
    """E = -1/2 s^T J s, exact for integer J (models/spectral_annealing.py:709-720, :962-976)."""
    s64 = s.astype(np.int64)
    e2 = int(s64 @ (Jint @ s64))
    assert e2 % 2 == 0
    return -e2 // 2


def energies_float(Jn, s_cols: np.ndarray, c: float) -> np.ndarray:
    """fp32_energies: FP32 couplings J/c, FP64 accumulation, times c (:962-976)."""
    S = s_cols.astype(np.float64)
    J64 = Jn.astype(np.float32).astype(np.float64)
    return -0.5 * c * np.sum(S * (J64 @ S), axis=0)


# ---------------------------------------------------------------------------------------------
# Colouring and the deterministic local search
# ---------------------------------------------------------------------------------------------


def colour_priority(indptr: np.ndarray, seed: int):
    """Priority keys (degree, hash, index): larger first."""
    n = indptr.size - 1
    deg = np.diff(indptr).astype(np.int64)
    h = mix3(u64(seed), u64(COLOUR_SALT), np.arange(n, dtype=np.uint64))
    return deg, h


def _order_desc(deg, h):
    """Node order by descending (degree, hash, index)."""
    idx = np.arange(deg.size)
    return np.lexsort((idx, h, deg))[::-1]


def greedy_colouring(indptr: np.ndarray, indices: np.ndarray, seed: int) -> np.ndarray:
    """Sequential greedy colouring in descending (degree, hash, index) order: each node takes the
    smallest colour not used by an already coloured neighbour (cf. local_search_guards/common.py:136-159,
    which orders by degree and index only)."""
    n = indptr.size - 1
    colours = np.full(n, -1, dtype=np.int64)
    for v in _order_desc(*colour_priority(indptr, seed)):
        nb = indices[indptr[v]:indptr[v + 1]]
        used = colours[nb]
        used = np.unique(used[used >= 0])
        c = 0
        for u in used:
            if u != c:
                break
            c += 1
        colours[v] = c
    return colours.astype(np.int32)


def class_order(seed: int, q: int, num_classes: int) -> np.ndarray:
    """Visiting order of the colour classes in pass q: argsort of mix3(seed ^ salt, q, class)."""
    keys = mix3(u64(int(seed) ^ ORDER_SALT), u64(q), np.arange(num_classes, dtype=np.uint64))
    return np.argsort(keys, kind="stable")


def selected(seed: int, pass_id: int, nodes: np.ndarray, frac: float) -> np.ndarray:
    """Node i takes part in pass pass_id iff frac >= 1 or (mix3(seed, pass_id, i) >> 11) < floor(frac 2^53)."""
    if frac >= 1.0:
        return np.ones(nodes.size, dtype=bool)
    thresh = np.uint64(int(frac * TWO53))
    h = mix3(u64(seed), u64(pass_id), nodes.astype(np.uint64))
    return (h >> np.uint64(11)) < thresh


@dataclass
class LocalSearchResult:
    spins: np.ndarray          # k x n int8
    passes: int                # main passes (the reference's `local_search_iterations`)
    e2_start: np.ndarray       # sum_i s_i h_i before (int64); energy = -e2/2
    e2_final: np.ndarray
    total_passes: int = 0


def local_search(Jint: sp.csr_matrix, cand: np.ndarray, seed: int, colours: np.ndarray | None = None,
                 max_iterations: int = 1000, subset_fraction: float = 0.9, final_sweep_fraction: float = 1.0,
                 final_sweep_iters: int = 25, num_rounds: int = 25) -> LocalSearchResult:
    """Batched single-flip descent with the termination schedule of flip_refinement and
    gpu_final_sweep (models/spectral_annealing.py:1039-1112, :983-1036), made deterministic: a pass
    visits the colour classes in class_order(seed, q) and, within a class, flips every selected node
    with s_i h_i < 0 (all at once, which equals one at a time since a class is independent), then
    updates the neighbours' fields. Main pass `it` selects nodes with selected(seed, it, ., subset);
    final-sweep round r with selected(seed, FINAL_PASS_BASE + r, ., final fraction) for all its passes.
    """
    Jint = sp.csr_matrix(Jint)
    n = Jint.shape[0]
    cand = np.atleast_2d(cand)
    k = cand.shape[0]
    if colours is None:
        colours = greedy_colouring(Jint.indptr, Jint.indices, seed)
    C = int(colours.max()) + 1
    classes = [np.flatnonzero(colours == q) for q in range(C)]
    S = cand.T.astype(np.int64).copy()            # n x k
    if Jint.dtype.kind not in "iu":
        # floating-point couplings (update_spins_kernel_float32/64, :821-919): FP32 values, FP64 fields
        Jint = sp.csr_matrix((Jint.data.astype(np.float32).astype(np.float64), Jint.indices, Jint.indptr),
                             shape=Jint.shape)
        S = S.astype(np.float64)
    H = np.asarray(Jint @ S)                      # fields (int64, or float64)
    e2_start = (S * H).sum(axis=0)
    q_counter = [0]

    def run_pass(sel_id: int, frac: float) -> np.ndarray:
        flips = np.zeros(k, dtype=bool)
        order = class_order(seed, q_counter[0], C)
        q_counter[0] += 1
        for cls in order:
            nodes = classes[cls]
            if frac < 1.0:
                nodes = nodes[selected(seed, sel_id, nodes, frac)]
            if nodes.size == 0:
                continue
            s = S[nodes]
            h = H[nodes]
            fl = (s * h) < 0
            if not fl.any():
                continue
            rows = np.flatnonzero(fl.any(axis=1))
            nodes, s, fl = nodes[rows], s[rows], fl[rows]
            flips |= fl.any(axis=0)
            snew = np.where(fl, -s, s)
            S[nodes] = snew
            delta = np.where(fl, 2 * snew, 0)                    # |nodes| x k
            R = Jint[nodes]                                       # rows of the flipped nodes
            rr = np.repeat(np.arange(nodes.size), np.diff(R.indptr))
            np.add.at(H, R.indices, R.data[:, None].astype(H.dtype) * delta[rr])
        return flips

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
                fixed_point = final_sweep_fraction >= 1.0
                break
    e2_final = (S * H).sum(axis=0)
    return LocalSearchResult(np.rint(S.T).astype(np.int8).copy(), passes, e2_start, e2_final, q_counter[0])


def local_search_scalar(Jint: sp.csr_matrix, cand: np.ndarray, seed: int, colours: np.ndarray, **kw):
    """One node at a time restatement of local_search (tiny inputs; the pin of the vectorised one)."""
    Jint = sp.csr_matrix(Jint)
    n = Jint.shape[0]
    cand = np.atleast_2d(cand)
    k = cand.shape[0]
    C = int(colours.max()) + 1
    S = [list(map(int, row)) for row in cand]
    if Jint.dtype.kind in "iu":
        dense = Jint.toarray().astype(np.int64)
    else:   # FP32 couplings, FP64 field sums
        dense = Jint.toarray().astype(np.float32).astype(np.float64)
    q = 0

    def field(c, i):
        return sum(dense[i, j] * S[c][j] for j in range(n))

    def run_pass(sel_id, frac):
        nonlocal q
        flips = [False] * k
        order = class_order(seed, q, C)
        q += 1
        for cls in order:
            for i in range(n):
                if colours[i] != cls:
                    continue
                if frac < 1.0 and not selected(seed, sel_id, np.array([i]), frac)[0]:
                    continue
                for c in range(k):
                    if S[c][i] * field(c, i) < 0:
                        S[c][i] = -S[c][i]
                        flips[c] = True
        return np.array(flips)

    max_iterations = kw.get("max_iterations", 1000)
    subset = kw.get("subset_fraction", 0.9)
    final_frac = kw.get("final_sweep_fraction", 1.0)
    active = np.ones(k, dtype=bool)
    passes = 0
    for it in range(max_iterations):
        passes += 1
        active &= run_pass(it, subset)
        if not active.any():
            break
    for r in range(kw.get("num_rounds", 25)):
        stop = False
        for _ in range(kw.get("final_sweep_iters", 25)):
            if not run_pass(FINAL_PASS_BASE + r, final_frac).any():
                stop = final_frac >= 1.0
                break
        if stop:
            break
    return np.array(S, dtype=np.int8), passes


# ---------------------------------------------------------------------------------------------
# A numpy restatement of the build's Lanczos schedule (FP32/FP64 vectors, FP64 coefficients,
# per-step Ritz test, explicit restart from the Ritz vector every ncv steps).
# ---------------------------------------------------------------------------------------------


def lanczos_n(matvec, v0: np.ndarray, tol: float, maxiter: int, ncv: int, dtype=np.float32):
    """Largest eigenpair of a symmetric operator; returns (theta, y, steps, res)."""
    from scipy.linalg import eigh_tridiagonal
    n = v0.size
    steps = 0
    v = v0.astype(np.float64)
    v = (v / np.linalg.norm(v)).astype(dtype)
    while True:
        V = np.zeros((ncv + 1, n), dtype=dtype)
        V[0] = v
        a, b = [], []
        for j in range(ncv):
            w = matvec(V[j]).astype(dtype)
            steps += 1
            h = V[: j + 1].astype(np.float64) @ w.astype(np.float64)
            a.append(h[j])
            w = (w.astype(np.float64) - h @ V[: j + 1].astype(np.float64)).astype(dtype)
            beta = float(np.linalg.norm(w.astype(np.float64)))
            th, Sv = eigh_tridiagonal(np.array(a), np.array(b), select="i", select_range=(j, j))
            res = beta * abs(Sv[-1, 0])
            if res <= tol or steps >= maxiter:
                y = Sv[:, 0] @ V[: j + 1].astype(np.float64)
                return float(th[0]), y / np.linalg.norm(y), steps, res
            b.append(beta)
            V[j + 1] = (w.astype(np.float64) / beta).astype(dtype)
        y = Sv[:, 0] @ V[:ncv].astype(np.float64)
        v = (y / np.linalg.norm(y)).astype(dtype)


# ---------------------------------------------------------------------------------------------
# The CPU path of SpectralAnnealing.solve: scipy ARPACK + numpy predictors + the local search
# ---------------------------------------------------------------------------------------------


@dataclass
class SweepResult:
    best_energy: float
    best_spins: np.ndarray | None
    rows: list = field(default_factory=list)
    candidates: list = field(default_factory=list)
    sweep_seconds: float = 0.0
    post_sweep_seconds: float = 0.0
    matvecs: int = 0


def cpu_sweep(J: sp.spmatrix, *, precision: int = 32, num_alphas: int = 128, schedule: str = "sqrt",
              interval=(0.0, 1.0), pred_mode: str = "diagonal", pred_tol: float = 1e-3, shift_eps: float = 1e-3,
              eigsh_tol: float = 1e-6, eigsh_maxiter: int = 1000, ncv: int | None = None,
              use_local_search: bool = True, max_local_iterations: int = 1000, seed: int = 42,
              alpha_limit: int | None = None) -> SweepResult:
    """The reference's solve (models/spectral_annealing.py:262-525) on the CPU.

    The eigensolve is scipy's ARPACK eigsh on the S_rho LinearOperator with which='SA' (the
    reference's call, :671-677); ARPACK counts restarts, so its maxiter is eigsh_maxiter / (ncv - 1)
    to give the same operator-application budget. alpha_limit stops after that many alphas (a
    bounded sample for timing).
    """
    dtype = np.float32 if precision == 32 else np.float64
    t0 = time.time()
    Jint = integer_couplings(J)
    Jn, d, c = normalize(J, dtype)
    n = J.shape[0]
    alphas = alpha_values(num_alphas, interval, schedule)
    if alpha_limit is not None:
        alphas = alphas[: int(alpha_limit)]
    ncv_eff = min(33, n - 1) if ncv is None else ncv
    arpack_maxiter = max(1, int(math.ceil(eigsh_maxiter / max(1, ncv_eff - 1))))
    y_last = y_prev = lam_prev = a_prev = None
    seen: dict = {}
    rows, cands = [], []
    matvecs = [0]
    for ai, alpha in enumerate(alphas):
        s, dsum, g = alpha_scaling(d, alpha)
        rho = shift_rho(lam_prev, g, shift_eps)
        J64 = Jn.astype(dtype)

        def mv(x, s=s, rho=rho):
            matvecs[0] += 1
            x = x.astype(dtype)
            return (dtype(rho) * x - s * (J64 @ (s * x))).astype(dtype)

        op = spla.LinearOperator((n, n), matvec=mv, dtype=dtype)
        pred_res, pred_ok = float("nan"), None
        if ai == 0 or y_last is None:
            y0 = normal_vector(n, seed, ai * 4)
            y0 /= np.linalg.norm(y0)
        elif pred_mode == "prev":
            y0 = y_last
        else:
            y0 = (diagonal_predictor(y_last, d, alpha - a_prev) if pred_mode == "diagonal"
                  else secant_predictor(y_last, y_prev))
            if y0 is None:
                y0 = normal_vector(n, seed, ai * 4 + 1)
                y0 /= np.linalg.norm(y0)
            v = n_matvec(Jn, s, y0)
            matvecs[0] += 1
            lam_est = float(y0 @ v)
            pred_res = float(np.linalg.norm(v - lam_est * y0))
            pred_ok = not pred_res > pred_tol
            if not pred_ok:
                y0 = y_last
        try:
            w, V = spla.eigsh(op, k=1, which="SA", tol=eigsh_tol, maxiter=arpack_maxiter, v0=y0.astype(dtype),
                              ncv=ncv_eff)
        except spla.ArpackNoConvergence as e:
            if e.eigenvalues.size == 0:
                continue
            w, V = e.eigenvalues, e.eigenvectors
        mu = float(w[0])
        lam = rho - mu
        y = V[:, 0].astype(np.float64)
        if y_last is not None and y @ y_last < 0:
            y = -y
        y /= np.linalg.norm(y)
        y_prev, y_last, lam_prev, a_prev = y_last, y, lam, alpha
        spins = round_spins(s.astype(np.float64) * y)
        key = spin_bits(spins).tobytes()
        if key in seen:
            cand_idx = seen[key]
            raw = cands[cand_idx][1]
        else:
            raw = energy_int64(Jint, spins)
            cand_idx = len(cands)
            seen[key] = cand_idx
            cands.append((spins, raw))
        rows.append(dict(alpha=alpha, raw_energy=float(raw), energy=float(raw), lambda1=lam, degree_scale=dsum,
                         energy_bound=-0.5 * lam * dsum * c, pred_residual=pred_res, pred_accepted=pred_ok,
                         candidate=cand_idx))
    t_sweep = time.time()
    best_e, best_s = float("inf"), None
    if use_local_search and cands:
        K = np.stack([cs for cs, _ in cands])
        res = local_search(Jint, K, seed, max_iterations=max_local_iterations)
        improved = []
        for i, (cs, raw) in enumerate(cands):
            ref_e = -int(res.e2_final[i]) // 2
            improved.append((cs, raw) if ref_e > raw else (res.spins[i], ref_e))
        for r in rows:
            e = improved[r["candidate"]][1]
            r["energy"] = float(e)
            r["local_search_iterations"] = res.passes
            if e < best_e:
                best_e, best_s = float(e), improved[r["candidate"]][0]
    else:
        for r in rows:
            if r["energy"] < best_e:
                best_e, best_s = r["energy"], cands[r["candidate"]][0]
    t_end = time.time()
    return SweepResult(best_e, best_s, rows, cands, t_sweep - t0, t_end - t_sweep, matvecs[0])


# ---------------------------------------------------------------------------------------------
# Guarded local-search policies (models/local_search_guards/*, models/spectral_annealing_safe.py)
# ---------------------------------------------------------------------------------------------
GUARD_SALT = 0x6A09E667F3BCC908

# local_search_guards/policies.py:21-37 (the settings the policies run with)
GUARD_SETTINGS = dict(subset_fraction=0.9, final_sweep_fraction=1.0, final_sweep_iters=25, final_sweep_rounds=25,
                      absolute_tolerance=1e-10, relative_tolerance=1e-12, safe_patience=3,
                      minimum_subset_fraction=0.01, batch_shrink_factor=0.5, batch_growth_factor=1.25,
                      oracle_max_steps_multiplier=100)


@dataclass
class GuardResult:
    spins: np.ndarray          # k x n int8
    iterations: int            # the policy's reported count (policies.py:47-49)
    energies: np.ndarray       # final energies: int64 (integer couplings) or float64 (units of J / c)
    energies_start: np.ndarray
    extra: dict = field(default_factory=dict)


def _guard_setup(J: sp.spmatrix, cand: np.ndarray):
    """(couplings, S, H, E, exact): integer couplings exactly, others as FP32 values in FP64; S k x n,
    H = fields k x n, E = -sum S H / 2 per candidate."""
    J = sp.csr_matrix(J)
    exact = J.dtype.kind in "iu" or bool(np.all(J.data == np.rint(J.data)))
    if exact:
        Jc = sp.csr_matrix((J.data.astype(np.int64), J.indices, J.indptr), shape=J.shape)
        S = np.atleast_2d(cand).astype(np.int64).copy()
    else:
        Jc = sp.csr_matrix((J.data.astype(np.float32).astype(np.float64), J.indices, J.indptr), shape=J.shape)
        S = np.atleast_2d(cand).astype(np.float64).copy()
    H = np.asarray((Jc @ S.T).T)
    e2 = (S * H).sum(axis=1)
    E = -(e2 // 2) if exact else -0.5 * e2
    return Jc, S, H, E, exact


def guard_tolerance(E: np.ndarray, scale: float, abs_tol: float, rel_tol: float) -> np.ndarray:
    """models/safe_update_policy.py:19-23 on the normalised energy E / scale."""
    En = np.asarray(E, dtype=np.float64) / scale
    return abs_tol + rel_tol * np.maximum(1.0, np.abs(En))


def _flip_threshold(E, scale, abs_tol, rel_tol, exact):
    """A flip (change 2 s h) passes iff s h < -thr: thr = tol scale / 2, floored for integer fields."""
    x = guard_tolerance(E, scale, abs_tol, rel_tol) * scale * 0.5
    return np.minimum(np.floor(x), 2147483647.0) if exact else x


def guard_safe(J: sp.spmatrix, cand: np.ndarray, seed: int, *, adaptive: bool, scale: float = 1.0,
               max_iterations: int = 1000, subset_fraction: float = 0.9, final_sweep_fraction: float = 1.0,
               final_sweep_iters: int = 25, final_sweep_rounds: int = 25, absolute_tolerance: float = 1e-10,
               relative_tolerance: float = 1e-12, patience: int = 3, minimum_subset_fraction: float = 0.01,
               batch_shrink_factor: float = 0.5, batch_growth_factor: float = 1.25) -> GuardResult:
    """Fixed / Adaptive Guard (safe_flip_refinement, models/spectral_annealing_safe.py:1035-1335; phases
    :223-797; acceptance and adaptation models/safe_update_policy.py:4-95 with the per-candidate scope
    safe_common.py:16-44 uses). Each round, candidate c proposes the Jacobi flips of the nodes selected by
    (mix3(seed ^ GUARD_SALT, round, i) >> 11) < floor(f_c 2^53) (a hashed share f_c of the nodes in place
    of the reference's random permutation prefix) whose spin disagrees with the sign of its field; the
    proposal is kept iff its exact energy change is below minus the tolerance of the current energy.
    Main phase: max_iterations rounds at subset_fraction; final phase: final_sweep_iters *
    final_sweep_rounds rounds at final_sweep_fraction; each phase starts with every candidate active
    and leaves when none is (patience). scale: c, the normalisation of the reference's energies."""
    Jc, S, H, E, exact = _guard_setup(J, cand)
    E0 = E.copy()
    k, n = S.shape
    gseed = u64(int(seed) ^ GUARD_SALT)
    nodes = np.arange(n, dtype=np.uint64)
    rid = 0
    rounds_done = {}
    for phase, rounds, f0 in (("main", int(max_iterations), float(subset_fraction)),
                              ("final", int(final_sweep_iters) * int(final_sweep_rounds), float(final_sweep_fraction))):
        done = 0
        if rounds > 0:
            active = np.ones(k, dtype=bool)
            streak = np.zeros(k, dtype=np.int64)
            frac = np.full(k, f0)
            for _ in range(rounds):
                if not active.any():
                    break
                hsel = mix3(gseed, u64(rid), nodes) >> np.uint64(11)
                thresh = np.array([int(f * TWO53) if a else 0 for f, a in zip(frac, active)], dtype=np.uint64)
                flip = (hsel[None, :] < thresh[:, None]) & ((S * H) < 0)
                counts = flip.sum(axis=1)
                T = np.where(flip, S, 0)
                inner = np.asarray((Jc @ T.T).T)
                dE = (2 * T * H).sum(axis=1) - (2 * T * inner).sum(axis=1)
                has = active & (counts > 0)
                tol = guard_tolerance(E, scale, absolute_tolerance, relative_tolerance)
                accepted = has & (np.asarray(dE, dtype=np.float64) / scale < -tol)
                rejected = has & ~accepted
                no_flips = active & ~has
                if accepted.any():
                    S[accepted] = np.where(flip[accepted], -S[accepted], S[accepted])
                    H = np.asarray((Jc @ S.T).T)
                    E = E + np.where(accepted, dE, 0)
                if adaptive:
                    newf = np.where(rejected, np.maximum(minimum_subset_fraction, frac * batch_shrink_factor), frac)
                    newf = np.where(accepted, np.minimum(f0, frac * batch_growth_factor), newf)
                    shrunk = rejected & (newf < frac)
                    stalled = no_flips | (rejected & ~shrunk)
                    frac = newf
                else:
                    stalled = active & ~accepted
                streak = np.where(active, np.where(accepted, 0, np.where(stalled, streak + 1, 0)), streak)
                active = active & (streak < patience)
                rid += 1
                done += 1
        rounds_done[phase] = done
    return GuardResult(np.rint(S).astype(np.int8), rounds_done["main"], E, E0, dict(final_rounds=rounds_done["final"]))


def largest_first_colouring(indptr: np.ndarray, indices: np.ndarray) -> np.ndarray:
    """greedy_largest_first_coloring (local_search_guards/common.py:136-159): nodes by descending degree,
    then ascending index; each takes the smallest colour no coloured neighbour has."""
    n = indptr.size - 1
    deg = np.diff(indptr)
    order = np.lexsort((np.arange(n), -deg))
    colours = np.full(n, -1, dtype=np.int64)
    for v in order:
        nb = indices[indptr[v]:indptr[v + 1]]
        used = set(int(u) for u in colours[nb] if u >= 0)
        c = 0
        while c in used:
            c += 1
        colours[v] = c
    return colours.astype(np.int32)


def guard_independent_set(J: sp.spmatrix, cand: np.ndarray, seed: int, *, scale: float = 1.0,
                          max_updates: int = 1000 + 625, absolute_tolerance: float = 1e-10,
                          relative_tolerance: float = 1e-12) -> GuardResult:
    """Independent-Set (conflict_free.py:82-296): colour classes of the largest-first colouring visited
    in np.random.default_rng(seed).permutation order, a new permutation per sweep; in a step every
    active candidate flips the class's nodes whose single-flip change 2 s h is below minus the
    tolerance of its current energy; after a sweep without a flip a candidate is certified (inactive).
    At most max_updates steps; iterations = steps run (kernel_updates)."""
    Jc, S, H, E, exact = _guard_setup(J, cand)
    E0 = E.copy()
    k, n = S.shape
    J = sp.csr_matrix(J)
    colours = largest_first_colouring(J.indptr, J.indices)
    C = int(colours.max()) + 1
    classes = [np.flatnonzero(colours == q) for q in range(C)]
    rng = np.random.default_rng(int(seed))
    order = rng.permutation(C)
    pos = 0
    active = np.ones(k, dtype=bool)
    sweep = np.zeros(k, dtype=np.int64)
    steps = 0
    for _ in range(int(max_updates)):
        if not active.any():
            break
        nodes = classes[int(order[pos])]
        thr = _flip_threshold(E, scale, absolute_tolerance, relative_tolerance, exact)
        V = S[:, nodes] * H[:, nodes]
        flip = active[:, None] & (V < -thr[:, None])
        E = E + (2 * V * flip).sum(axis=1)
        sweep += flip.sum(axis=1)
        if flip.any():
            S[:, nodes] = np.where(flip, -S[:, nodes], S[:, nodes])
            H = np.asarray((Jc @ S.T).T)
        steps += 1
        pos += 1
        if pos == C:
            active &= ~(sweep == 0)
            sweep[:] = 0
            order = rng.permutation(C)
            pos = 0
    return GuardResult(np.rint(S).astype(np.int8), steps, E, E0, dict(colours=colours))


def guard_sequential(J: sp.spmatrix, cand: np.ndarray, *, scale: float = 1.0, max_steps: int | None = None,
                     absolute_tolerance: float = 1e-10, relative_tolerance: float = 1e-12) -> GuardResult:
    """Sequential-Oracle (sequential_oracle.py:22-121): each step every candidate flips its node of
    smallest s_i h_i (first on ties) while 2 s h is below minus the tolerance of its energy, until none
    improves (a one-flip local optimum). iterations = steps with a flip; converged False if flips still
    improve after max_steps (default 100 n) steps."""
    Jc, S, H, E, exact = _guard_setup(J, cand)
    E0 = E.copy()
    k, n = S.shape
    max_steps = 100 * n if max_steps is None else int(max_steps)
    steps = 0
    converged = False
    rows = np.arange(k)
    while True:
        V = S * H
        best = np.argmin(V, axis=1)
        bv = V[rows, best]
        thr = _flip_threshold(E, scale, absolute_tolerance, relative_tolerance, exact)
        improving = bv < -thr
        if not improving.any():
            converged = True
            break
        if steps == max_steps:
            break
        for c in np.flatnonzero(improving):
            S[c, best[c]] = -S[c, best[c]]
            E[c] = E[c] + 2 * bv[c]
        H = np.asarray((Jc @ S.T).T)
        steps += 1
    return GuardResult(np.rint(S).astype(np.int8), steps, E, E0, dict(converged=converged))
