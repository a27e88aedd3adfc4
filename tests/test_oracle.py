"""Pins of the CPU oracle against closed forms, brute force and library routines (CPU only)."""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import specann_oracle as O
from specann_amd import graphs


def small_signed(n=12, p=0.4, seed=0):
    return graphs.erdos_renyi(n, p, seed=seed, kind="pm1")


def test_splitmix64_known_values():
    # SplitMix64 outputs for seed 0 (stream x_k = k * golden gamma, finalised): published sequence
    # 0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F
    x = np.array([0, 0x9E3779B97F4A7C15, (2 * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF], dtype=np.uint64)
    out = O.splitmix64(x)
    assert [int(v) for v in out] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_normal_vector_statistics():
    v = O.normal_vector(200000, 3, 1)
    assert abs(v.mean()) < 0.01 and abs(v.std() - 1) < 0.01
    assert not np.array_equal(v[:100], O.normal_vector(100, 3, 2))


def test_normalize_closed_form():
    J = sp.csr_matrix(np.array([[0, -2, 1], [-2, 0, 0], [1, 0, 0]], dtype=float))
    Jn, d, c = O.normalize(J, np.float64)
    assert c == 3.0
    np.testing.assert_allclose(d, [1.0, 2 / 3, 1 / 3])
    np.testing.assert_allclose(Jn.toarray(), J.toarray() / 3)


def test_alpha_values_schedule():
    a = O.alpha_values(128)
    assert a[0] == 0.0 and a[-1] == 1.0 and len(a) == 128
    np.testing.assert_allclose(np.square(a), np.linspace(0, 1, 128))
    b = O.alpha_values(51, (0, 1.25), "sqrt")
    assert b[-1] == pytest.approx(np.sqrt(1.25))
    assert O.alpha_values(5, (0, 1), "linear") == [0.0, 0.25, 0.5, 0.75, 1.0]


def test_alpha_scaling_and_gershgorin():
    J = small_signed(30)
    Jn, d, c = O.normalize(J, np.float64)
    for alpha in (0.0, 0.3, 1.0, 1.2):
        s, dsum, g = O.alpha_scaling(d, alpha)
        np.testing.assert_allclose(s, d ** (-alpha / 2))
        assert dsum == pytest.approx(np.sum(d ** alpha))
        # the similarity D^{a/2} N D^{-a/2} = D^{-a} J has absolute row sums d^{1-a}: g bounds lambda_max
        lam, _ = O.leading_eigpair(Jn, s)
        assert lam <= g * (1 + 1e-12)
        assert g == pytest.approx(max(1.0, d.min() ** (1 - alpha)) if alpha > 1 else 1.0)


def test_shift_rho_cases():
    assert O.shift_rho(None, 1.0) == pytest.approx(1.001)
    assert O.shift_rho(float("nan"), 1.0) == pytest.approx(1.001)
    assert O.shift_rho(-3.0, 1.0) == pytest.approx(1.001)
    assert O.shift_rho(2.0, 1.0) == pytest.approx(2.002)
    assert O.shift_rho(None, -1.0, eps=1e-3) == pytest.approx(1.001)


def test_shift_rho_build_equals_closed_forms():
    """The build's compute_shift_rho (the host scalar of the sweep) on the closed-form cases, and
    equal to the oracle's over a grid of inputs (models/spectral_annealing.py:56-67)."""
    from specann_amd.spectral_annealing import compute_shift_rho as build
    assert build(None, 1.0) == pytest.approx(1.001)
    assert build(float("inf"), 0.8) == pytest.approx(0.8008)
    assert build(0.0, 0.7) == pytest.approx(0.7007)                 # lambda1_prev <= 0: g alone
    assert build(2.0, 1.0) == pytest.approx(2.002)                  # max(lambda1_prev, g)
    assert build(0.5, 1.0, eps=0.01) == pytest.approx(1.01)
    assert build(None, -2.0) == pytest.approx(1.001)                # rho <= 0: (1 + eps) max(g, 1)
    assert build(None, 0.0, eps=0.5) == pytest.approx(1.5)
    for lp in (None, float("nan"), -1.0, 0.0, 0.3, 1.7):
        for g in (-1.0, 0.0, 0.2, 1.0, 3.0):
            for eps in (1e-3, 0.25):
                assert build(lp, g, eps) == O.shift_rho(lp, g, eps)


def test_check_eigenpair_closed_forms():
    """_check_eigenpair (models/spectral_annealing.py:679-691): the tolerance is max(eps/2, 1e-5)
    relative to g, and any non-finite value fails."""
    y = np.ones(4) / 2
    g = 0.8
    assert O.check_eigenpair(g, y, g, 1e-3)
    assert O.check_eigenpair(g * (1 + 4.9e-4), y, g, 1e-3)             # within eps / 2 = 5e-4
    assert not O.check_eigenpair(g * (1 + 5.1e-4), y, g, 1e-3)
    assert O.check_eigenpair(g * (1 + 0.9e-5), y, g, 1e-6)             # eps / 2 < 1e-5: 1e-5 applies
    assert not O.check_eigenpair(g * (1 + 1.1e-5), y, g, 1e-6)
    assert O.check_eigenpair(-5.0, y, g, 1e-3)                         # no lower limit
    assert not O.check_eigenpair(float("nan"), y, g, 1e-3)
    assert not O.check_eigenpair(float("inf"), y, g, 1e-3)
    bad = y.copy()
    bad[2] = np.nan
    assert not O.check_eigenpair(0.5, bad, g, 1e-3)
    bad[2] = np.inf
    assert not O.check_eigenpair(0.5, bad, g, 1e-3)


def test_build_eigenpair_check_matches_oracle():
    """The sweep's own check (SpectralAnnealing._check) rejects exactly what the oracle rejects."""
    from types import SimpleNamespace
    from specann_amd._lib import SA_BROKEN, SA_CONVERGED, EigensolverBreakdown
    from specann_amd.spectral_annealing import SpectralAnnealing
    y = np.ones(3)
    for eps in (1e-3, 1e-6):
        sa = SpectralAnnealing("i", "d", shift_eps=eps, verbose=False)
        for theta in (0.7, 0.8, 0.8 * (1 + 4e-4), 0.8 * (1 + 6e-4), 0.8 * (1 + 2e-5), float("nan"), float("inf")):
            ok = O.check_eigenpair(theta, y, 0.8, eps)
            res = SimpleNamespace(status=SA_CONVERGED, theta=theta)
            if ok:
                sa._check(res, 0.8)
            else:
                with pytest.raises(EigensolverBreakdown):
                    sa._check(res, 0.8)
        with pytest.raises(EigensolverBreakdown, match="breakdown"):
            sa._check(SimpleNamespace(status=SA_BROKEN, theta=0.5), 0.8)


def test_energies_float_exact_and_brute_force():
    """fp32_energies (models/spectral_annealing.py:962-976): -1/2 c s^T (J/c)_fp32 s with FP64
    accumulation. Pinned by exact rational arithmetic on the FP32 couplings, by the integer energies
    when J/c is exact in FP32 (c a power of two), and by the ground state of a tiny instance."""
    from fractions import Fraction
    rng = np.random.default_rng(4)
    n = 10
    A = np.triu(rng.uniform(-2.0, 2.0, (n, n)) * (rng.random((n, n)) < 0.5), 1)
    J = sp.csr_matrix(A + A.T)
    Jn, d, c = O.normalize(J, np.float32)
    S = np.where(rng.random((n, 6)) < 0.5, 1, -1)
    got = O.energies_float(Jn, S, c)
    dense = Jn.astype(np.float32).toarray()
    for k in range(S.shape[1]):
        exact = Fraction(0)
        for i in range(n):
            for j in range(n):
                if dense[i, j] != 0:
                    exact += Fraction(float(dense[i, j])) * int(S[i, k]) * int(S[j, k])
        exact = -Fraction(1, 2) * Fraction(c) * exact
        assert abs(got[k] - float(exact)) <= 1e-12 * max(1.0, abs(float(exact)))
    # integer couplings with c = 8 (J/c exact in FP32): equal to the int64 energies
    B = np.zeros((6, 6))
    for (i, j, w) in [(0, 1, 3), (0, 2, -2), (1, 2, 1), (2, 3, 2), (3, 4, -1), (4, 5, 1), (0, 5, 2), (1, 4, 1)]:
        B[i, j] = B[j, i] = w
    B[3, 5] = B[5, 3] = 2
    B[0, 4] = B[4, 0] = 1    # node 0 then has the largest absolute degree, 3 + 2 + 2 + 1 = 8
    Ji = sp.csr_matrix(B)
    Jn, d, c = O.normalize(Ji, np.float32)
    assert c == 8.0
    S = np.array(list(itertools.product([-1, 1], repeat=6))).T
    e = O.energies_float(Jn, S, c)
    Jint = O.integer_couplings(Ji)
    ints = [O.energy_int64(Jint, S[:, k]) for k in range(S.shape[1])]
    assert [float(v) for v in ints] == list(e)
    # brute-force ground state agrees with the minimum over all 2^n states
    assert e.min() == min(ints)


def test_matvec_dense():
    J = small_signed(40)
    Jn, d, c = O.normalize(J, np.float64)
    s, _, _ = O.alpha_scaling(d, 0.6)
    x = np.random.default_rng(0).standard_normal(40)
    N = np.diag(s) @ Jn.toarray() @ np.diag(s)
    np.testing.assert_allclose(O.n_matvec(Jn, s, x), N @ x, atol=1e-14)
    np.testing.assert_allclose(O.s_matvec(Jn, s, x, 1.5), 1.5 * x - N @ x, atol=1e-14)


def test_leading_eigpair_arpack_matches_dense():
    J = graphs.erdos_renyi(4000, 10 / 3999, seed=1, kind="neg")
    Jn, d, c = O.normalize(J, np.float64)
    s, _, _ = O.alpha_scaling(d, 0.5)
    lam_a, y_a = O.leading_eigpair(Jn, s)               # n > 3000: ARPACK
    w, V = np.linalg.eigh(O.dense_n(Jn, s))
    assert lam_a == pytest.approx(w[-1], abs=1e-10)
    assert abs(abs(y_a @ V[:, -1]) - 1) < 1e-8


def test_lanczos_restatement_converges_to_dense():
    J = small_signed(300, 0.05, seed=2)
    Jn, d, c = O.normalize(J, np.float64)
    s, _, _ = O.alpha_scaling(d, 0.4)
    lam, y = O.leading_eigpair(Jn, s)
    th, yl, steps, res = O.lanczos_n(lambda x: O.n_matvec(Jn, s, x), np.ones(300), 1e-10, 5000, 33, np.float64)
    assert th == pytest.approx(lam, abs=1e-9)
    assert abs(abs(yl @ y) - 1) < 1e-6


def test_diagonal_predictor_closed_form_and_first_order():
    rng = np.random.default_rng(1)
    d = rng.uniform(0.1, 1.0, 50)
    y = rng.standard_normal(50)
    y /= np.linalg.norm(y)
    p = O.diagonal_predictor(y, d, 0.2)
    ref = d ** 0.1 * y
    np.testing.assert_allclose(p, ref / np.linalg.norm(ref))
    # first order: y + da * P (1/2 ln d (.) y) + O(da^2)
    da = 1e-4
    q = O.diagonal_predictor(y, d, da)
    u = 0.5 * np.log(d) * y
    lin = y + da * (u - (y @ u) * y)
    assert np.linalg.norm(q - lin) < 10 * da ** 2
    assert O.diagonal_predictor(np.zeros(5), np.ones(5), 0.1) is None


def test_secant_predictor_cases():
    yl = np.array([1.0, 0.0])
    yp = np.array([-np.sqrt(0.5), -np.sqrt(0.5)])     # opposite sign: aligned to (0.707, 0.707)
    out = O.secant_predictor(yl, yp)
    ref = yl + (yl - np.array([np.sqrt(0.5), np.sqrt(0.5)]))
    np.testing.assert_allclose(out, ref / np.linalg.norm(ref))
    assert O.secant_predictor(yl, None) is yl
    assert O.secant_predictor(yl, yl.copy()) is yl


def test_tangent_matches_finite_differences():
    J = small_signed(60, 0.2, seed=3)
    Jn, d, c = O.normalize(J, np.float64)
    alpha, h = 0.5, 1e-4
    lam, y, yp = O.tangent(Jn, d, alpha)

    def vec(a):
        s, _, _ = O.alpha_scaling(d, a)
        _, v = O.leading_eigpair(Jn, s)
        return v if v @ y > 0 else -v

    fd = (vec(alpha + h) - vec(alpha - h)) / (2 * h)
    assert np.linalg.norm(fd - yp) <= 1e-5 * max(1.0, np.linalg.norm(yp))


def _small_problem(n=120, seed=3):
    J = graphs.erdos_renyi(n, 12 / (n - 1), seed=seed, kind="pm1")
    Jn, d, c = O.normalize(J, np.float64)
    return Jn, d


def test_minres_deflated_matches_scipy_minres_iterates():
    """The restated recurrence equals scipy's MINRES on the explicitly deflated operator, iterate by
    iterate (same algorithm, FP64), and reaches the pseudo-inverse solution."""
    from scipy.sparse.linalg import LinearOperator, minres
    Jn, d = _small_problem()
    alpha = 0.4
    s, _, _ = O.alpha_scaling(d, alpha)
    N = O.dense_n(Jn, s)
    w, V = np.linalg.eigh(N)
    lam, y = w[-1], V[:, -1]
    rng = np.random.default_rng(0)
    r = rng.standard_normal(N.shape[0])
    P = np.eye(N.shape[0]) - np.outer(y, y)
    A = P @ (N - lam * np.eye(N.shape[0])) @ P
    op = LinearOperator(A.shape, matvec=lambda v: A @ v, dtype=np.float64)
    b = P @ r
    for k in (1, 2, 5, 9):
        x, itn, _, _ = O.minres_deflated(lambda v: N @ v, y, r, lam, 0.0, k)
        xs, _ = minres(op, b, maxiter=k, rtol=0.0)
        xs = xs - (y @ xs) * y
        assert itn == k
        assert np.linalg.norm(x - xs) <= 1e-9 * np.linalg.norm(xs), k
    x, itn, phibar, beta1 = O.minres_deflated(lambda v: N @ v, y, r, lam, 1e-13, 400)
    ref = np.linalg.pinv(N - lam * np.eye(N.shape[0]), rcond=1e-12) @ b
    assert np.linalg.norm(x - ref) <= 1e-7 * np.linalg.norm(ref)
    assert phibar <= 1e-13 * beta1 * 10


def test_tangent_minres_equals_dense_tangent():
    Jn, d = _small_problem(150, 5)
    alpha = 0.6
    lam, y, yp = O.tangent(Jn, d, alpha)
    t, z, itn = O.tangent_minres(Jn, d, alpha, y, lam, steps=500, tol=1e-13)
    assert np.linalg.norm(t - yp) <= 1e-7 * np.linalg.norm(yp)
    # a planted sign error in the correction term is visible
    bad = t - 4.0 * lam * z
    assert np.linalg.norm(bad - yp) > 1e-3 * np.linalg.norm(yp)


def test_round_and_bits():
    x = np.array([0.0, -0.0, -1e-30, 2.0, -3.0])
    np.testing.assert_array_equal(O.round_spins(x), [1, 1, -1, 1, -1])
    s = np.where(np.random.default_rng(0).random(130) < 0.5, 1, -1)
    words = O.spin_bits(s)
    assert words.size == 3
    for i in range(130):
        assert ((int(words[i // 64]) >> (i % 64)) & 1) == (s[i] > 0)
    assert O.bits_signature(words) != O.bits_signature(O.spin_bits(-s))


def test_energy_int64_brute_force_and_ground_state():
    J = graphs.erdos_renyi(14, 0.5, seed=5, kind="pm1")
    Jint = O.integer_couplings(J)
    Jd = J.toarray()
    best = None
    for bits in itertools.product([-1, 1], repeat=14):
        s = np.array(bits)
        e = O.energy_int64(Jint, s)
        assert e == -0.5 * s @ Jd @ s
        assert e == O.energy_int64(Jint, -s)
        best = e if best is None else min(best, e)
    # the local search never goes below the ground state and ends 1-flip optimal
    cands = np.stack([O.round_spins(O.normal_vector(14, 1, i)) for i in range(4)])
    res = O.local_search(Jint, cands, seed=7)
    for c in range(4):
        e = O.energy_int64(Jint, res.spins[c])
        assert e >= best
        h = Jint @ res.spins[c].astype(np.int64)
        assert np.all(res.spins[c] * h >= 0)


def test_greedy_colouring_matches_networkx():
    nx = pytest.importorskip("networkx")
    J = graphs.barabasi_albert(500, 4, seed=3)
    col = O.greedy_colouring(J.indptr, J.indices, 42)
    order = O._order_desc(*O.colour_priority(J.indptr, 42))
    G = nx.from_scipy_sparse_array(J)
    ref = nx.greedy_color(G, strategy=lambda g, colors: list(order))
    np.testing.assert_array_equal(col, [ref[i] for i in range(500)])
    rows = np.repeat(np.arange(500), np.diff(J.indptr))
    assert not np.any(col[rows] == col[J.indices])


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_local_search_vectorised_equals_scalar(seed):
    J = graphs.erdos_renyi(24, 0.3, seed=seed, kind="pm1")
    Jint = O.integer_couplings(J)
    colours = O.greedy_colouring(J.indptr, J.indices, seed)
    cands = np.stack([O.round_spins(O.normal_vector(24, seed, i)) for i in range(3)])
    vec = O.local_search(Jint, cands, seed, colours, max_iterations=50)
    sc, passes = O.local_search_scalar(Jint, cands, seed, colours, max_iterations=50)
    np.testing.assert_array_equal(vec.spins, sc)
    assert vec.passes == passes


def _real_couplings(n, p, seed):
    rng = np.random.default_rng(seed)
    A = np.triu(rng.uniform(-1.5, 1.5, (n, n)) * (rng.random((n, n)) < p), 1)
    return sp.csr_matrix(A + A.T)


@pytest.mark.parametrize("seed", [1, 2])
def test_float_local_search_vectorised_equals_scalar(seed):
    """The floating-point local search (FP32 J / c, FP64 fields): the vectorised restatement equals the
    one-node-at-a-time one bit for bit, never raises the energy, ends 1-flip optimal after a full pass,
    and its FP64 fields are exact (equal to exact rational sums)."""
    from fractions import Fraction
    J = _real_couplings(24, 0.3, seed)
    Jn, d, c = O.normalize(J, np.float32)
    assert np.abs(Jn.data).min() >= 2.0 ** -30
    colours = O.greedy_colouring(Jn.indptr, Jn.indices, seed)
    cands = np.stack([O.round_spins(O.normal_vector(24, seed, i)) for i in range(3)])
    vec = O.local_search(Jn, cands, seed, colours, max_iterations=50)
    sc, passes = O.local_search_scalar(Jn, cands, seed, colours, max_iterations=50)
    np.testing.assert_array_equal(vec.spins, sc)
    assert vec.passes == passes
    e0 = O.energies_float(Jn, cands.T, c)
    e1 = O.energies_float(Jn, vec.spins.T, c)
    assert np.all(e1 <= e0 + 1e-12)
    np.testing.assert_allclose(-0.5 * c * vec.e2_final, e1, rtol=1e-12)
    H = Jn.astype(np.float64) @ vec.spins.T.astype(np.float64)
    assert np.all(vec.spins.T * H >= 0)
    dense = Jn.toarray().astype(np.float64)
    for i in range(24):                                     # the FP64 field is the exact sum
        exact = sum(Fraction(float(dense[i, j])) * int(vec.spins[0, j]) for j in range(24))
        assert H[i, 0] == float(exact)


def test_local_search_planted_mistake_is_caught():
    """A wrong field update (sign flipped) in a copy of the pass must break 1-flip optimality."""
    J = graphs.erdos_renyi(40, 0.2, seed=4, kind="pm1")
    Jint = O.integer_couplings(J)
    cands = np.stack([O.round_spins(O.normal_vector(40, 4, i)) for i in range(3)])
    good = O.local_search(Jint, cands, 3)
    bad_J = sp.csr_matrix((-Jint.data, Jint.indices, Jint.indptr), shape=Jint.shape)
    bad = O.local_search(bad_J, cands, 3)
    h = Jint @ bad.spins.T.astype(np.int64)
    assert not np.all(bad.spins.T * h >= 0)
    h = Jint @ good.spins.T.astype(np.int64)
    assert np.all(good.spins.T * h >= 0)


def test_cpu_sweep_small_invariants():
    J = graphs.erdos_renyi(200, 10 / 199, seed=42, kind="neg")
    res = O.cpu_sweep(J, num_alphas=8, seed=42)
    Jint = O.integer_couplings(J)
    assert res.best_energy == O.energy_int64(Jint, res.best_spins)
    for r in res.rows:
        # lambda1 s^T (J/c) s <= ... : the energy bound is a lower bound of every energy
        assert r["energy"] >= r["energy_bound"] - 1e-6
        assert r["energy"] <= r["raw_energy"]
