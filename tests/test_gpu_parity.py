"""Parity of the per-alpha control flow on the GPU against the oracle: the predicted start vectors
(diagonal, secant, tangent) built inside sa_lanczos_start, the predictor gate's residual (Lanczos
step 0), the maxiter path's true residual, the eigenpair rejection and the cold retry, FP64
eigenpairs at a relative residual of 1e-10, and the rounding agreement at the eigen tolerance."""
import numpy as np
import pytest
import torch

from oracle import specann_oracle as O
from specann_amd import graphs
from specann_amd._lib import SA_REJECTED
from specann_amd.device import DeviceProblem
from specann_amd.eigsh import LanczosSolver
from specann_amd.spectral_annealing import SpectralAnnealing, unpack_bits
from specann_amd.tangent import TangentPredictor

pytestmark = pytest.mark.gpu


def _setup(gpu, J, alpha_prev, alpha, tdtype=torch.float32):
    """Device problem, s(alpha), the FP64 oracle eigenvectors at alpha_prev (y_last) and at the alpha
    before it (y_prev), both rounded to the vectors' type."""
    dt = np.float32 if tdtype == torch.float32 else np.float64
    Jn, d, c = O.normalize(J, dt)
    prob = DeviceProblem(J, tdtype)
    s = torch.empty(prob.n, dtype=tdtype, device=gpu)
    prob.alpha_scaling(alpha, 1e-12, s)
    _, y_last = O.leading_eigpair(Jn, O.alpha_scaling(d, alpha_prev)[0])
    _, y_prev = O.leading_eigpair(Jn, O.alpha_scaling(d, alpha_prev - 0.05)[0])
    if y_prev @ y_last < 0:
        y_prev = -y_prev
    return prob, s, Jn, d, y_last.astype(dt), y_prev.astype(dt)


def _start_vector(lz):
    """The unit start vector sa_lanczos_start built: basis row 0 times sigma_0 = row 0 / ||row 0||."""
    st = lz.work.head()
    v = lz.V[0].double().cpu().numpy()
    return v / st.r[3], st


def _gate_residual(Jn, d, alpha, y0):
    """||N y0 - (y0 . N y0) y0|| in FP64 for a unit y0 (models/spectral_annealing.py:605-612)."""
    s, _, _ = O.alpha_scaling(d, alpha)
    v = O.n_matvec(Jn, s, y0)
    return float(np.linalg.norm(v - (y0 @ v) * y0))


@pytest.mark.parametrize("alpha_prev,alpha", [(0.3, 0.35), (0.6, 0.7), (0.9, 1.0)])
def test_diagonal_start_and_gate_residual(gpu, alpha_prev, alpha):
    """The fused diagonal predictor D^(da/2) y_last (utils/predictors.py:6-26) and beta_0 of the gated
    first step, which the sweep reports as pred_residual, against the oracle."""
    J = graphs.barabasi_albert(2500, 4, seed=11, kind="pm1")
    print(f"diagonal predictor alpha {alpha_prev} -> {alpha}", flush=True)
    prob, s, Jn, d, y_last, _ = _setup(gpu, J, alpha_prev, alpha)
    lz = LanczosSolver(prob, 20)
    lz.begin(1e-6, 200, bound=3.0, pred_tol=1.0)
    lz.start(s, torch.from_numpy(y_last).to(gpu), d=prob.d, pexp=0.5 * (alpha - alpha_prev), gate=1)
    got, st = _start_vector(lz)
    ref = O.diagonal_predictor(y_last, d, alpha - alpha_prev)
    assert np.linalg.norm(got - ref) <= 1e-6
    res = lz.run(s, 200, first_check=1)
    want = _gate_residual(Jn, d, alpha, ref)
    assert abs(res.first_beta - want) <= 2e-6 + 1e-4 * want, (res.first_beta, want)


def test_gate_rejects_above_pred_tol(gpu):
    """A predicted start whose residual exceeds pred_tol stops the solve with SA_REJECTED after step 0,
    and the residual it reports is the oracle's."""
    J = graphs.barabasi_albert(2500, 4, seed=11, kind="pm1")
    prob, s, Jn, d, y_last, _ = _setup(gpu, J, 0.3, 0.6)
    lz = LanczosSolver(prob, 20)
    ref = O.diagonal_predictor(y_last, d, 0.3)
    want = _gate_residual(Jn, d, 0.6, ref)
    assert want > 1e-6
    lz.begin(1e-6, 200, bound=3.0, pred_tol=0.5 * want)
    lz.start(s, torch.from_numpy(y_last).to(gpu), d=prob.d, pexp=0.15, gate=1)
    res = lz.run(s, 200, first_check=1)
    assert res.status == SA_REJECTED and res.steps == 1
    assert abs(res.first_beta - want) <= 2e-6 + 1e-4 * want


def test_secant_start_vector(gpu):
    """The secant combination 2 y_last - sign y_prev (gamma = 1, utils/predictors.py:29-56) as the
    sweep queues it, against the oracle's secant_predictor."""
    J = graphs.erdos_renyi(2000, 10 / 1999, seed=3, kind="pm1")
    prob, s, Jn, d, y_last, y_prev = _setup(gpu, J, 0.5, 0.55)
    for flip in (1.0, -1.0):                  # y_prev with either sign: the predictor aligns it
        yp = (flip * y_prev).astype(np.float32)
        lz = LanczosSolver(prob, 20)
        lz.begin(1e-6, 200, bound=3.0, pred_tol=1.0)
        sa = SpectralAnnealing("t", "d", pred_mode="secant", verbose=False)
        sa._secant_start(prob, lz, s, torch.from_numpy(y_last).to(gpu), torch.from_numpy(yp).to(gpu))
        got, _ = _start_vector(lz)
        ref = O.secant_predictor(y_last, yp)
        assert np.linalg.norm(got - ref) <= 2e-6
        res = lz.run(s, 200, first_check=1)
        want = _gate_residual(Jn, d, 0.55, ref)
        assert abs(res.first_beta - want) <= 2e-6 + 1e-4 * want
    # y_prev equal to y_last: the secant step is zero and the predictor returns y_last itself
    lz = LanczosSolver(prob, 20)
    lz.begin(1e-6, 200, bound=3.0, pred_tol=1.0)
    yl = torch.from_numpy(y_last).to(gpu)
    SpectralAnnealing("t", "d", pred_mode="secant", verbose=False)._secant_start(prob, lz, s, yl, yl.clone())
    got, _ = _start_vector(lz)
    assert np.linalg.norm(got - y_last / np.linalg.norm(y_last)) <= 1e-6


def test_tangent_start_vector(gpu):
    """The tangent predictor D^(da/2) (y + 2 lambda da z) as the sweep queues it: z from the device
    MINRES, the combination against the same formula on the host, and the step-0 residual against
    the oracle. It also predicts the next eigenvector better than the diagonal predictor does."""
    J = graphs.erdos_renyi(2000, 10 / 1999, seed=3, kind="pm1")
    a0, a1 = 0.5, 0.55
    prob, s, Jn, d, y_last, _ = _setup(gpu, J, a0, a1)
    s0 = torch.empty_like(s)
    prob.alpha_scaling(a0, 1e-12, s0)
    lam0, _ = O.leading_eigpair(Jn, O.alpha_scaling(d, a0)[0])
    tp = TangentPredictor(prob, steps=8, tol=1e-6)
    yl = torch.from_numpy(y_last).to(gpu)
    tp.compute(s0, yl, lam0)
    z = tp.z.double().cpu().numpy()
    da = a1 - a0
    lz = LanczosSolver(prob, 20)
    lz.begin(1e-6, 200, bound=3.0, pred_tol=1.0)
    lz.start(s, yl, d=prob.d, pexp=0.5 * da, y=tp.z, cy=2.0 * tp.lam * da, gate=1)
    got, _ = _start_vector(lz)
    w = np.power(np.maximum(d.astype(np.float64), 1e-12), 0.5 * da)
    ref = w * (y_last.astype(np.float64) + 2.0 * lam0 * da * z)
    ref /= np.linalg.norm(ref)
    assert np.linalg.norm(got - ref) <= 2e-6
    res = lz.run(s, 200, first_check=1)
    want = _gate_residual(Jn, d, a1, ref)
    assert abs(res.first_beta - want) <= 2e-6 + 1e-4 * want
    _, y1 = O.leading_eigpair(Jn, O.alpha_scaling(d, a1)[0])
    diag = O.diagonal_predictor(y_last, d, da)
    assert 1 - abs(ref @ y1) < 1 - abs(diag @ y1)


class _Capture:
    """Wraps SpectralAnnealing._eigensolve to keep each alpha's final (unnormalised) Ritz vector."""

    def __init__(self, monkeypatch):
        self.vectors = {}
        orig = SpectralAnnealing._eigensolve

        def wrapped(sa, prob, lz, s, x_rand, alpha_idx, alpha, *a):
            eo = orig(sa, prob, lz, s, x_rand, alpha_idx, alpha, *a)
            self.vectors[alpha_idx] = (lz.Y.double().cpu().numpy() / eo.y_norm, eo)
            return eo

        monkeypatch.setattr(SpectralAnnealing, "_eigensolve", wrapped)


def test_maxiter_path_reports_true_residual(gpu, monkeypatch):
    """With a small eigsh_maxiter the solves stop unconverged: the row's residual is the true
    ||N y - theta y|| / ||y|| of the final Ritz vector, eigsh_converged is residual <= tol and the
    energy bound is NaN (models/spectral_annealing.py:640-647, :326-329)."""
    J = graphs.erdos_renyi(3000, 20 / 2999, seed=5, kind="pm1")
    cap = _Capture(monkeypatch)
    prob = DeviceProblem(J, torch.float32)
    sa = SpectralAnnealing("t", "d", random_seed=3, num_alphas=6, eigsh_maxiter=12, eigsh_ncv=8, verbose=False)
    out = sa.run_on_device(prob)
    Jn, d, c = O.normalize(J, np.float32)
    unconverged = 0
    for r in out.rows:
        y, eo = cap.vectors[r["alpha_index"]]
        s, _, _ = O.alpha_scaling(d, r["alpha"])
        true = float(np.linalg.norm(O.n_matvec(Jn, s, y) - r["lambda1"] * y) / np.linalg.norm(y))
        assert r["eigsh_iterations"] <= 12
        if r["residual"] is None:
            assert r["eigsh_converged"] and r["residual_estimate"] <= 1e-6
            continue
        unconverged += 1
        assert r["residual"] == pytest.approx(true, rel=1e-3, abs=1e-7)
        assert r["eigsh_converged"] == (r["residual"] <= 1e-6)
        if not r["eigsh_converged"]:
            assert np.isnan(r["energy_bound"])
    assert unconverged >= 3


def test_spurious_eigenvalue_is_rejected_then_skipped(gpu, monkeypatch):
    """An eigenvalue above the Gershgorin bound fails _check_eigenpair (:679-691): the alpha is retried
    cold, rejected again and skipped. Forced from the host by under-reporting g for one alpha."""
    J = graphs.erdos_renyi(800, 12 / 799, seed=5, kind="pm1")
    prob = DeviceProblem(J, torch.float32)
    orig = DeviceProblem.alpha_scaling
    target = 3

    def scaling(self, alpha, clamp, out):
        dsum, g = orig(self, alpha, clamp, out)
        if abs(alpha - np.sqrt(target / 7.0)) < 1e-12:
            g *= 0.01                           # lambda1 (0.3 to 0.6 here) now exceeds the reported bound
        return dsum, g

    monkeypatch.setattr(DeviceProblem, "alpha_scaling", scaling)
    sa = SpectralAnnealing("t", "d", random_seed=3, num_alphas=8, verbose=False)
    out = sa.run_on_device(prob)
    assert sa._alphas_skipped == 1 and sa._alpha_errors == 0
    assert [r["alpha_index"] for r in out.rows] == [0, 1, 2, 4, 5, 6, 7]


def test_nonfinite_ritz_vector_is_retried_cold(gpu, monkeypatch):
    """A zero or non-finite final Ritz vector counts as a breakdown inside the retry scope: the alpha is
    solved again from a random vector and kept (eigsh_retry), as the reference does for a non-finite
    eigenvector (:616-636, :679-691)."""
    J = graphs.erdos_renyi(800, 12 / 799, seed=5, kind="pm1")
    prob = DeviceProblem(J, torch.float32)
    orig = LanczosSolver.final_vector
    calls = {"n": 0}

    def final_vector(self, y_last):
        calls["n"] += 1
        nrm, dot = orig(self, y_last)
        return (float("nan"), dot) if calls["n"] == 3 else (nrm, dot)

    monkeypatch.setattr(LanczosSolver, "final_vector", final_vector)
    sa = SpectralAnnealing("t", "d", random_seed=3, num_alphas=6, verbose=False)
    out = sa.run_on_device(prob)
    assert len(out.rows) == 6 and sa._alphas_skipped == 0
    assert [r["eigsh_retry"] for r in out.rows] == [False, False, True, False, False, False]


@pytest.mark.parametrize("reorth", ["selective", "full"])
def test_fp64_eigenpairs_relative_residual_1e10(gpu, monkeypatch, reorth):
    """SpecAnn-64 at eigsh_tol = 1e-10: every eigenpair's true relative residual ||N y - lambda y|| /
    (||N|| ||y||) is at most 1e-10 (||N|| <= g = 1 on this schedule), lambda agrees with dense eigh to
    1e-11, and the angle to the dense eigenvector is within the residual-over-gap bound."""
    J = graphs.barabasi_albert(1500, 6, seed=3, kind="pm1")
    cap = _Capture(monkeypatch)
    prob = DeviceProblem(J, torch.float64)
    sa = SpectralAnnealing("t", "d", random_seed=42, precision=64, num_alphas=10, eigsh_tol=1e-10, verbose=False,
                           reorth=reorth)
    out = sa.run_on_device(prob)
    Jn, d, c = O.normalize(J, np.float64)
    for r in out.rows:
        assert r["eigsh_converged"]
        y, eo = cap.vectors[r["alpha_index"]]
        s, _, g = O.alpha_scaling(d, r["alpha"])
        rel = float(np.linalg.norm(O.n_matvec(Jn, s, y) - r["lambda1"] * y) / np.linalg.norm(y))
        assert rel <= 1e-10 * max(1.0, g), (r["alpha"], rel)
        # the row reports the true residual the solver checked on the GPU
        assert r["residual"] is not None and r["residual"] == pytest.approx(rel, rel=1e-3, abs=1e-14)
        w, V = np.linalg.eigh(O.dense_n(Jn, s))
        lam, yd, gap = w[-1], V[:, -1], w[-1] - w[-2]
        assert abs(r["lambda1"] - lam) <= 1e-11 * max(1.0, abs(lam))
        y = y / np.linalg.norm(y)
        angle = float(np.linalg.norm(y - (y @ yd) * yd))     # sin of the angle, without cancellation
        print(f"alpha {r['alpha']:.3f}: residual {rel:.2e}, gap {gap:.3e}, angle {angle:.2e}")
        assert angle <= 2.0 * rel / gap + 1e-12


def test_rounding_agrees_above_the_eigen_tolerance(gpu, monkeypatch):
    """FP32 sweep at the default eigsh_tol: each alpha's rounded candidate equals the sign of the FP64
    oracle eigenvector (the reference's rounding, :345) at every node where |y_i| exceeds 1e-5, the
    FP32 relative residual target; the GPU's own eigenvector has a true residual within it."""
    J = graphs.barabasi_albert(3000, 5, seed=7, kind="pm1")
    cap = _Capture(monkeypatch)
    prob = DeviceProblem(J, torch.float32)
    out = SpectralAnnealing("t", "d", random_seed=42, num_alphas=12, verbose=False).run_on_device(prob)
    Jn, d, c = O.normalize(J, np.float32)
    bits = out.cand_bits.cpu().numpy()
    worst = 0.0
    for r in out.rows:
        s, _, _ = O.alpha_scaling(d, r["alpha"])
        yg, _ = cap.vectors[r["alpha_index"]]
        assert np.linalg.norm(O.n_matvec(Jn, s, yg) - r["lambda1"] * yg) <= 1e-5
        _, y = O.leading_eigpair(Jn, s)
        got = unpack_bits(bits[r["candidate"]], J.shape[0]).astype(np.int64)
        ref = O.round_spins(y).astype(np.int64)
        if np.sum(got * ref) < 0:
            ref = -ref
            y = -y
        differ = got != ref
        if differ.any():
            worst = max(worst, float(np.max(np.abs(y[differ]))))
    print(f"largest |y_i| at a node whose rounding differs: {worst:.3g}")
    assert worst <= 1e-5


@pytest.mark.parametrize("pred", ["tangent", "diagonal"])
def test_fp32_eigenpairs_relative_residual_1e5(gpu, pred):
    """FP32 sweeps at the reference's tol 1e-6: every converged alpha's eigenpair has a true relative
    residual ||N y - lambda y|| / ||y|| <= 1e-5 (the FP32 bar), recomputed here in FP64 on the host from
    the sweep's own unit vectors, and lambda within 1e-5 of the dense eigenvalue."""
    J = graphs.erdos_renyi(1024, 40 / 1023, seed=42, kind="neg")
    Jn, d, c = O.normalize(J, np.float32)
    got = {}

    def hook(ai, alpha, prob, lz, s, scale, eo):
        got[ai] = (alpha, (lz.Y * scale).double().cpu().numpy(), eo.lambda1, eo.converged)

    sa = SpectralAnnealing("t", "synthetic", random_seed=42, num_alphas=128, pred_mode=pred, verbose=False)
    sa.eigvec_hook = hook
    sa.run_on_device(DeviceProblem(J, torch.float32))
    assert len(got) == 128
    worst = 0.0
    for ai, (alpha, y, lam, conv) in got.items():
        if not conv:
            continue
        s, _, _ = O.alpha_scaling(d, alpha)
        rel = float(np.linalg.norm(O.n_matvec(Jn, s, y) - lam * y) / np.linalg.norm(y))
        worst = max(worst, rel)
        if ai % 16 == 0:
            w = np.linalg.eigvalsh(O.dense_n(Jn, s))
            assert abs(lam - w[-1]) <= 1e-5
    assert worst <= 1e-5, worst
