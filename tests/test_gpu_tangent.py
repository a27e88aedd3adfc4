"""The eigenvector tangent and its adjoint (deflated MINRES on the GPU) against the oracle (GPU)."""
import numpy as np
import pytest
import torch

from oracle import specann_oracle as O
from specann_amd import graphs
from specann_amd.device import DeviceProblem
from specann_amd.tangent import CLAMP, DeflatedMinres, EigenContext, LeadingEigvec, TangentPredictor

pytestmark = pytest.mark.gpu


def _graph(n=400, deg=10, seed=7):
    return graphs.erdos_renyi(n, deg / (n - 1), seed=seed, kind="pm1")


def _dense_pair(J, alpha, dt):
    Jn, d, c = O.normalize(J, dt)
    s, _, _ = O.alpha_scaling(d.astype(np.float64), alpha)
    N = O.dense_n(Jn.astype(np.float64), s)
    w, V = np.linalg.eigh(N)
    y = V[:, -1]
    if y[np.argmax(np.abs(y))] < 0:
        y = -y
    return Jn, d, N, w[-1], y


@pytest.mark.parametrize("tdtype,rtol", [(torch.float64, 1e-9), (torch.float32, 2e-3)])
@pytest.mark.parametrize("k", [1, 3, 8])
def test_minres_iterates_match_oracle(gpu, tdtype, rtol, k):
    """k iterations of the device MINRES equal k iterations of the oracle's restatement."""
    J = _graph()
    alpha = 0.45
    dt = np.float32 if tdtype == torch.float32 else np.float64
    Jn, d, N, lam, y = _dense_pair(J, alpha, dt)
    print(f"minres dtype={tdtype} k={k} n={J.shape[0]}", flush=True)
    prob = DeviceProblem(J, tdtype)
    s = torch.empty(prob.n, dtype=tdtype, device=gpu)
    prob.alpha_scaling(alpha, CLAMP, s)
    r = O.normal_vector(prob.n, 11, 0)
    mr = DeflatedMinres(prob, chunk=2)
    yt = torch.from_numpy(y.astype(dt)).to(gpu)
    rt = torch.from_numpy(r.astype(dt)).to(gpu)
    res = mr.solve(s, yt, rt, lam, 0.0, k)
    assert res.iterations == k
    x_ref, itn, phibar, beta1 = O.minres_deflated(lambda v: N @ v, y.astype(dt).astype(np.float64),
                                                  r.astype(dt).astype(np.float64), lam, 0.0, k)
    x = mr.x.cpu().numpy().astype(np.float64)
    assert np.linalg.norm(x - x_ref) <= rtol * np.linalg.norm(x_ref)
    assert res.rnorm == pytest.approx(phibar, rel=max(rtol, 1e-9))
    assert abs(float(x @ y)) <= 10 * rtol * np.linalg.norm(x)


def test_minres_zero_rhs_and_maxit_zero(gpu):
    J = _graph(200, 8, 3)
    prob = DeviceProblem(J, torch.float64)
    s = torch.empty(prob.n, dtype=torch.float64, device=gpu)
    prob.alpha_scaling(0.3, CLAMP, s)
    _, _, _, lam, y = _dense_pair(J, 0.3, np.float64)
    yt = torch.from_numpy(y).to(gpu)
    mr = DeflatedMinres(prob)
    res = mr.solve(s, yt, torch.zeros_like(yt), lam, 1e-8, 10)      # zero right-hand side: x = 0 at once
    assert res.iterations == 0 and float(torch.abs(mr.x).max()) == 0.0
    res = mr.solve(s, yt, torch.ones_like(yt), lam, 1e-8, 0)
    assert res.iterations == 0 and float(torch.abs(mr.x).max()) == 0.0


@pytest.mark.parametrize("alpha", [0.2, 0.7])
def test_tangent_matches_dense_pinv_and_finite_differences(gpu, alpha):
    """dy/dalpha from the device MINRES equals the dense pseudo-inverse formula, and central finite
    differences (h = 1e-3) of FP64 eigenvectors (GPU Lanczos and the oracle's dense eigh)."""
    J = _graph(500, 12, 9)
    print(f"tangent alpha={alpha}", flush=True)
    prob = DeviceProblem(J, torch.float64)
    eig = EigenContext(prob, tol=1e-12, adjoint_steps=400, adjoint_tol=1e-13)
    y, lam, _ = eig.solve(alpha)
    tp = TangentPredictor(prob, steps=400, tol=1e-13)
    tp.compute(eig.s, y, lam)
    t = tp.full_tangent(y).cpu().numpy()
    Jn, d, c = O.normalize(J, np.float64)
    lam_o, y_o, yp_o = O.tangent(Jn, d, alpha)
    sgn = 1.0 if float(y.cpu().numpy() @ y_o) > 0 else -1.0
    assert abs(lam - lam_o) <= 1e-10 * abs(lam_o)
    assert np.linalg.norm(t - sgn * yp_o) <= 1e-7 * np.linalg.norm(yp_o)
    h = 1e-3
    yp, _, _ = eig.solve(alpha + h, sign_ref=y)
    ym, _, _ = eig.solve(alpha - h, sign_ref=y)
    fd = (yp - ym).cpu().numpy() / (2 * h)
    assert np.linalg.norm(fd - t) <= 1e-4 * np.linalg.norm(t)
    # the oracle's own FD agrees as well (sign-aligned dense eigenvectors)
    _, _, Np, _, yo_p = _dense_pair(J, alpha + h, np.float64)
    _, _, Nm, _, yo_m = _dense_pair(J, alpha - h, np.float64)
    yo_p *= np.sign(yo_p @ y_o)
    yo_m *= np.sign(yo_m @ y_o)
    assert np.linalg.norm((yo_p - yo_m) / (2 * h) - yp_o) <= 1e-4 * np.linalg.norm(yp_o)


@pytest.mark.parametrize("tdtype,rtol", [(torch.float64, 1e-6), (torch.float32, 5e-2)])
def test_autograd_adjoint_matches_finite_differences(gpu, tdtype, rtol):
    """LeadingEigvec.backward (the adjoint MINRES) against central differences of
    L(alpha) = c . y(alpha) + 0.3 lambda(alpha)."""
    J = _graph(600, 12, 4)
    print(f"adjoint dtype={tdtype}", flush=True)
    prob = DeviceProblem(J, tdtype)
    eig = EigenContext(prob, adjoint_steps=600, adjoint_tol=1e-13 if tdtype == torch.float64 else 1e-6)
    cvec = torch.from_numpy(O.normal_vector(prob.n, 3, 1)).to(gpu)
    alpha0 = 0.55
    y0, _, _ = eig.solve(alpha0)
    ref = y0.clone()
    a = torch.tensor(alpha0, dtype=torch.float64, requires_grad=True)
    y, lam = LeadingEigvec.apply(a, eig, ref)
    loss = (cvec * y.double()).sum() + 0.3 * lam
    loss.backward()
    g = float(a.grad)

    def L(al):
        yy, ll, _ = eig.solve(al, sign_ref=ref)
        return float((cvec * yy.double()).sum()) + 0.3 * ll

    # the exact derivative: fourth-order central differences of the dense FP64 eigenpairs (h = 3e-4,
    # truncation and rounding below 1e-9 here), the eigenvector's sign aligned with the device's y0
    Jn, d, _ = O.normalize(J, np.float64)
    cv = cvec.cpu().numpy()
    y0d = _dense_pair(J, alpha0, np.float64)[4]
    sgn = np.sign(float(ref.double().cpu().numpy() @ y0d))

    def L_dense(al):
        _, _, _, lam_d, y_d = _dense_pair(J, al, np.float64)
        return float(cv @ (sgn * np.sign(y_d @ y0d) * y_d)) + 0.3 * lam_d

    hd = 3e-4
    exact = (-L_dense(alpha0 + 2 * hd) + 8 * L_dense(alpha0 + hd) - 8 * L_dense(alpha0 - hd)
             + L_dense(alpha0 - 2 * hd)) / (12 * hd)
    # the same differences of device solves: each eigenvector carries about residual / gap of error
    # (tolerance 1e-12, gap 2.3e-3 here), which the differences amplify by 1.5 / h, so this check is
    # looser than the one against the exact derivative
    h = 1e-3 if tdtype == torch.float64 else 1e-2
    fd = (-L(alpha0 + 2 * h) + 8 * L(alpha0 + h) - 8 * L(alpha0 - h) + L(alpha0 - 2 * h)) / (12 * h)
    print(f"adjoint gradient {g:.12g} exact {exact:.12g} device finite differences {fd:.12g}", flush=True)
    assert g == pytest.approx(exact, rel=rtol, abs=rtol * 1e-2)
    assert g == pytest.approx(fd, rel=max(rtol, 1e-5), abs=rtol * 1e-2)


def test_tangent_predictor_beats_diagonal(gpu):
    """The tangent-corrected start vector has a smaller eigen-residual at the next alpha than the
    diagonal predictor's (the quantity the predictor gate tests)."""
    J = graphs.erdos_renyi(20000, 20 / 19999, seed=12, kind="pm1")
    prob = DeviceProblem(J, torch.float32)
    eig = EigenContext(prob, tol=1e-7)
    a0, a1 = 0.50, 0.53
    y0, lam0, _ = eig.solve(a0)
    tp = TangentPredictor(prob, steps=12, tol=1e-3)
    tp.compute(eig.s, y0, lam0)
    d = prob.d.double()
    base = torch.pow(torch.clamp(d, min=CLAMP), 0.5 * (a1 - a0))
    diag = base * y0.double()
    tang = base * (y0.double() + 2.0 * lam0 * (a1 - a0) * tp.z.double())
    y1, lam1, _ = eig.solve(a1, sign_ref=y0)

    def residual(v):
        v = (v / torch.linalg.norm(v)).to(prob.dtype)
        w = torch.empty_like(v)
        prob.matvec(eig.s, v, 0.0, 1.0, w, torch.empty_like(v))
        th = float(v.double() @ w.double())
        return float(torch.linalg.norm(w.double() - th * v.double()))

    r_diag, r_tang = residual(diag), residual(tang)
    print(f"predictor residuals: diagonal {r_diag:.3e} tangent {r_tang:.3e}", flush=True)
    assert r_tang < 0.5 * r_diag
