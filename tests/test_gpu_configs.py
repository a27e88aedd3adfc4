"""The benchmark configurations at their real shapes on the GPU, one test each, with the checks that
hold at any size: eigenvalues against FP64 ARPACK, true residuals on the CPU, exact energies of the
candidates, the FP32/FP64 energy gap, converged fractions, and the packed local search bit-exact
against the oracle on a sub-batch.

    C2  signed Barabasi-Albert n = 2^20, m = 20, J = +-1
    C3  256 Erdos-Renyi / Barabasi-Albert instances, n log-uniform in [2^10, 2^16), J = -1, packed
    C4  Erdos-Renyi n = 2^22, mean degree 40, J = +-1 (tangent predictor, FP64 tangent equation)
    C5  Barabasi-Albert n = 2^25, m = 10, J = -1
"""
import time

import numpy as np
import pytest
import scipy.sparse.linalg as spla
import torch

import bench
from oracle import specann_oracle as O
from specann_amd import graphs
from specann_amd.device import DeviceProblem
from specann_amd.spectral_annealing import SpectralAnnealing, unpack_bits

pytestmark = pytest.mark.gpu


def log(msg):
    print(f"[config] {msg}", flush=True)


class _Capture:
    """Keeps the unit Ritz vector (on the host) of the alphas listed in `keep`."""

    def __init__(self, monkeypatch, keep):
        self.vectors = {}
        orig = SpectralAnnealing._eigensolve

        def wrapped(sa, prob, lz, s, x_rand, alpha_idx, alpha, *a):
            eo = orig(sa, prob, lz, s, x_rand, alpha_idx, alpha, *a)
            if alpha_idx in keep:
                self.vectors[alpha_idx] = lz.Y.double().cpu().numpy() / eo.y_norm
            return eo

        monkeypatch.setattr(SpectralAnnealing, "_eigensolve", wrapped)


def _true_residual(Jn32, d, alpha, lam, y):
    s, _, _ = O.alpha_scaling(d, alpha)
    return float(np.linalg.norm(O.n_matvec(Jn32, s, y) - lam * y))


def _check_candidate_energies(J, out, sample):
    """Exact int64 energies (scipy on the host) of a sample of the sweep's distinct candidates."""
    Jint = O.integer_couplings(J)
    bits = out.cand_bits.cpu().numpy()
    for c in sample:
        if c < bits.shape[0]:
            s = unpack_bits(bits[c], J.shape[0]).astype(np.int64)
            assert out.raw_energies[c] == O.energy_int64(Jint, s), c
    assert out.best_energy == O.energy_int64(Jint, out.best_spins.astype(np.int64))


def test_c2_signed_ba_fp32_and_fp64(gpu, monkeypatch):
    t0 = time.time()
    J = graphs.barabasi_albert_native(1 << 20, 20, seed=42, kind="pm1")
    log(f"C2 graph n={J.shape[0]} nnz={J.nnz} in {time.time() - t0:.1f}s")
    sampled = (16, 48, 96, 127)
    cap = _Capture(monkeypatch, sampled)
    p32 = DeviceProblem(J, torch.float32)
    p32.autotune()
    out32 = SpectralAnnealing("c2", "synthetic", random_seed=42, pred_mode="tangent", verbose=False
                              ).run_on_device(p32)
    y32 = dict(cap.vectors)
    del p32
    p64 = DeviceProblem(J, torch.float64)
    p64.autotune()
    out64 = SpectralAnnealing("c2", "synthetic", random_seed=42, precision=64, verbose=False).run_on_device(p64)
    del p64
    log(f"C2 sweeps FP32 {out32.sweep_time:.2f}s FP64 {out64.sweep_time:.2f}s; best {out32.best_energy} / "
        f"{out64.best_energy}")
    assert len(out32.rows) == 128 and len(out64.rows) == 128
    assert all(r["eigsh_converged"] for r in out32.rows) and all(r["eigsh_converged"] for r in out64.rows)
    assert abs(out32.best_energy - out64.best_energy) <= 1e-3 * abs(out64.best_energy)
    Jn64, d64, _ = O.normalize(J, np.float64)
    Jn32, d32, _ = O.normalize(J, np.float32)
    by_alpha = {r["alpha_index"]: r for r in out32.rows}
    by_alpha64 = {r["alpha_index"]: r for r in out64.rows}
    for i in sampled:
        r, r64 = by_alpha[i], by_alpha64[i]
        s, _, _ = O.alpha_scaling(d64, r["alpha"])
        op = spla.LinearOperator(J.shape, matvec=lambda x, s=s: s * (Jn64 @ (s * x)), dtype=np.float64)
        w, _ = spla.eigsh(op, k=1, which="LA", tol=1e-12, v0=y32[i], ncv=40)
        lam = float(w[0])
        log(f"alpha {r['alpha']:.4f}: ARPACK {lam:.12f} FP32 {r['lambda1']:.12f} FP64 {r64['lambda1']:.12f}")
        assert abs(r["lambda1"] - lam) <= 2e-6 * max(1.0, abs(lam))
        assert abs(r64["lambda1"] - lam) <= 1e-9 * max(1.0, abs(lam))
        assert _true_residual(Jn32, d32, r["alpha"], r["lambda1"], y32[i]) <= 1e-5
    _check_candidate_energies(J, out32, range(0, out32.num_candidates, max(1, out32.num_candidates // 6)))


def test_c4_er22_tangent_sweep_and_tangent_equation(gpu, monkeypatch):
    """The first 16 alphas of the 128-alpha schedule with the tangent predictor: every alpha converged,
    true residuals <= 1e-5 at sampled alphas (which also covers the half-precision shadow basis of the
    reorthogonalisation), exact energies; then the FP64 tangent at one alpha satisfies its defining
    equation (N - lambda I) y' = -P N' y on the host to the MINRES tolerance."""
    from specann_amd.tangent import CLAMP, DeflatedMinres
    t0 = time.time()
    J = graphs.erdos_renyi_native(1 << 22, 40 / ((1 << 22) - 1), seed=42, kind="pm1")
    log(f"C4 graph n={J.shape[0]} nnz={J.nnz} in {time.time() - t0:.1f}s")
    sampled = (1, 8, 15)
    cap = _Capture(monkeypatch, sampled)
    p32 = DeviceProblem(J, torch.float32)
    p32.autotune()
    out = SpectralAnnealing("c4", "synthetic", random_seed=42, pred_mode="tangent", num_alphas=16,
                            schedule_interval=(0.0, 15 / 127), verbose=False).run_on_device(p32)
    log(f"C4 16 alphas: {out.sweep_time:.2f}s, {out.matvecs} + {out.tangent_matvecs} applications")
    assert np.allclose([r["alpha"] for r in out.rows], O.alpha_values(128)[:16])
    # an alpha may stop at the reference's maxiter (1000 applications) unconverged; its row then carries
    # the true residual, as the reference's does (models/spectral_annealing.py:640-647)
    for r in out.rows:
        log(f"alpha {r['alpha']:.4f}: {r['eigsh_iterations']} applications, converged {r['eigsh_converged']}, "
            f"residual {r['residual']}")
        assert r["eigsh_converged"] or (r["residual"] is not None and r["eigsh_iterations"] >= 1000)
    assert sum(r["eigsh_converged"] for r in out.rows) >= 14
    Jn32, d32, _ = O.normalize(J, np.float32)
    for i in sampled:
        r = next(q for q in out.rows if q["alpha_index"] == i)
        res = _true_residual(Jn32, d32, r["alpha"], r["lambda1"], cap.vectors[i])
        log(f"alpha {r['alpha']:.4f}: true residual {res:.3e} (estimate {r['residual_estimate']:.3e})")
        if r["eigsh_converged"]:
            assert res <= 1e-5
        else:
            assert res == pytest.approx(r["residual"], rel=1e-2)
    _check_candidate_energies(J, out, range(0, out.num_candidates, max(1, out.num_candidates // 4)))
    del p32
    # FP64 tangent at alpha = 0.3: y from the FP64 solver, z from 60 deflated MINRES steps
    from specann_amd.eigsh import ScaledOperator, eigsh
    alpha = 0.3
    p64 = DeviceProblem(J, torch.float64)
    s = torch.empty(p64.n, dtype=torch.float64, device=gpu)
    p64.alpha_scaling(alpha, CLAMP, s)
    w, v, _ = eigsh(ScaledOperator(p64, s, None), which="LA", tol=1e-11, maxiter=20000)
    lam = float(w[0])
    y = v[:, 0].contiguous()
    u = torch.empty_like(y)
    from specann_amd._lib import call
    from specann_amd.device import ptr, stream
    call("sa_tangent_rhs_f64", p64.n, ptr(y), ptr(p64.d), CLAMP, ptr(u), stream())
    mr = DeflatedMinres(p64, chunk=10)
    res = mr.solve(s, y, u, lam, 1e-8, 60)
    z = mr.x.cpu().numpy()
    yh, uh = y.cpu().numpy(), u.cpu().numpy()
    Jn64, d64, _ = O.normalize(J, np.float64)
    sv, _, _ = O.alpha_scaling(d64, alpha)
    pu = uh - (yh @ uh) * yh
    # MINRES solves P (N - lambda I) P z = P u; its reported residual is ||P u - P (N - lambda I) z||
    r = O.n_matvec(Jn64, sv, z) - lam * z
    r -= (yh @ r) * yh
    true = float(np.linalg.norm(pu - r))
    log(f"FP64 tangent: lambda {lam:.10f}, MINRES {res.iterations} its, reported {res.rnorm:.3e}, host {true:.3e}, "
        f"|Pu| {np.linalg.norm(pu):.3e}")
    assert true <= max(2.0 * res.rnorm, 1e-12) + 1e-10 * np.linalg.norm(pu)
    assert abs(z @ yh) <= 1e-8 * np.linalg.norm(z)


def test_c3_packed_batch(gpu):
    """256 instances packed block-diagonally, 128 alphas: per-instance eigenvalues against the oracle on a
    sample, exact energies, every alpha converged; the packed local search bit-exact against the oracle
    on a sub-batch of eight instances."""
    from specann_amd.batch import BatchProblem, BatchSweep, pack
    t0 = time.time()
    Js = bench.make_batch(bench.WORKLOADS["c3"])
    log(f"C3 {len(Js)} instances, n total {sum(J.shape[0] for J in Js)}, in {time.time() - t0:.1f}s")
    bp = BatchProblem(Js, torch.float32)
    out = BatchSweep(random_seed=42, num_alphas=128, pred_mode="secant").run(bp)
    log(f"C3 batched sweep done in {time.time() - t0:.1f}s")
    conv = np.mean([r["eigsh_converged"] for res in out.results for r in res.rows])
    assert conv >= 0.99
    sizes = np.array([J.shape[0] for J in Js])
    sample = [int(np.argmin(sizes)), int(np.argmax(sizes)), 17, 128]
    for b in sample:
        J, res = Js[b], out.results[b]
        Jn, d, _ = O.normalize(J, np.float64)
        for r in res.rows[:: 32]:
            s, _, _ = O.alpha_scaling(d, r["alpha"])
            lam, _ = O.leading_eigpair(Jn, s)
            assert abs(r["lambda1"] - lam) <= 2e-5 * max(1.0, abs(lam)), (b, r["alpha"])
        Jint = O.integer_couplings(J)
        assert res.best_energy == O.energy_int64(Jint, res.best_spins.astype(np.int64))
    del bp, out
    sub = [Js[i] for i in np.argsort(sizes)[:8]]
    bp = BatchProblem(sub, torch.float32)
    out = BatchSweep(random_seed=42, num_alphas=32, pred_mode="secant").run(bp)
    Jp, row_off, sz = pack(sub)
    K = out.packed_candidates.shape[0]
    words = out.packed_candidates.cpu().numpy()
    cand = np.stack([O.round_spins(np.unpackbits(w.view(np.uint8), bitorder="little")[: Jp.shape[0]] - 0.5)
                     for w in words])
    ref = O.local_search(O.integer_couplings(Jp), cand, 42)
    np.testing.assert_array_equal(out.ls_state.spins[:, :K].cpu().numpy().T, ref.spins)
    assert out.ls_state.passes == ref.passes


def test_c5_ba25_properties(gpu, monkeypatch):
    """Barabasi-Albert n = 2^25, m = 10, J = -1, tangent predictor, the first 8 alphas of the schedule:
    every alpha converged, a true residual <= 1e-5 on the host, exact energies of candidates."""
    t0 = time.time()
    J = graphs.barabasi_albert_native(1 << 25, 10, seed=42, kind="neg")
    log(f"C5 graph n={J.shape[0]} nnz={J.nnz} in {time.time() - t0:.1f}s")
    cap = _Capture(monkeypatch, (7,))
    p32 = DeviceProblem(J, torch.float32)
    log(f"C5 uploaded in {time.time() - t0:.1f}s")
    p32.autotune()
    log(f"C5 layout {p32.tuning['choice']} at {time.time() - t0:.1f}s")
    out = SpectralAnnealing("c5", "synthetic", random_seed=42, pred_mode="tangent", num_alphas=8,
                            schedule_interval=(0.0, 7 / 127), verbose=False).run_on_device(p32)
    log(f"C5 8 alphas: {out.sweep_time:.2f}s, {out.matvecs} applications, local search {out.post_sweep_time:.2f}s")
    assert all(r["eigsh_converged"] for r in out.rows)
    del p32
    r = out.rows[-1]
    Jn32, d32, _ = O.normalize(J, np.float32)
    res = _true_residual(Jn32, d32, r["alpha"], r["lambda1"], cap.vectors[7])
    log(f"alpha {r['alpha']:.4f}: true residual {res:.3e}")
    assert res <= 1e-5
    del Jn32
    _check_candidate_energies(J, out, (0, out.num_candidates - 1))
