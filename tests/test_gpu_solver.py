"""Solver-level parity of the MI355X SpecAnn against the CPU oracle (GPU)."""
import json
import os
from pathlib import Path

import numpy as np
import pandas as pd
import pytest
import scipy.sparse as sp
import torch

from oracle import specann_oracle as O
from specann_amd import graphs
from specann_amd.device import DeviceProblem
from specann_amd.endpoints import Alpha0GPU, Alpha1GPU
from specann_amd.spectral_annealing import SpectralAnnealing, unpack_bits

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def c1_graph():
    return graphs.erdos_renyi(1024, 40 / 1023, seed=42, kind="neg")


def sweep(J, **kw):
    kw.setdefault("random_seed", 42)
    kw.setdefault("verbose", False)
    prec = kw.get("precision", 32)
    prob = DeviceProblem(J, torch.float32 if prec == 32 and kw.get("fp32_spmv", True) is not False
                         else torch.float64)
    return prob, SpectralAnnealing("t", "synthetic", **kw).run_on_device(prob)


def check_rows_against_dense(J, out, dtype, tol_lambda):
    Jn, d, c = O.normalize(J, dtype)
    for r in out.rows:
        s, dsum, g = O.alpha_scaling(d, r["alpha"])
        lam, _ = O.leading_eigpair(Jn, s)
        if r["eigsh_converged"]:
            assert abs(r["lambda1"] - lam) <= tol_lambda, (r["alpha"], r["lambda1"], lam)
        assert r["degree_scale"] == pytest.approx(dsum, rel=1e-5)
        assert r["lambda1"] <= g * (1 + 1e-5)
        if r["eigsh_converged"]:
            assert r["energy_bound"] <= r["energy"] + 1e-6


def check_candidates_exact(J, out):
    Jint = O.integer_couplings(J)
    bits = out.cand_bits.cpu().numpy().view(np.uint64)
    n = J.shape[0]
    cands = np.stack([unpack_bits(b.view(np.int64), n) for b in bits])
    raw = [O.energy_int64(Jint, c) for c in cands]
    assert raw == [int(x) for x in out.raw_energies]
    return Jint, cands


def test_c1_full_sweep_fp32():
    J = c1_graph()
    print("C1 full sweep: ER n=1024 p=40/1023 J=-1 seed 42, 128 alphas", flush=True)
    prob, out = sweep(J)
    assert len(out.rows) == 128
    check_rows_against_dense(J, out, np.float32, 2e-5)
    Jint, cands = check_candidates_exact(J, out)
    # the post-sweep local search equals the oracle's on the same candidates, bit for bit
    ref = O.local_search(Jint, cands, 42)
    assert [int(-e // 2) for e in ref.e2_final] == [int(x) for x in out.refined_energies]
    assert out.local_passes == ref.passes
    assert out.best_energy == min(min(out.refined_energies), min(out.raw_energies))
    assert out.best_energy == O.energy_int64(Jint, out.best_spins.astype(np.int64))
    # dedup: one candidate per distinct rounding, in first-appearance order
    assert sorted(set(r["candidate"] for r in out.rows)) == list(range(out.num_candidates))
    first = [next(i for i, r in enumerate(out.rows) if r["candidate"] == c) for c in range(out.num_candidates)]
    assert first == sorted(first)
    # the CPU path (ARPACK + same local search) lands within 1% of the same best energy
    cpu = O.cpu_sweep(J, seed=42)
    print(f"best energy GPU {out.best_energy} CPU {cpu.best_energy}; {out.matvecs} operator applications")
    assert abs(out.best_energy - cpu.best_energy) <= 0.01 * abs(cpu.best_energy)


def test_rounded_spins_match_fp64_eigenvector():
    """Rounding parity: the GPU candidate of each alpha equals sign(x) of the FP64 oracle eigenvector
    except where |y_i| is below the eigen tolerance scale."""
    J = graphs.barabasi_albert(3000, 5, seed=7, kind="pm1")
    prob, out = sweep(J, num_alphas=12, use_local_search=True)
    Jn, d, c = O.normalize(J, np.float32)
    bits = out.cand_bits.cpu().numpy()
    for r in out.rows:
        s, _, _ = O.alpha_scaling(d, r["alpha"])
        _, y = O.leading_eigpair(Jn, s)
        got = unpack_bits(bits[r["candidate"]], J.shape[0]).astype(np.int64)
        ref = O.round_spins(y).astype(np.int64)
        if np.sum(got * ref) < 0:
            ref = -ref
        firm = np.abs(y) > 1e-3
        assert np.array_equal(got[firm], ref[firm]), r["alpha"]


def test_fp64_form_tight_residuals():
    J = graphs.barabasi_albert(4000, 6, seed=3, kind="pm1")
    prob, out = sweep(J, precision=64, num_alphas=10, eigsh_tol=1e-10)
    Jn, d, c = O.normalize(J, np.float64)
    for r in out.rows:
        assert r["eigsh_converged"]
        s, _, _ = O.alpha_scaling(d, r["alpha"])
        lam, _ = O.leading_eigpair(Jn, s)
        assert abs(r["lambda1"] - lam) <= 1e-9 * max(1.0, abs(lam))
    check_candidates_exact(J, out)


@pytest.mark.parametrize("opts", [dict(pred_mode="secant"), dict(pred_mode="prev"), dict(disable_warm_start=True),
                                  dict(pred_mode="tangent"), dict(pred_mode="tangent", tangent_steps=3),
                                  dict(use_local_search=False), dict(schedule="linear", num_alphas=9),
                                  dict(ls_batch=False, num_alphas=6), dict(eigsh_ncv=8),
                                  dict(schedule_interval=(0, 1.25), num_alphas=7)])
def test_sweep_variants(opts):
    J = graphs.erdos_renyi(600, 12 / 599, seed=5, kind="pm1")
    kw = dict(num_alphas=16, verbose=True)
    kw.update(opts)
    print(f"variant {opts}", flush=True)
    prob, out = sweep(J, **kw)
    for r in out.rows:
        print(r["alpha_index"], r["eigsh_iterations"], r["eigsh_retry"], r["pred_residual"], r["lambda1"])
    assert len(out.rows) == kw["num_alphas"]
    check_rows_against_dense(J, out, np.float32, 5e-5)
    check_candidates_exact(J, out)
    Jint = O.integer_couplings(J)
    assert out.best_energy == O.energy_int64(Jint, out.best_spins.astype(np.int64))
    if opts.get("pred_mode") == "prev" or opts.get("disable_warm_start"):
        assert all(np.isnan(r["pred_residual"]) for r in out.rows)
    else:
        # the predictor gate ran at every warm-started alpha and its verdict matches its residual
        for r in out.rows[1:]:
            if r["eigsh_retry"]:
                continue
            assert np.isfinite(r["pred_residual"])
            assert r["pred_accepted"] == (not r["pred_residual"] > 1e-3)


def test_graph_replay_equals_eager_launches():
    """Lanczos chunks replayed from a captured HIP graph give bit-identical sweeps to eager launches
    (the graph is replayed many times, with eager work and host synchronisations in between)."""
    J = graphs.erdos_renyi(900, 16 / 899, seed=8, kind="pm1")
    outs = []
    for g in (True, False):
        prob = DeviceProblem(J, torch.float32)
        outs.append(SpectralAnnealing("i", "d", random_seed=5, num_alphas=12, verbose=False,
                                      lanczos_graphs=g).run_on_device(prob))
    a, b = outs
    for r1, r2 in zip(a.rows, b.rows):
        assert r1["eigsh_iterations"] == r2["eigsh_iterations"]
        assert r1["lambda1"] == r2["lambda1"]
        assert r1["candidate"] == r2["candidate"]
    assert a.best_energy == b.best_energy
    assert a.raw_energies == b.raw_energies and a.refined_energies == b.refined_energies


def test_tiny_and_edge_graphs():
    # two spins, one coupling: ncv = n = 2
    J = sp.csr_matrix(np.array([[0.0, -1.0], [-1.0, 0.0]]))
    prob, out = sweep(J, num_alphas=4)
    assert out.best_energy == -1.0
    # a star (one long row) and a graph with an isolated node
    n = 5000
    rows = np.zeros(n - 2, dtype=np.int64)
    cols = np.arange(1, n - 1, dtype=np.int64)
    Jstar = graphs.symmetric_csr(n, rows, cols, -np.ones(n - 2))
    prob, out = sweep(Jstar, num_alphas=6)
    Jint = O.integer_couplings(Jstar)
    assert out.best_energy == O.energy_int64(Jint, out.best_spins.astype(np.int64))
    assert out.best_energy == -(n - 2)     # the star is bipartite: every edge cut


def test_endpoints_match_dense():
    J = graphs.erdos_renyi(800, 16 / 799, seed=9, kind="pm1")
    Jn, d, c = O.normalize(J, np.float64)
    Jint = O.integer_couplings(J)
    for cls, alpha, clamp in ((Alpha0GPU, 0.0, 1e-12), (Alpha1GPU, 1.0, 1e-7)):
        solver = cls("t.txt", "synthetic", random_seed=3, verbose=False)
        e, spins, _ = solver.solve(J)
        s, _, _ = O.alpha_scaling(d, alpha, clamp)
        lam, _ = O.leading_eigpair(Jn, s)
        assert abs(solver._mu - (1.0 - lam)) <= 2e-5
        assert e == O.energy_int64(Jint, spins.astype(np.int64))
        assert solver._eigsh_converged


def test_cli_on_bundled_instances(tmp_path, monkeypatch):
    from main import main
    monkeypatch.chdir(tmp_path)
    inst = str(GOLDEN / "instances")
    rc = main(["--solver", "SpecAnn", "--dataset", "barabasi_albert_3ins_m10_tiny", "--datasets-path", inst,
               "--seed", "42", "--save-spins"])
    assert rc == 0
    ground = json.loads((GOLDEN / "ground_states.json").read_text())
    df = pd.read_csv(tmp_path / "results/barabasi_albert_3ins_m10_tiny/SpecAnn/ba_20_1_0.csv")
    cols = ['instance_name', 'dataset', 'seed', 'solver_name', 'energy', 'time', 'local_search',
            'local_search_iterations', 'alpha', 'beta', 'mode', 'precision', 'fp32_spmv', 'ls_policy',
            'alphas_skipped', 'alphas_retried', 'num_candidates', 'sweep_time', 'post_sweep_time',
            'gpu_pool_bytes', 'timestamp']
    assert cols == list(df.columns[:len(cols)])
    e = float(df.energy.iloc[0])
    assert e >= ground["ba_20_1_0.txt"]["ground_energy"]
    assert e <= 0.97 * ground["ba_20_1_0.txt"]["ground_energy"]     # within 3% of the optimum
    alphas = pd.read_csv(tmp_path / "results/barabasi_albert_3ins_m10_tiny/SpecAnn/alphas/ba_20_1_0_alphas.csv")
    assert len(alphas) == 128 and 'eigsh_iterations' in alphas.columns
    sols = list((tmp_path / "results/barabasi_albert_3ins_m10_tiny/SpecAnn/solutions").glob("*.sol"))
    assert len(sols) == 3
    for run in ("SpecAdj", "SpecSignedLap", "SpecAnn-64"):
        assert main(["--solver", run, "--dataset", "barabasi_albert_3ins_m10_tiny", "--datasets-path", inst,
                     "--instance", "ba_25_1_1.txt", "--seed", "1"]) == 0
    # the guarded policy's run name (configs/default.yml: ls_policy adaptive_guard)
    assert main(["--solver", "SpecAnn-guard-adaptive", "--dataset", "barabasi_albert_3ins_m10_tiny",
                 "--datasets-path", inst, "--instance", "ba_25_1_1.txt", "--seed", "1"]) == 0
    g = pd.read_csv(tmp_path / "results/barabasi_albert_3ins_m10_tiny/SpecAnn-guard-adaptive/ba_25_1_1.csv")
    assert g.ls_policy.iloc[0] == "adaptive_guard"
    assert float(g.energy.iloc[0]) < 0 and int(g.local_search_iterations.iloc[0]) >= 1


def test_breakdown_detection_from_the_host(gpu):
    """The Ritz-matrix guard (an oversized recurrence coefficient) and a non-finite start vector stop
    the solve with SA_BROKEN; every kernel still completes normally."""
    from specann_amd._lib import SA_BROKEN
    from specann_amd.eigsh import LanczosSolver
    J = graphs.erdos_renyi(500, 10 / 499, seed=2, kind="pm1")
    prob = DeviceProblem(J, torch.float32)
    s = torch.empty(prob.n, dtype=torch.float32, device=gpu)
    prob.alpha_scaling(0.5, 1e-12, s)
    lz = LanczosSolver(prob, 20)
    x = torch.from_numpy(O.normal_vector(prob.n, 1, 0)).to(gpu, torch.float32)
    lz.begin(1e-6, 200, bound=1e-9)                 # |alpha_0| > 10 bound: the guard trips at step 0
    lz.start(s, x)
    assert lz.run(s, 200).status == SA_BROKEN
    bad = x.clone()
    bad[7] = float("nan")
    lz.begin(1e-6, 200, bound=10.0)
    lz.start(s, bad)                                 # non-finite norm: broken before the first step
    assert lz.run(s, 200).status == SA_BROKEN
    lz.begin(1e-6, 200, bound=10.0)                  # and the solver recovers on the next begin
    lz.start(s, x)
    assert lz.run(s, 200).status != SA_BROKEN


@pytest.mark.parametrize("fail_calls", [1, 2])
def test_breakdown_retry_and_skip(gpu, monkeypatch, fail_calls):
    """A breakdown at one alpha is retried once from a random vector (the row says eigsh_retry);
    when the retry breaks down too, the alpha is skipped and counted (models/spectral_annealing.py:616-647,
    :323-329)."""
    from specann_amd import eigsh as E
    J = graphs.erdos_renyi(600, 12 / 599, seed=5, kind="pm1")
    orig = E.LanczosSolver.begin
    calls = {"n": 0}

    def begin(self, tol, maxiter, bound, pred_tol=float("inf")):
        calls["n"] += 1
        # the third alpha's begin calls: trip the breakdown guard with a tiny bound
        if 3 <= calls["n"] < 3 + fail_calls:
            bound = 1e-12
        return orig(self, tol, maxiter, bound, pred_tol)

    monkeypatch.setattr(E.LanczosSolver, "begin", begin)
    sa = SpectralAnnealing("t", "synthetic", random_seed=42, verbose=False, num_alphas=8, pred_mode="prev")
    prob = DeviceProblem(J, torch.float32)
    out = sa.run_on_device(prob)
    if fail_calls == 1:
        assert len(out.rows) == 8
        assert out.rows[2]["eigsh_retry"] and not any(r["eigsh_retry"] for i, r in enumerate(out.rows) if i != 2)
        assert sa._alphas_skipped == 0
    else:
        assert len(out.rows) == 7 and sa._alphas_skipped == 1
        assert [r["alpha_index"] for r in out.rows] == [0, 1, 3, 4, 5, 6, 7]
    check_rows_against_dense(J, out, np.float32, 5e-5)
