"""The batched (block-diagonal) sweep against one-instance sweeps and the CPU oracle (GPU)."""
import numpy as np
import scipy.sparse as sp
import pytest
import torch

from oracle import specann_oracle as O
from specann_amd import graphs
from specann_amd.batch import TERMINAL, BatchProblem, BatchSweep, pack
from specann_amd.spectral_annealing import SpectralAnnealing
from specann_amd.device import DeviceProblem

pytestmark = pytest.mark.gpu


def _instances():
    return [graphs.erdos_renyi(300, 12 / 299, seed=1, kind="neg"),
            graphs.barabasi_albert(517, 4, seed=2, kind="neg"),
            graphs.erdos_renyi(1000, 20 / 999, seed=3, kind="pm1"),
            graphs.barabasi_albert(130, 3, seed=4, kind="pm1"),
            graphs.erdos_renyi(2200, 10 / 2199, seed=5, kind="neg")]


@pytest.mark.parametrize("tdtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("pred", ["diagonal", "secant"])
def test_batched_sweep_matches_single_instance_sweeps(gpu, tdtype, pred):
    Js = _instances()
    print(f"batch dtype={tdtype} pred={pred} sizes={[J.shape[0] for J in Js]}", flush=True)
    bp = BatchProblem(Js, tdtype, rows_per_block=256)
    tol = 1e-6 if tdtype == torch.float32 else 1e-10
    out = BatchSweep(random_seed=7, num_alphas=10, pred_mode=pred, eigsh_tol=tol).run(bp)
    prec = 32 if tdtype == torch.float32 else 64
    for b, J in enumerate(Js):
        res = out.results[b]
        assert len(res.rows) == 10
        Jn, d, c = O.normalize(J, np.float64)
        for r in res.rows:
            assert r["eigsh_converged"], (b, r)
            s, _, g = O.alpha_scaling(d, r["alpha"])
            lam, _ = O.leading_eigpair(Jn, s)
            assert abs(r["lambda1"] - lam) <= (2e-5 if prec == 32 else 1e-8) * max(1.0, abs(lam)), (b, r["alpha"])
            assert r["gershgorin"] == pytest.approx(g, rel=1e-5)
        # the same instance swept alone gives the same eigenvalues
        single = SpectralAnnealing("i", "d", random_seed=7, precision=prec, num_alphas=10, pred_mode=pred,
                                   eigsh_tol=tol, verbose=False).run_on_device(DeviceProblem(J, tdtype))
        for r1, r2 in zip(res.rows, single.rows):
            assert r1["lambda1"] == pytest.approx(r2["lambda1"], rel=2e-5 if prec == 32 else 1e-8)
        # energies are exact and the best spins realise the best energy
        Jint = O.integer_couplings(J)
        assert res.best_spins is not None and res.best_spins.shape == (J.shape[0],)
        assert res.best_energy == O.energy_int64(Jint, res.best_spins.astype(np.int64))
        assert res.best_energy <= min(res.raw_energies)
        assert min(res.refined_energies) >= res.best_energy


def test_batched_local_search_equals_oracle_on_packed_graph(gpu):
    Js = _instances()[:3]
    bp = BatchProblem(Js, torch.float32, rows_per_block=256)
    out = BatchSweep(random_seed=3, num_alphas=8).run(bp)
    Jp, row_off, sizes = pack(Js)
    Jint = O.integer_couplings(Jp)
    K = out.packed_candidates.shape[0]
    words = out.packed_candidates.cpu().numpy()
    cand = np.stack([O.round_spins(np.unpackbits(w.view(np.uint8), bitorder="little")[: Jp.shape[0]] - 0.5)
                     for w in words])                                  # K x n
    ref = O.local_search(Jint, cand, 3, max_iterations=1000)
    state = out.ls_state
    got = state.spins[:, :K].cpu().numpy().T                             # K x n
    np.testing.assert_array_equal(got, ref.spins)
    assert state.passes == ref.passes
    spins = got.T.astype(np.int64)
    cand = cand.T
    for b, J in enumerate(Js):
        rows = slice(int(row_off[b]), int(row_off[b] + sizes[b]))
        Jb = O.integer_couplings(J)
        for c in range(out.results[b].num_candidates):
            assert out.results[b].refined_energies[c] == O.energy_int64(Jb, spins[rows, c])
            assert out.results[b].raw_energies[c] == O.energy_int64(Jb, cand[rows, c].astype(np.int64))


def test_batch_energy_entry_points_agree(gpu):
    """Per-instance exact energies of random candidate bits: the candidate-bit kernel
    (sa_batch_energy_bits_i64) and the local search's initial fields summed per instance
    (sa_batch_ls_energy, which the batched sweep uses) agree with each other and the oracle."""
    from specann_amd import local_search as LS
    from specann_amd._lib import call
    from specann_amd.device import ptr, stream
    Js = _instances()
    bp = BatchProblem(Js, torch.float32, rows_per_block=256)
    prob = bp.prob
    nwords = (prob.n + 63) // 64
    K = 70
    words = np.random.default_rng(5).integers(0, 1 << 62, size=(K, nwords), dtype=np.int64)
    packed = torch.from_numpy(words).to(gpu)
    e2a = torch.zeros((len(Js), K), dtype=torch.int64, device=gpu)
    call("sa_batch_energy_bits_i64", prob.csr_ref, bp.ref, K, ptr(packed), nwords, ptr(e2a), stream())
    init = LS.init_fields(prob, packed, K)
    e2b = torch.zeros_like(e2a)
    call("sa_batch_ls_energy", bp.ref, K, init[2], ptr(init[0]), ptr(init[1]), ptr(e2b), stream())
    assert torch.equal(e2a, e2b)
    Jp, row_off, sizes = pack(Js)
    bits = np.unpackbits(words.view(np.uint8), bitorder="little").reshape(K, -1)[:, : Jp.shape[0]]
    got = -(e2a.cpu().numpy() // 2)
    for b, J in enumerate(Js):
        rows = slice(int(row_off[b]), int(row_off[b] + sizes[b]))
        Jb = O.integer_couplings(J)
        for c in range(0, K, 7):
            assert got[b, c] == O.energy_int64(Jb, bits[c, rows].astype(np.int64) * 2 - 1)


def test_batch_edge_cases(gpu):
    """One instance; instances smaller than a 64-row word and than the Lanczos basis."""
    tiny = graphs.erdos_renyi(12, 0.5, seed=9, kind="neg")
    star = graphs.barabasi_albert(40, 1, seed=3, kind="pm1")
    for Js in ([graphs.erdos_renyi(700, 10 / 699, seed=6, kind="pm1")], [tiny, star, tiny]):
        bp = BatchProblem(Js, torch.float64, rows_per_block=64)
        out = BatchSweep(random_seed=1, num_alphas=6, eigsh_tol=1e-9).run(bp)
        for b, J in enumerate(Js):
            res = out.results[b]
            Jn, d, c = O.normalize(J, np.float64)
            for r in res.rows:
                assert r["eigsh_converged"], (b, r)
                s, _, _ = O.alpha_scaling(d, r["alpha"])
                lam, _ = O.leading_eigpair(Jn, s)
                assert abs(r["lambda1"] - lam) <= 1e-7 * max(1.0, abs(lam))
            Jint = O.integer_couplings(J)
            assert res.best_energy == O.energy_int64(Jint, res.best_spins.astype(np.int64))
            if J.shape[0] <= 16:   # brute force the ground state: the sweep's best is never below it
                best = min(O.energy_int64(Jint, np.array([1 if (m >> i) & 1 else -1 for i in range(J.shape[0])]))
                           for m in range(1 << J.shape[0]))
                assert res.best_energy >= best


def _rows_equal(r1, r2):
    keys = ("alpha_index", "eigsh_iterations", "eigsh_converged", "lambda1", "rho", "gershgorin", "energy_bound",
            "residual_estimate", "lanczos_restarts", "failed", "candidate", "energy", "raw_energy")
    for k in keys:
        a, b = r1.get(k), r2.get(k)
        if isinstance(a, float) and isinstance(b, float) and np.isnan(a) and np.isnan(b):
            continue
        assert a == b, (k, a, b)
    a, b = r1["pred_residual"], r2["pred_residual"]
    assert (np.isnan(a) and np.isnan(b)) or a == b, ("pred_residual", a, b)


@pytest.mark.parametrize("tdtype,pred", [(torch.float32, "diagonal"), (torch.float32, "secant"),
                                         (torch.float32, "prev"), (torch.float64, "secant")])
def test_async_schedule_gives_the_lockstep_bits(gpu, tdtype, pred):
    """Each instance's arithmetic does not depend on when the others move on: moving every instance to
    its next alpha as soon as its own solve stops gives the lockstep schedule's eigenvalues, step
    counts, candidates and energies bit for bit."""
    Js = _instances()
    print(f"async vs lockstep dtype={tdtype} pred={pred}", flush=True)
    bp = BatchProblem(Js, tdtype, rows_per_block=256)
    tol = 1e-6 if tdtype == torch.float32 else 1e-10
    kw = dict(random_seed=11, num_alphas=12, pred_mode=pred, eigsh_tol=tol)
    sweep = BatchSweep(**kw)
    fast = sweep.run(bp)
    slow = BatchSweep(lockstep=True, **kw).run(bp)
    assert sweep.retries == 0 and sweep.fallbacks == 0
    assert fast.matvecs == slow.matvecs
    for b in range(len(Js)):
        f, s = fast.results[b], slow.results[b]
        assert len(f.rows) == len(s.rows) == 12
        for r1, r2 in zip(f.rows, s.rows):
            _rows_equal(r1, r2)
        assert f.raw_energies == s.raw_energies
        assert f.refined_energies == s.refined_energies
        assert f.best_energy == s.best_energy and f.best_alpha == s.best_alpha
        assert np.array_equal(f.best_spins, s.best_spins)


def test_async_schedule_retries_a_rejected_solve_cold(gpu, monkeypatch):
    """A solve the acceptance test refuses is solved again from a random vector at the same alpha (the
    one-instance sweep's retry); the retried solve converges to the same eigenvalue and the other
    instances are untouched."""
    Js = _instances()[:3]
    bp = BatchProblem(Js, torch.float32, rows_per_block=256)
    kw = dict(random_seed=5, num_alphas=6, pred_mode="secant")
    ref = BatchSweep(lockstep=True, **kw).run(bp)
    seen = []
    orig = BatchSweep._ok

    def ok(self, st, theta, g, ai):
        res = orig(self, st, theta, g, ai)
        if ai[1] == 3 and st[1] in TERMINAL and not seen and res[1]:   # refuse instance 1 at alpha 3, once
            res = res.copy()
            res[1] = False
            seen.append(1)
        return res

    monkeypatch.setattr(BatchSweep, "_ok", ok)
    sweep = BatchSweep(**kw)
    out = sweep.run(bp)
    assert seen == [1] and sweep.retries == 1
    row = out.results[1].rows[3]
    assert row["retry"] and not row["failed"] and row["eigsh_converged"]
    assert row["lambda1"] == pytest.approx(ref.results[1].rows[3]["lambda1"], rel=2e-5)
    assert sum(r["retry"] for r in out.results[1].rows) == 1
    for b in (0, 2):
        for r1, r2 in zip(out.results[b].rows, ref.results[b].rows):
            _rows_equal(r1, r2)


def test_masked_batch_entry_points(gpu):
    """A mask limits the per-instance launches to its instances; the per-instance alpha and predictor
    exponent forms give the scalar forms' bits; the step refuses a mask."""
    import ctypes
    from specann_amd import _lib
    from specann_amd._lib import SaBatch, call
    from specann_amd.device import ptr, stream
    Js = _instances()
    bp = BatchProblem(Js, torch.float32, rows_per_block=256)
    prob, B, n, dev = bp.prob, bp.B, bp.n, bp.device
    mask = np.array([1, 0, 1, 0, 1], dtype=np.uint8)
    mdev = torch.from_numpy(mask).to(dev)
    desc = SaBatch.from_buffer_copy(bp.desc)
    desc.inst_mask = mdev.data_ptr()
    M = ctypes.byref(desc)
    alphas = np.array([0.3, 0.6, 0.45, 0.9, 0.7])
    s_ref = torch.full((n,), 7.0, device=dev)
    s_got = torch.full((n,), 7.0, device=dev)
    for b in range(B):                       # scalar form, one alpha at a time, kept on instance b's rows
        tmp = torch.empty(n, device=dev)
        call("sa_batch_alpha_scaling_f32", bp.ref, ptr(prob.d), float(alphas[b]), 1e-12, ptr(tmp), stream())
        if mask[b]:
            s_ref[bp.rows_of(b)] = tmp[bp.rows_of(b)]
    a_dev = torch.from_numpy(alphas).to(dev)
    call("sa_batch_alpha_scaling_each_f32", M, ptr(prob.d), ptr(a_dev), 1e-12, ptr(s_got), stream())
    assert torch.equal(s_got, s_ref)
    x = torch.randn(n, device=dev)
    y = torch.full((n,), -1.0, device=dev)
    call("sa_batch_copy_f32", M, ptr(x), ptr(y), stream())
    for b in range(B):
        r = bp.rows_of(b)
        assert torch.equal(y[r], x[r] if mask[b] else torch.full_like(x[r], -1.0))
    # per-instance predictor exponent against the scalar one
    ldv = _lib.padded_ld(n, 4)
    V1 = torch.zeros((2, ldv), device=dev)
    V2 = torch.zeros((2, ldv), device=dev)
    Z1, Z2 = torch.zeros(n, device=dev), torch.zeros(n, device=dev)
    bounds = torch.full((B,), 10.0, dtype=torch.float64, device=dev)
    call("sa_batch_lanczos_begin", bp.ref, 8, 100, 1e-6, ptr(bounds), 1e-3, 2, 0.0, stream())
    pexp = np.array([0.05, 0.0, 0.11, 0.3, 0.02])
    for b in range(B):
        if mask[b]:
            d1 = SaBatch.from_buffer_copy(bp.desc)
            one = torch.from_numpy((np.arange(B) == b).astype(np.uint8)).to(dev)
            d1.inst_mask = one.data_ptr()
            call("sa_batch_lanczos_start_f32", ctypes.byref(d1), ptr(s_got), ptr(x), ptr(prob.d), float(pexp[b]),
                 1.0, None, 0.0, ptr(V1), ldv, ptr(Z1), 1, -1, stream())
    p_dev = torch.from_numpy(pexp).to(dev)
    call("sa_batch_lanczos_start_each_f32", M, ptr(s_got), ptr(x), ptr(prob.d), ptr(p_dev), ptr(V2), ldv, ptr(Z2), 1,
         stream())
    assert torch.equal(V1, V2) and torch.equal(Z1, Z2)
    heads = bp.heads()
    assert np.all(heads["status"][mask == 0] == _lib.SA_RUNNING)
    assert np.all(heads["r"][mask == 0, 3] == 0.0) and np.all(heads["r"][mask == 1, 3] > 0.0)  # start norms
    W = torch.zeros(n, device=dev)
    with pytest.raises(RuntimeError, match="mask"):
        call("sa_batch_lanczos_step_f32", prob.csr_ref, M, ptr(s_got), ptr(V2), ldv, ptr(Z2), ptr(W), stream())


@pytest.mark.parametrize("tdtype,mixed", [(torch.float32, False), (torch.float64, False), (torch.float32, True)])
def test_instance_staged_spmv_matches_the_csr_product(gpu, tdtype, mixed):
    """The instance-staged layout (instance slices of the gathered vector staged in LDS, one or more
    passes) gives the fused product y = a x + b s (.) (J/c) g of an FP64 host reference, on instances of
    one to three passes, for several workgroup heights and lanes per row; padding rows get a x."""
    from specann_amd._lib import call
    from specann_amd.device import ptr, stream
    big = graphs.erdos_renyi(40000, 8 / 39999, seed=11, kind="pm1")
    Js = [big] + _instances()[:3]
    if mixed:   # one coupling of magnitude 3: entries keep their value stream (no sign packing)
        J = Js[2].tolil()
        i, j = sp.triu(J, 1).nonzero()
        J[i[0], j[0]] = J[j[0], i[0]] = 3
        Js[2] = J.tocsr()
    print(f"instance-staged spmv dtype={tdtype} sizes={[J.shape[0] for J in Js]}", flush=True)
    bp = BatchProblem(Js, tdtype, rows_per_block=256)
    prob, n = bp.prob, bp.n
    Jp, row_off, sizes = pack(Js)
    c = np.zeros(n)
    for b, J in enumerate(Js):
        c[row_off[b]:row_off[b] + sizes[b]] = abs(J).sum(axis=1).max()
    Jn = sp.diags(1.0 / np.where(c > 0, c, 1.0)) @ Jp
    rng = np.random.default_rng(3)
    s_h, g_h, x_h = rng.uniform(0.5, 1.5, n), rng.standard_normal(n), rng.standard_normal(n)
    dev = bp.device
    s, g, x = (torch.from_numpy(v).to(dev, tdtype) for v in (s_h, g_h, x_h))
    s64, g64, x64 = (v.double().cpu().numpy() for v in (s, g, x))
    want = 0.25 * x64 + -1.5 * s64 * (Jn @ g64)
    scale = 0.25 * np.abs(x64) + 1.5 * s64 * (abs(Jn) @ np.abs(g64))
    tol = 1e-6 if tdtype == torch.float32 else 1e-13
    sfx = "f32" if tdtype == torch.float32 else "f64"
    assert bp.inst_eligible()
    for rows, lanes in ((2048, 8), (1024, 4), (64, 32), (512, 16)):
        bp.set_inst_layout(rows, lanes)
        assert bp.layout["passes"] == (2 if tdtype == torch.float32 else 3)
        assert bp.layout["signed"] == (not mixed)
        y = torch.full((n,), 9.0, dtype=tdtype, device=dev)
        call(f"sa_spmv_gather_{sfx}", prob.csr_ref, ptr(s), ptr(g), 0.25, ptr(x), -1.5, ptr(y), stream())
        got = y.double().cpu().numpy()
        assert np.all(np.abs(got - want) <= tol * (scale + 1e-30) + 1e-300), (rows, lanes)
    bp.set_tiling(1024, 0)
    y = torch.full((n,), 9.0, dtype=tdtype, device=dev)
    call(f"sa_spmv_gather_{sfx}", prob.csr_ref, ptr(s), ptr(g), 0.25, ptr(x), -1.5, ptr(y), stream())
    assert np.all(np.abs(y.double().cpu().numpy() - want) <= tol * (scale + 1e-30) + 1e-300)


@pytest.mark.parametrize("tdtype", [torch.float32, torch.float64])
def test_batched_sweep_on_the_instance_staged_layout(gpu, tdtype):
    """The sweep on the instance-staged layout finds the plain-tile layout's eigenvalues at every alpha
    (the SpMV differs only in the order of its FP64 row sums)."""
    Js = _instances()
    tol = 1e-6 if tdtype == torch.float32 else 1e-10
    kw = dict(random_seed=7, num_alphas=8, pred_mode="secant", eigsh_tol=tol)
    bp = BatchProblem(Js, tdtype, rows_per_block=256)
    ref = BatchSweep(**kw).run(bp)
    bp.set_inst_layout(1024, 8)
    out = BatchSweep(**kw).run(bp)
    rel = 2e-5 if tdtype == torch.float32 else 1e-8
    for b, J in enumerate(Js):
        for r1, r2 in zip(out.results[b].rows, ref.results[b].rows):
            assert r1["eigsh_converged"] and r2["eigsh_converged"]
            assert r1["lambda1"] == pytest.approx(r2["lambda1"], rel=rel)
        Jint = O.integer_couplings(J)
        res = out.results[b]
        assert res.best_energy == O.energy_int64(Jint, res.best_spins.astype(np.int64))


@pytest.mark.parametrize("mixed", [False, True])
def test_instance_streamed_spmv_matches_the_csr_product(gpu, mixed):
    """The instance-streamed layout (FP32; each workgroup's rows of one instance streamed by column,
    FP64 row sums): the fused product y = a x + b s (.) (J/c) g of an FP64 host reference for several
    workgroup heights, padding rows y = a x, bitwise repeatable (the batched sweep's test runs it with
    tile lists of the instances still sweeping); sign-packed entries with per-instance magnitudes when
    every instance's couplings have one magnitude, values otherwise."""
    from specann_amd._lib import call
    from specann_amd.device import ptr, stream
    big = graphs.erdos_renyi(40000, 8 / 39999, seed=11, kind="pm1")
    Js = [big] + _instances()[:3]
    if mixed:
        J = Js[2].tolil()
        i, j = sp.triu(J, 1).nonzero()
        J[i[0], j[0]] = J[j[0], i[0]] = 3
        Js[2] = J.tocsr()
    bp = BatchProblem(Js, torch.float32, rows_per_block=256)
    prob, n = bp.prob, bp.n
    Jp, row_off, sizes = pack(Js)
    c = np.zeros(n)
    for b, J in enumerate(Js):
        c[row_off[b]:row_off[b] + sizes[b]] = abs(J).sum(axis=1).max()
    Jn = sp.diags(1.0 / np.where(c > 0, c, 1.0)) @ Jp
    rng = np.random.default_rng(3)
    s_h, g_h, x_h = rng.uniform(0.5, 1.5, n), rng.standard_normal(n), rng.standard_normal(n)
    dev = bp.device
    s, g, x = (torch.from_numpy(v).to(dev, torch.float32) for v in (s_h, g_h, x_h))
    s64, g64, x64 = (v.double().cpu().numpy() for v in (s, g, x))
    want = 0.25 * x64 + -1.5 * s64 * (Jn @ g64)
    scale = 0.25 * np.abs(x64) + 1.5 * s64 * (abs(Jn) @ np.abs(g64))
    assert bp.istream_eligible()
    for rows in (20416, 4096, 1024, 64):
        print(f"instance-streamed spmv rows={rows} sizes={[J.shape[0] for J in Js]}", flush=True)
        assert bp.set_istream_layout(rows)
        assert bp.layout["kind"] == "instance-streamed" and prob.csr.imode == 2
        # one magnitude per instance (J = +-1 or -1 each): sign-packed entries; the mixed batch keeps values
        assert bp.layout["signed"] == (not mixed) and prob.csr.isigned == int(not mixed)
        ys = []
        for _ in range(2):
            y = torch.full((n,), 9.0, dtype=torch.float32, device=dev)
            call("sa_spmv_gather_f32", prob.csr_ref, ptr(s), ptr(g), 0.25, ptr(x), -1.5, ptr(y), stream())
            ys.append(y.cpu().numpy())
        np.testing.assert_array_equal(ys[0], ys[1])
        got = ys[0].astype(np.float64)
        assert np.all(np.abs(got - want) <= 1e-6 * (scale + 1e-30) + 1e-300), rows


def test_batched_sweep_on_the_instance_streamed_layout(gpu):
    """The sweep on the instance-streamed layout (instances finish at different steps, so the SpMV
    runs on tile lists of the instances still sweeping) finds the plain-tile layout's eigenvalues at
    every alpha and exact energies of its spins."""
    Js = _instances()
    kw = dict(random_seed=7, num_alphas=8, pred_mode="secant", eigsh_tol=1e-6)
    bp = BatchProblem(Js, torch.float32, rows_per_block=256)
    ref = BatchSweep(**kw).run(bp)
    assert bp.set_istream_layout(4096)
    out = BatchSweep(**kw).run(bp)
    for b, J in enumerate(Js):
        for r1, r2 in zip(out.results[b].rows, ref.results[b].rows):
            assert r1["eigsh_converged"] and r2["eigsh_converged"]
            assert r1["lambda1"] == pytest.approx(r2["lambda1"], rel=2e-5)
        Jint = O.integer_couplings(J)
        res = out.results[b]
        assert res.best_energy == O.energy_int64(Jint, res.best_spins.astype(np.int64))
