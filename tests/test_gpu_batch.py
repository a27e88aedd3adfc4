"""The batched (block-diagonal) sweep against one-instance sweeps and the CPU oracle (GPU)."""
import numpy as np
import pytest
import torch

from oracle import specann_oracle as O
from specann_amd import graphs
from specann_amd.batch import BatchProblem, BatchSweep, pack
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
