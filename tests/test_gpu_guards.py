"""Guarded local-search policies on the GPU against the oracle (GPU): spins, reported iterations and
energies bit for bit (integer couplings) or to FP64 rounding of the tracked energies (FP32 couplings
J / c), over one and several 64-candidate groups."""
import numpy as np
import pytest
import torch

from oracle import specann_oracle as O
from specann_amd import graphs
from specann_amd import guards as G
from specann_amd.device import DeviceProblem

pytestmark = pytest.mark.gpu

GRAPHS = {
    "er_500_pm1": lambda: graphs.erdos_renyi(500, 12 / 499, seed=3, kind="pm1"),
    "ba_1500_neg": lambda: graphs.barabasi_albert(1500, 3, seed=4, kind="neg"),
}


def _cands(n, k, seed=7):
    return np.stack([O.round_spins(O.normal_vector(n, seed, i)) for i in range(k)])


def _bits(cands, dev):
    return torch.from_numpy(np.stack([O.spin_bits(c) for c in cands]).view(np.int64)).to(dev)


def _float_graph(seed=2):
    rng = np.random.default_rng(seed)
    J = graphs.erdos_renyi(400, 10 / 399, seed=seed, kind="pm1").astype(np.float64)
    J.data = J.data * rng.choice([0.5, 1.5, 2.25], J.nnz)
    return ((J + J.T) * 0.5).tocsr()


def _oracle(policy, J, cands, seed, scale, max_iterations):
    if policy in ("fixed_guard", "adaptive_guard"):
        return O.guard_safe(J, cands, seed, adaptive=policy == "adaptive_guard", scale=scale,
                            max_iterations=max_iterations)
    if policy == "independent_set":
        return O.guard_independent_set(J, cands, seed, scale=scale, max_updates=max_iterations + 625)
    return O.guard_sequential(J, cands, scale=scale)


@pytest.mark.parametrize("policy", list(G.POLICIES))
@pytest.mark.parametrize("case,k", [("er_500_pm1", 3), ("ba_1500_neg", 70)])
def test_policy_matches_oracle_integer(gpu, policy, case, k):
    J = GRAPHS[case]()
    print(f"guard policy={policy} case={case} k={k}", flush=True)
    prob = DeviceProblem(J, torch.float32)
    cands = _cands(prob.n, k)
    st = G.refine(policy, prob, _bits(cands, gpu), k, max_iterations=200, seed=17)
    ref = _oracle(policy, O.integer_couplings(J), cands, 17, prob.c, 200)
    np.testing.assert_array_equal(st.spins[:, :k].cpu().numpy().T, ref.spins)
    assert st.iterations == ref.iterations
    np.testing.assert_array_equal(st.energies_final, ref.energies)
    np.testing.assert_array_equal(st.energies_start, ref.energies_start)
    assert np.all(st.energies_final <= st.energies_start)
    for c in range(min(k, 4)):
        assert st.energies_final[c] == O.energy_int64(O.integer_couplings(J), ref.spins[c])


@pytest.mark.parametrize("policy", list(G.POLICIES))
def test_policy_matches_oracle_float(gpu, policy):
    J = _float_graph()
    print(f"guard policy={policy} float", flush=True)
    prob = DeviceProblem(J, torch.float32)
    assert prob.jint is None
    Jn, _, c = O.normalize(J, np.float32)
    k = 5
    cands = _cands(prob.n, k, seed=3)
    st = G.refine(policy, prob, _bits(cands, gpu), k, max_iterations=150, seed=5)
    ref = _oracle(policy, Jn, cands, 5, 1.0, 150)
    np.testing.assert_array_equal(st.spins[:, :k].cpu().numpy().T, ref.spins)
    assert st.iterations == ref.iterations
    np.testing.assert_allclose(st.energies_final, ref.energies * c, rtol=1e-10)


def test_run_policy_reference_signature(gpu):
    """policies.run_policy(policy, J_csr, spins_list, max_iterations, dtype, seed) -> (n x k int8, iters,
    summary), as SpectralAnnealing._refine_spins calls it (models/spectral_annealing.py:543-547)."""
    J = GRAPHS["er_500_pm1"]()
    cands = _cands(J.shape[0], 4)
    spins_list = [torch.from_numpy(c.astype(np.float32)).to(gpu) for c in cands]
    refined, iters, summary = G.run_policy("adaptive_guard", J, spins_list, max_iterations=100,
                                           dtype=np.float32, seed=9)
    assert refined.shape == (J.shape[0], 4) and refined.dtype == torch.int8
    ref = O.guard_safe(O.integer_couplings(J), cands, 9, adaptive=True,
                       scale=float(np.abs(J).sum(axis=1).max()), max_iterations=100)
    np.testing.assert_array_equal(refined.cpu().numpy().T, ref.spins)
    assert iters == ref.iterations == summary["main_iterations"]
    out, it2 = G.refine_candidates("sequential_oracle", J, spins_list[:1], max_iterations=100, seed=0)
    assert out.shape == (J.shape[0], 1) and it2 >= 1


@pytest.mark.parametrize("policy", list(G.POLICIES))
def test_sweep_with_guard_policy(gpu, policy):
    """SpectralAnnealing with ls_policy: the distinct candidates of the sweep go through the policy
    (models/spectral_annealing.py:423-456, :540-547); refined energies and the reported iterations
    equal the oracle's on the same candidates, and every row keeps the better of raw and refined."""
    from specann_amd.spectral_annealing import SpectralAnnealing, unpack_bits
    J = graphs.erdos_renyi(600, 20 / 599, seed=8, kind="neg")
    print(f"sweep policy={policy}", flush=True)
    prob = DeviceProblem(J, torch.float32)
    out = SpectralAnnealing("t", "synthetic", random_seed=42, num_alphas=16, ls_policy=policy,
                            max_local_iterations=100, verbose=False).run_on_device(prob)
    bits = out.cand_bits.cpu().numpy()
    cands = np.stack([unpack_bits(b, prob.n) for b in bits])
    ref = _oracle(policy, O.integer_couplings(J), cands, 42, prob.c, 100)
    assert out.refined_energies == [int(e) for e in ref.energies]
    assert out.raw_energies == [int(e) for e in ref.energies_start]
    assert all(r["local_search_iterations"] == ref.iterations for r in out.rows if "local_search_iterations" in r)
    assert out.best_energy == min(min(out.refined_energies), min(out.raw_energies))
    assert O.energy_int64(O.integer_couplings(J), out.best_spins.astype(np.int64)) == out.best_energy
