"""The persistent local search (one cooperative launch for the whole schedule) against the
per-class launches and the oracle (GPU): spins, pass counts and exact energies, over schedules that
reach every branch of the device's schedule (main phase cut by max_iterations or by the active
candidates, final rounds ending at a fixed point, at a pass without flips on a subset, or at
final_sweep_iters; no final rounds; no main passes)."""
import numpy as np
import pytest
import torch

from oracle import specann_oracle as O
from specann_amd import graphs
from specann_amd import local_search as LS
from specann_amd.device import DeviceProblem

pytestmark = pytest.mark.gpu

SCHEDULES = {
    "default": dict(max_iterations=1000),
    "main_cut": dict(max_iterations=2, subset_fraction=0.5),
    "subset_final": dict(max_iterations=3, final_sweep_fraction=0.6, final_sweep_iters=4, num_rounds=5),
    "no_final": dict(max_iterations=50, num_rounds=0),
    "final_only": dict(max_iterations=0, final_sweep_iters=3, num_rounds=4),
    "full_main": dict(max_iterations=40, subset_fraction=1.0, final_sweep_iters=1, num_rounds=3),
}


def _cands(n, k, seed=9):
    return np.stack([O.round_spins(O.normal_vector(n, seed, i)) for i in range(k)])


def _bits(cands, dev):
    return torch.from_numpy(np.stack([O.spin_bits(c) for c in cands]).view(np.int64)).to(dev)


@pytest.mark.parametrize("name", list(SCHEDULES))
def test_persistent_equals_per_class_and_oracle(gpu, name):
    kw = SCHEDULES[name]
    J = graphs.erdos_renyi(3000, 30 / 2999, seed=4, kind="pm1")
    print(f"persistent ls schedule={name}", flush=True)
    prob = DeviceProblem(J, torch.float32)
    k = 5
    cands = _cands(prob.n, k)
    col = LS.colour_graph(prob, 77)
    a = LS.flip_refinement(prob, _bits(cands, gpu), k, seed=77, colouring=col, persistent=True, **kw)
    b = LS.flip_refinement(prob, _bits(cands, gpu), k, seed=77, colouring=col, persistent=False, **kw)
    np.testing.assert_array_equal(a.spins[:, :k].cpu().numpy(), b.spins[:, :k].cpu().numpy())
    assert (a.passes, a.total_passes) == (b.passes, b.total_passes)
    np.testing.assert_array_equal(a.e2_final, b.e2_final)
    ref = O.local_search(O.integer_couplings(J), cands, 77, **kw)
    np.testing.assert_array_equal(a.spins[:, :k].cpu().numpy().T, ref.spins)
    assert a.passes == ref.passes and a.total_passes == ref.total_passes
    np.testing.assert_array_equal(a.e2_final, ref.e2_final)


@pytest.mark.parametrize("groups_k", [64, 130])
def test_persistent_many_groups_and_hubs(gpu, groups_k):
    """Several 64-candidate groups and a heavy-tailed degree distribution (BA hubs)."""
    J = graphs.barabasi_albert(20000, 4, seed=8, kind="pm1")
    print(f"persistent ls k={groups_k}", flush=True)
    prob = DeviceProblem(J, torch.float32)
    cands = _cands(prob.n, groups_k, seed=3)
    col = LS.colour_graph(prob, 5)
    a = LS.flip_refinement(prob, _bits(cands, gpu), groups_k, seed=5, colouring=col, persistent=True)
    b = LS.flip_refinement(prob, _bits(cands, gpu), groups_k, seed=5, colouring=col, persistent=False)
    np.testing.assert_array_equal(a.spins[:, :groups_k].cpu().numpy(), b.spins[:, :groups_k].cpu().numpy())
    assert (a.passes, a.total_passes) == (b.passes, b.total_passes)
    np.testing.assert_array_equal(a.e2_final, b.e2_final)
    assert np.all(a.e2_final >= a.e2_start)


def test_persistent_float_couplings(gpu):
    rng = np.random.default_rng(2)
    J = graphs.erdos_renyi(2500, 20 / 2499, seed=6, kind="pm1").astype(np.float64)
    J.data = J.data * rng.choice([0.25, 0.5, 1.5, 2.0], J.nnz)
    J = ((J + J.T) * 0.5).tocsr()
    print("persistent ls float", flush=True)
    prob = DeviceProblem(J, torch.float32)
    assert prob.jint is None
    k = 7
    cands = _cands(prob.n, k, seed=4)
    col = LS.colour_graph(prob, 11)
    a = LS.flip_refinement(prob, _bits(cands, gpu), k, seed=11, colouring=col, persistent=True,
                           final_sweep_fraction=0.7)
    b = LS.flip_refinement(prob, _bits(cands, gpu), k, seed=11, colouring=col, persistent=False,
                           final_sweep_fraction=0.7)
    assert not a.exact
    np.testing.assert_array_equal(a.spins[:, :k].cpu().numpy(), b.spins[:, :k].cpu().numpy())
    assert (a.passes, a.total_passes) == (b.passes, b.total_passes)
    np.testing.assert_array_equal(a.e2_final, b.e2_final)


def test_persistent_argument_errors(gpu):
    J = graphs.erdos_renyi(500, 10 / 499, seed=1, kind="pm1")
    prob = DeviceProblem(J, torch.float32)
    cands = _cands(prob.n, 2)
    with pytest.raises(Exception, match="fractions"):
        LS.flip_refinement(prob, _bits(cands, gpu), 2, seed=1, subset_fraction=0.0)
