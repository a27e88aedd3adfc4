"""The oracle's guarded local-search policies (models/local_search_guards/*,
models/spectral_annealing_safe.py), pinned by checks independent of their own arithmetic: networkx's
largest-first greedy colouring, energy differences computed from scratch, brute-force replays of the
step rules on tiny graphs, one-flip optimality and the reference's guard invariants."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import specann_oracle as O
from specann_amd import graphs


def _tiny(seed, n=40, p=0.2, kind="pm1"):
    return graphs.erdos_renyi(n, p, seed=seed, kind=kind)


def _energy(J, s):
    """-s^T J s / 2 from scratch (FP64 or exact int)."""
    J = sp.csr_matrix(J)
    s = np.asarray(s)
    if J.dtype.kind in "iu" or np.all(J.data == np.rint(J.data)):
        return -int(s.astype(np.int64) @ (J.astype(np.int64) @ s.astype(np.int64))) // 2
    return -0.5 * float(s @ (J @ s))


def _cands(n, k, seed):
    return np.stack([O.round_spins(O.normal_vector(n, seed, i)) for i in range(k)])


def test_largest_first_colouring_matches_networkx():
    nx = pytest.importorskip("networkx")
    for seed in range(3):
        J = graphs.barabasi_albert(300, 3, seed=seed, kind="pm1")
        G = nx.from_scipy_sparse_array(J)
        ref = nx.greedy_color(G, strategy="largest_first")
        got = O.largest_first_colouring(J.indptr, J.indices)
        assert [ref[i] for i in range(J.shape[0])] == got.tolist()


def test_jacobi_delta_formula_against_scratch_energies():
    """The proposal change sum_F 2 s h - 2 sum_F s_i sum_{j in F} J_ij s_j equals E(s') - E(s)."""
    rng = np.random.default_rng(1)
    for kind in ("pm1", "neg"):
        J = O.integer_couplings(_tiny(3, kind=kind))
        for _ in range(20):
            s = rng.choice([-1, 1], J.shape[0]).astype(np.int64)
            F = rng.random(J.shape[0]) < 0.4
            h = J @ s
            T = np.where(F, s, 0)
            dE = int((2 * T * h).sum() - (2 * T * (J @ T)).sum())
            s2 = np.where(F, -s, s)
            assert dE == _energy(J, s2) - _energy(J, s)


@pytest.mark.parametrize("adaptive", [False, True])
def test_safe_policy_invariants_and_exact_energies(adaptive):
    J = O.integer_couplings(_tiny(5, n=60, p=0.3))
    c = float(np.abs(J).sum(axis=1).max())
    cands = _cands(60, 6, 4)
    r = O.guard_safe(J, cands, 11, adaptive=adaptive, scale=c, max_iterations=50, final_sweep_iters=5,
                     final_sweep_rounds=3)
    for i in range(6):
        assert r.energies[i] == _energy(J, r.spins[i])      # tracked energy = scratch energy
        assert r.energies_start[i] == _energy(J, cands[i])
    assert np.all(r.energies <= r.energies_start)
    assert 1 <= r.iterations <= 50


def test_safe_rejects_increasing_proposals():
    """A dense graph from a random start: full Jacobi proposals overshoot, and the guard rejects them
    (no accepted round may raise the energy; the adaptive share shrinks and still makes progress)."""
    J = O.integer_couplings(_tiny(7, n=50, p=0.6))
    c = float(np.abs(J).sum(axis=1).max())
    cands = _cands(50, 4, 2)
    r_fixed = O.guard_safe(J, cands, 3, adaptive=False, scale=c, subset_fraction=1.0, max_iterations=30,
                           final_sweep_iters=0)
    r_adapt = O.guard_safe(J, cands, 3, adaptive=True, scale=c, subset_fraction=1.0, max_iterations=30,
                           final_sweep_iters=0)
    assert np.all(r_fixed.energies <= r_fixed.energies_start)
    assert np.all(r_adapt.energies <= r_adapt.energies_start)
    assert r_adapt.energies.sum() <= r_fixed.energies.sum()


def _is_one_flip_optimal(J, s):
    h = J @ s.astype(J.dtype)
    return bool(np.all(s * h >= 0))


def test_independent_set_certified_candidates_are_one_flip_optimal():
    J = O.integer_couplings(graphs.barabasi_albert(200, 3, seed=2, kind="pm1"))
    c = float(np.abs(J).sum(axis=1).max())
    cands = _cands(200, 5, 8)
    r = O.guard_independent_set(J, cands, 21, scale=c, max_updates=100000)
    for i in range(5):
        assert _is_one_flip_optimal(J, r.spins[i].astype(np.int64))
        assert r.energies[i] == _energy(J, r.spins[i])
    # the colouring is proper
    col = r.extra["colours"]
    rows = np.repeat(np.arange(200), np.diff(J.indptr))
    assert not np.any(col[rows] == col[J.indices])


def test_independent_set_replay_by_scratch_energies():
    """Brute-force replay of the step rule on a tiny graph: per step, each node of the class flips iff
    flipping it alone lowers the energy (computed from scratch); the classes in default_rng order."""
    J = O.integer_couplings(_tiny(9, n=24, p=0.3))
    cand = _cands(24, 1, 5)
    r = O.guard_independent_set(J, cand, 4, scale=1.0, max_updates=40)
    col = O.largest_first_colouring(J.indptr, J.indices)
    C = int(col.max()) + 1
    rng = np.random.default_rng(4)
    order, pos, s, sweep, active = rng.permutation(C), 0, cand[0].astype(np.int64).copy(), 0, True
    for _ in range(40):
        if not active:
            break
        cls = np.flatnonzero(col == order[pos])
        base = _energy(J, s)
        flips = [i for i in cls if _energy(J, np.where(np.arange(24) == i, -s, s)) < base]
        for i in flips:
            s[i] = -s[i]
        sweep += len(flips)
        pos += 1
        if pos == C:
            active = sweep > 0
            sweep, pos, order = 0, 0, rng.permutation(C)
    np.testing.assert_array_equal(r.spins[0], s)


def test_sequential_oracle_replay_and_optimality():
    J = O.integer_couplings(_tiny(13, n=30, p=0.25))
    cands = _cands(30, 3, 6)
    r = O.guard_sequential(J, cands, scale=1.0)
    assert r.extra["converged"]
    for i in range(3):
        s = cands[i].astype(np.int64).copy()
        steps = 0
        while True:
            base = _energy(J, s)
            d = [_energy(J, np.where(np.arange(30) == j, -s, s)) - base for j in range(30)]
            j = int(np.argmin(d))       # first node of the smallest change
            if d[j] >= 0:
                break
            s[j] = -s[j]
            steps += 1
        np.testing.assert_array_equal(r.spins[i], s)
        assert _is_one_flip_optimal(J, r.spins[i].astype(np.int64))
    assert r.iterations >= 1


def test_sequential_step_limit_reports_not_converged():
    J = O.integer_couplings(_tiny(13, n=30, p=0.25))
    r = O.guard_sequential(J, _cands(30, 2, 6), scale=1.0, max_steps=1)
    assert not r.extra["converged"] and r.iterations == 1


def test_float_couplings_use_tolerances():
    """FP32 couplings J / c: the policies track FP64 energies in units of J / c; results are one-flip
    optimal up to the tolerance and never rise."""
    rng = np.random.default_rng(3)
    J = _tiny(4, n=50, p=0.3).astype(np.float64)
    J.data *= rng.choice([0.5, 1.5, 2.25], J.nnz)
    J = ((J + J.T) * 0.5).tocsr()
    c = float(np.abs(J).sum(axis=1).max())
    Jn = (J / c).astype(np.float32)
    cands = _cands(50, 3, 1)
    for r in (O.guard_sequential(Jn, cands), O.guard_independent_set(Jn, cands, 2, max_updates=10000),
              O.guard_safe(Jn, cands, 1, adaptive=True, max_iterations=40, final_sweep_iters=4, final_sweep_rounds=2)):
        assert r.energies.dtype == np.float64
        assert np.all(r.energies <= r.energies_start + 1e-12)
        J64 = sp.csr_matrix(Jn).astype(np.float64)
        for i in range(3):
            assert r.energies[i] == pytest.approx(_energy(J64, r.spins[i].astype(np.float64)), abs=1e-9)
