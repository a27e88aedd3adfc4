"""Host-side logic on the CPU: loader vs the reference's own output, config merge, registry,
results CSVs, failure tagging, and the errors the solvers raise before touching the GPU."""
import json
import warnings
from pathlib import Path
from types import SimpleNamespace

import numpy as np
import pandas as pd
import pytest
import scipy.sparse as sp

from specann_amd import config as C
from specann_amd import graphs
from specann_amd.data_loader import DataLoader
from specann_amd.failures import NumericalFailure, is_numerical, is_out_of_memory
from specann_amd.results import store_alphas, store_general
from specann_amd.selector import SOLVERS, get_solver_class

GOLDEN = Path(__file__).resolve().parent / "golden"
DATASET = "barabasi_albert_3ins_m10_tiny"


@pytest.mark.parametrize("name", ["ba_20_1_0", "ba_25_1_1", "ba_60_1_2"])
def test_loader_matches_reference_vectors(name):
    ref = np.load(GOLDEN / f"loader_{name}.npz")
    J, index = DataLoader(DATASET, str(GOLDEN / "instances")).load_instance(name + ".txt", return_index=True)
    assert tuple(J.shape) == tuple(ref["shape"])
    np.testing.assert_array_equal(J.indptr, ref["indptr"])
    np.testing.assert_array_equal(J.indices, ref["indices"])
    np.testing.assert_array_equal(J.data, ref["data"])
    np.testing.assert_array_equal(index.node_ids, ref["node_ids"])
    assert index.bias_node_index == int(ref["bias"][0])
    assert index.n_original == int(ref["n_original"][0])
    assert index.J_sum == pytest.approx(float(ref["J_sum"][0]))


def test_loader_gset_negates_and_drops_isolated(tmp_path):
    d = tmp_path / "MyGset"
    d.mkdir()
    (d / "g.txt").write_text("4 3\n1 2 1\n2 3 2\n% comment\n1 1 5\n")
    with warnings.catch_warnings(record=True):
        warnings.simplefilter("always")
        J, idx = DataLoader("MyGset", str(tmp_path)).load_instance("g.txt", return_index=True)
    assert J.shape == (3, 3)                      # node 4 isolated, self-loop skipped
    assert J[0, 1] == -1 and J[1, 2] == -2
    assert idx.weights_negated and idx.bias_node_index == -1


def test_config_merge_and_float_resolver():
    cfg = C.load_config("default")
    assert isinstance(cfg["SpecAnn"]["params"]["eigsh_tol"], float)
    args = SimpleNamespace(solver="SpecAnn-64", dataset="X", instance=["a.txt"], save_spins=None,
                           num_seeds=None, seed=7, datasets_path="datasets")
    m = C.merge_args_with_config(args, cfg)
    assert m["precision"] == 64 and m["solver"] == "spectral_annealing" and m["run_name"] == "SpecAnn-64"
    assert m["random_seed"] == 7 and m["num_seeds"] == 1 and m["instance"] == ["a.txt"]
    with pytest.raises(ValueError):
        C.merge_args_with_config(SimpleNamespace(solver="Nope", dataset="X", instance=None), cfg)
    for run in ("SpecAnn", "SpecAnn-64", "SpecAnn-LS", "SpecAnn-cold", "SpecAnn-sec", "SpecAnn-guard-adaptive",
                "SpecAnn-sched-example", "SpecAdj", "SpecSignedLap"):
        assert run in cfg


def test_registry_and_constructor_signatures():
    import inspect
    assert set(SOLVERS) == {"spectral_annealing", "spec_adj", "spec_signed_lap"}
    sa = get_solver_class("spectral_annealing")
    params = inspect.signature(sa.__init__).parameters
    for name, default in [("random_seed", None), ("mode", 0), ("use_local_search", True),
                          ("max_local_iterations", 1000), ("eigsh_tol", 1e-6), ("eigsh_maxiter", 1000),
                          ("precision", 32), ("pred_mode", "diagonal"), ("pred_tol", 1e-3), ("shift_eps", 1e-3),
                          ("schedule", "sqrt"), ("num_alphas", 128), ("schedule_interval", (0.0, 1.0)),
                          ("eigsh_ncv", None), ("fp32_spmv", None), ("ls_policy", "default"), ("ls_batch", True)]:
        assert params[name].default == default, name
    with pytest.raises(ValueError):
        get_solver_class("bogus")


def test_constructor_validation():
    sa = get_solver_class("spectral_annealing")
    with pytest.raises(ValueError):
        sa("i", "d", pred_mode="bad", verbose=False)
    with pytest.raises(ValueError):
        sa("i", "d", precision=16, verbose=False)
    with pytest.raises(ValueError):
        sa("i", "d", num_alphas=1, verbose=False)
    with pytest.raises(ValueError):
        sa("i", "d", precision=64, fp32_spmv=True, verbose=False)
    with pytest.raises(ValueError):
        sa("i", "d", ls_policy="fixed_guard", ls_batch=False, verbose=False)
    a = sa("i", "d", schedule="linear", num_alphas=5, verbose=False)
    assert a._get_alpha_values(0) == [0.0, 0.25, 0.5, 0.75, 1.0]
    with pytest.raises(ValueError):
        a._get_alpha_values(1)


def test_guard_policies_registry_and_bit_packing():
    """ls_policy values (models/local_search_guards/policies.py:13-18, :55) and the candidate masks
    the guard kernels read (bit l of word g = candidate 64 g + l)."""
    from specann_amd import guards
    from specann_amd.spectral_annealing import LS_POLICIES
    assert guards.LS_POLICIES == LS_POLICIES
    assert set(guards.POLICIES) == {"fixed_guard", "adaptive_guard", "independent_set", "sequential_oracle"}
    sa = get_solver_class("spectral_annealing")
    assert sa("i", "d", ls_policy="adaptive_guard", verbose=False).ls_policy == "adaptive_guard"
    with pytest.raises(ValueError, match="ls_batch"):
        sa("i", "d", ls_policy="fixed_guard", ls_batch=False, verbose=False)
    rng = np.random.default_rng(0)
    for k in (1, 63, 64, 65, 130):
        m = rng.random(k) < 0.5
        groups = (k + 63) // 64
        words = guards._bits_per_group(m, groups)
        assert words.dtype == np.uint64 and words.size == groups
        back = [(int(words[c // 64]) >> (c % 64)) & 1 for c in range(k)]
        assert back == m.astype(int).tolist()


def test_failure_tagging():
    assert is_out_of_memory(MemoryError())
    try:
        try:
            raise RuntimeError("HIP out of memory: tried to allocate")
        except RuntimeError as inner:
            raise ValueError("wrapper") from inner
    except ValueError as e:
        assert is_out_of_memory(e)
    assert not is_out_of_memory(ValueError("x"))
    from specann_amd._lib import EigensolverBreakdown
    assert is_numerical(EigensolverBreakdown("x")) and not is_numerical(KeyError("x"))
    assert issubclass(NumericalFailure, RuntimeError)


def test_results_csv_upsert_and_alpha_replacement(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    row = dict(instance_name="a.txt", dataset="D", seed=1, solver_name="R", energy=-3.0)
    store_general("D", "R", "a.txt", row)
    store_general("D", "R", "a.txt", dict(row, energy=-5.0))
    df = pd.read_csv(tmp_path / "results/D/R/a.csv")
    assert len(df) == 1 and df.energy.iloc[0] == -5.0
    rows = [dict(instance_name="a.txt", dataset="D", seed=1, solver_name="R", alpha=a) for a in (0.5, 0.1)]
    store_alphas("D", "R", "a.txt", 1, rows)
    store_alphas("D", "R", "a.txt", 1, rows[:1])
    df = pd.read_csv(tmp_path / "results/D/R/alphas/a_alphas.csv")
    assert list(df.alpha) == [0.5]


def test_ground_state_fixture_present():
    g = json.loads((GOLDEN / "ground_states.json").read_text())
    assert g["ba_20_1_0.txt"]["ground_energy"] < 0


def _sol_spins(n):
    s = np.where((5 * np.arange(n) + 1) % 3 == 0, 1, -1).astype(np.int8)
    s[-1] = -1
    return s


@pytest.mark.parametrize("dataset,name", [(DATASET, "ba_20_1_0.txt"), ("tiny_gset", "g7.txt")])
def test_solution_file_matches_reference_writer(tmp_path, dataset, name):
    """The .sol writer produces the bytes the reference's save_solution wrote for the same vector
    (tests/golden/sol_*.sol), and the reader restores the header and the solver-space vector."""
    from specann_amd.solution_io import Solution, load_solution, save_solution, solution_path
    J, idx = DataLoader(dataset, str(GOLDEN / "instances")).load_instance(name, return_index=True)
    s = _sol_spins(J.shape[0])
    golden = GOLDEN / f"sol_{Path(name).stem}.sol"
    p = save_solution(solution_path(tmp_path, name, 7), s, J, idx, dataset=dataset, instance_name=name,
                      run_name="SpecAnn", seed=7)
    assert p.name == f"{Path(name).stem}__seed7.sol"
    assert p.read_bytes() == golden.read_bytes()
    sol = load_solution(golden)
    assert isinstance(sol, Solution)
    assert sol.header["run_name"] == "SpecAnn" and sol.header["seed"] == 7 and sol.header["dataset"] == dataset
    assert sol.energy == pytest.approx(-0.5 * s @ (J @ s.astype(np.float64)))
    np.testing.assert_array_equal(sol.to_solve_space(), s)
    if idx.bias_node_index >= 0:
        assert sol.header["gauge_flipped"] == 1 and sol.header["cut_value"] is None
        assert sol.node_ids.size == J.shape[0] - 1 and np.all(sol.spins[:3] == -s[:3])
    else:
        assert sol.header["weights_negated"] == 1
        # a cut of the unit-weight graph: (W - E) / 2 with W = 9 edges
        assert sol.header["cut_value"] == pytest.approx((9 - sol.energy) / 2)
        np.testing.assert_array_equal(sol.node_ids, np.arange(1, 8))


def test_generators_symmetric_and_simple():
    for J in (graphs.erdos_renyi(3000, 10 / 2999, seed=1, kind="pm1"), graphs.barabasi_albert(3000, 5, seed=2)):
        assert abs(J - J.T).sum() == 0
        assert J.diagonal().sum() == 0
        assert J.has_sorted_indices
        K = J.copy()
        K.sum_duplicates()
        assert K.nnz == J.nnz
    assert np.diff(graphs.barabasi_albert(3000, 5, seed=2).indptr).min() >= 1   # BA drops isolated nodes
    er = graphs.erdos_renyi(20000, 40 / 19999, seed=3)
    assert abs(er.nnz / 20000 - 40) < 1.0
    ba = graphs.barabasi_albert(20000, 20, seed=4)
    assert 38 < ba.nnz / 20000 <= 40
    J, off = graphs.packed_batch([graphs.erdos_renyi(50, 0.2, seed=i) for i in range(3)])
    assert J.shape == (150, 150) and list(off) == [0, 50, 100, 150]


def test_bench_refuses_multi_gpu(capsys):
    import bench
    assert bench.main(["--gpus", "2"]) == 2
    assert "one GPU" in capsys.readouterr().err


def test_run_pipeline_order_seeds_failures_and_warmup(tmp_path, monkeypatch):
    """The pipeline behind main.py with a stand-in solver: instance order by the first number in the
    name, constructor-param filtering, seed range, failure tagging, warm-up clean-up, spin files."""
    from specann_amd import core
    calls = []

    class FakeSolver:
        def __init__(self, instance_name, dataset, run_name=None, random_seed=None, eigsh_tol=1e-6, **kwargs):
            self.instance_name, self.seed, self.tol = instance_name, random_seed, eigsh_tol
            self.out = Path("results") / dataset / run_name

        def solve(self, J):
            calls.append((self.instance_name, self.seed, self.tol))
            self.out.mkdir(parents=True, exist_ok=True)
            (self.out / f"{Path(self.instance_name).stem}.csv").write_text("x\n")
            if self.instance_name == "ba_25_1_1.txt" and self.seed == 6:
                raise MemoryError("device exhausted")
            if self.instance_name == "ba_25_1_1.txt" and self.seed == 7:
                raise KeyError("boom")
            n = J.shape[0]
            return -1.0, np.ones(n, dtype=np.float32), 0.0

    monkeypatch.chdir(tmp_path)
    monkeypatch.setattr(core, "get_solver_class", lambda name: FakeSolver)
    cfg = {"solver": "fake", "dataset": DATASET, "datasets_path": str(GOLDEN / "instances"), "run_name": "R",
           "random_seed": 5, "num_seeds": 3, "warmup": True, "save_spins": True, "eigsh_tol": 1e-9,
           "not_a_param": 1, "instance": ["ba_60_1_2.txt", "ba_25_1_1.txt", "missing_3.txt"]}
    with pytest.warns(UserWarning, match="not_a_param"):
        out = core.run_solver(cfg)
    # the warm-up ran first and left no files behind
    assert calls[0] == (core.WARMUP_INSTANCE, 5, 1e-9)
    assert not list((tmp_path / "results").rglob("*warmup*"))
    # instances by size: missing_3 (load failure, one per seed), ba_25 (seeds 5..7), ba_60
    assert [c[:2] for c in calls[1:]] == [("ba_25_1_1.txt", 5), ("ba_25_1_1.txt", 6), ("ba_25_1_1.txt", 7),
                                          ("ba_60_1_2.txt", 5), ("ba_60_1_2.txt", 6), ("ba_60_1_2.txt", 7)]
    assert out["total_runs"] == 9 and out["failed_runs"] == 5 and out["successful_runs"] == 4
    msgs = {(i, s): m for i, s, m in out["failures"]}
    assert msgs[("ba_25_1_1.txt", 6)].startswith("out of memory")
    assert msgs[("ba_25_1_1.txt", 7)].startswith("KeyError")
    assert all(msgs[("missing_3.txt", s)].startswith("load: ") for s in (5, 6, 7))
    sols = sorted(p.name for p in (tmp_path / "results" / DATASET / "R" / "solutions").iterdir())
    assert sols == ["ba_25_1_1__seed5.sol", "ba_60_1_2__seed5.sol", "ba_60_1_2__seed6.sol", "ba_60_1_2__seed7.sol"]


def test_main_exit_status(tmp_path, monkeypatch):
    """main.py exits with status 2 when a run failed and 0 otherwise."""
    import main as cli
    from specann_amd import core

    class Ok:
        def __init__(self, instance_name, dataset, run_name=None, random_seed=None, **kw):
            self.name = instance_name

        def solve(self, J):
            if self.name == "ba_25_1_1.txt":
                raise RuntimeError("no")
            return 0.0, np.ones(J.shape[0]), 0.0

    monkeypatch.chdir(tmp_path)
    monkeypatch.setattr(core, "get_solver_class", lambda name: Ok)
    base = ["--solver", "SpecAnn", "--dataset", DATASET, "--datasets-path", str(GOLDEN / "instances")]
    assert cli.main(base + ["--instance", "ba_20_1_0.txt"]) == 0
    assert cli.main(base + ["--instance", "ba_25_1_1.txt", "--seed", "3"]) == 2


@pytest.mark.parametrize("kind", ["neg", "pm1"])
def test_native_generators_structure_and_thread_independence(kind):
    """libspecann's host generators: symmetric, sorted, no self-loops or duplicates, no isolated nodes,
    the expected mean degree, and the same graph for any thread count (counter-based draws)."""
    for make in (lambda t: graphs.erdos_renyi_native(20000, 40 / 19999, seed=3, kind=kind, threads=t),
                 lambda t: graphs.barabasi_albert_native(20000, 20, seed=4, kind=kind, threads=t)):
        J1, J8 = make(1), make(8)
        assert (J1 != J8).nnz == 0 and np.array_equal(J1.indptr, J8.indptr)
        assert abs(J1 - J1.T).sum() == 0 and J1.diagonal().sum() == 0
        K = J1.copy()
        K.sum_duplicates()
        assert K.nnz == J1.nnz and J1.has_sorted_indices
        assert np.all(np.diff(J1.indices[J1.indptr[0]:J1.indptr[1]]) > 0)
        assert np.diff(J1.indptr).min() >= 1
        assert set(np.unique(J1.data)) <= ({-1.0} if kind == "neg" else {-1.0, 1.0})
        assert 38.5 < J1.nnz / J1.shape[0] <= 40.5
    ba = graphs.barabasi_albert_native(20000, 20, seed=4, kind=kind)
    deg = np.diff(ba.indptr)
    assert deg.max() > 20 * np.median(deg)          # preferential attachment: heavy-tailed degrees
    er = graphs.erdos_renyi_native(20000, 40 / 19999, seed=3, kind=kind)
    assert np.diff(er.indptr).max() < 90            # binomial degrees: no hubs
    if kind == "pm1":
        assert abs(er.data.mean()) < 0.02


def test_int32_field_guard():
    """Integer couplings take the exact int32-field local search only while the largest absolute
    degree stays below 2^29 (fields are bounded by it and a flip adds 2 J_ij s_j)."""
    from specann_amd.device import INT_FIELD_LIMIT, exact_int_fields
    assert INT_FIELD_LIMIT == 2.0 ** 29
    assert exact_int_fields(True, 40.0) and exact_int_fields(True, 2.0 ** 29 - 64)
    assert not exact_int_fields(True, 2.0 ** 29) and not exact_int_fields(True, 2.0 ** 31)
    assert not exact_int_fields(False, 3.0)
    assert (2 ** 29 - 1) * 2 < 2 ** 31 - 1     # a field plus the largest update stays in int32
