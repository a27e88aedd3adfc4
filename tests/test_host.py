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


def test_guard_policy_fails_with_clear_error():
    sa = get_solver_class("spectral_annealing")
    s = sa("i", "d", ls_policy="adaptive_guard", verbose=False)
    with pytest.raises(NotImplementedError, match="guarded"):
        s.solve(graphs.erdos_renyi(20, 0.3, seed=1))


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


def test_solution_roundtrip(tmp_path):
    from specann_amd.solution_io import load_solution, save_solution, solution_path
    J, idx = DataLoader(DATASET, str(GOLDEN / "instances")).load_instance("ba_20_1_0.txt", return_index=True)
    s = np.where(np.arange(J.shape[0]) % 2 == 0, 1, -1)
    p = save_solution(solution_path(tmp_path, "ba_20_1_0.txt", 3), s, J, idx, dataset=DATASET,
                      instance_name="ba_20_1_0.txt", run_name="R", seed=3)
    header, ids, spins = load_solution(p)
    assert ids.size == J.shape[0] - (1 if idx.bias_node_index >= 0 else 0)
    assert float(header["energy"]) == pytest.approx(-0.5 * s @ (J @ s))


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
