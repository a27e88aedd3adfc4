"""Write the golden fixtures of tests/golden/ (run once, in the development container):

    python tests/golden/make_golden.py

* ground_states.json: exhaustive ground-state energies of the smallest bundled instances (brute
  force over all 2^n spin vectors, exact integer arithmetic);
* loader_<instance>.npz: the CSR (indptr, indices, data), node ids and bias-node row that the
  reference's own DataLoader (utils/data_loader.py, host-only numpy/scipy code) produces for the
  bundled fixture instances; the build's loader is compared with them;
* sol_<instance>.sol: the file the reference's own save_solution (utils/solution_io.py:43-76,
  host-only numpy code) writes for the spin vector sol_spins(n) on a bias-node instance and on a
  tiny Gset-format instance (weights negated, so the cut value is filled in); the build's writer
  must produce the same bytes and its reader the same header and vector.
The reference is imported from /root/reference only here; the fixtures are data.
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
DATASET = "barabasi_albert_3ins_m10_tiny"
INSTANCES = ["ba_20_1_0.txt", "ba_25_1_1.txt", "ba_60_1_2.txt"]
REFERENCE = Path("/root/reference")
SOL_CASES = [(DATASET, "ba_20_1_0.txt"), ("tiny_gset", "g7.txt")]


def sol_spins(n: int) -> np.ndarray:
    """The spin vector of the .sol fixtures: s_i = +1 iff (5 i + 1) mod 3 == 0; the last spin (the bias
    row of a bias-node instance) is -1, so the writer's gauge flip is exercised."""
    s = np.where((5 * np.arange(n) + 1) % 3 == 0, 1, -1).astype(np.int8)
    s[-1] = -1
    return s


def brute_force_ground(J) -> int:
    """min over s in {-1,+1}^n of -1/2 s^T J s (integer J), in chunks of 2^16 vectors."""
    Jd = np.asarray(J.toarray(), dtype=np.int64)
    n = Jd.shape[0]
    best = None
    chunk = 1 << 16
    ar = np.arange(chunk, dtype=np.int64)
    for base in range(0, 1 << (n - 1), chunk):          # s_0 = +1 suffices (E(s) = E(-s))
        idx = base + ar[: min(chunk, (1 << (n - 1)) - base)]
        bits = (idx[:, None] >> np.arange(n - 1)) & 1
        S = np.concatenate([np.ones((idx.size, 1), dtype=np.int64), 2 * bits - 1], axis=1)
        e = -((S @ Jd) * S).sum(axis=1) // 2
        m = int(e.min())
        best = m if best is None else min(best, m)
    return best


def main() -> None:
    sys.path.insert(0, str(REFERENCE))
    from utils.data_loader import DataLoader as RefLoader  # reference host code, numpy/scipy only

    ref = RefLoader(DATASET, datasets_path=str(HERE / "instances"))
    grounds = {}
    for name in INSTANCES:
        J, index = ref.load_instance(name, return_index=True)
        np.savez(HERE / f"loader_{Path(name).stem}.npz", indptr=J.indptr, indices=J.indices, data=J.data,
                 shape=np.array(J.shape), node_ids=index.node_ids, bias=np.array([index.bias_node_index]),
                 n_original=np.array([index.n_original]), J_sum=np.array([index.J_sum]))
        if J.shape[0] <= 22:
            grounds[name] = {"n_solve": int(J.shape[0]), "ground_energy": brute_force_ground(J)}
            print(name, grounds[name])
    (HERE / "ground_states.json").write_text(json.dumps(grounds, indent=1) + "\n")

    from utils.solution_io import save_solution as ref_save  # reference host code, numpy only
    for dataset, name in SOL_CASES:
        J, index = RefLoader(dataset, datasets_path=str(HERE / "instances")).load_instance(name, return_index=True)
        out = ref_save(HERE / f"sol_{Path(name).stem}.sol", sol_spins(J.shape[0]), J, index, dataset=dataset,
                       instance_name=name, run_name="SpecAnn", seed=7)
        print("wrote", out.name)


if __name__ == "__main__":
    main()
