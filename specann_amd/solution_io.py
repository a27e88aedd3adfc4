"""Best spin vectors as plain-text .sol files (the format of utils/solution_io.py:1-122).

    results/<dataset>/<run>/solutions/<instance>__seed<seed>.sol

A '#' header of key: value lines, then "node_id spin" lines in the file's node numbering. With a
bias node, all spins are flipped so that it is +1 and it is dropped. The file is written to a
temporary name and moved into place.
"""
from __future__ import annotations

import os
from pathlib import Path

import numpy as np


def ising_energy(J, s) -> float:
    v = np.asarray(s, dtype=np.float64).ravel()
    return float(-0.5 * v @ (J @ v))


def cut_value(energy: float, J_sum: float, weights_negated: bool):
    if not weights_negated:
        return None
    return (-0.5 * J_sum - energy) / 2.0


def solution_path(results_dir, instance_name: str, seed) -> Path:
    base = Path(instance_name).name
    if base.endswith('.txt'):
        base = base[:-4]
    return Path(results_dir) / "solutions" / f"{base}__seed{seed}.sol"


def save_solution(path, spins, J, index, *, dataset: str, instance_name: str, run_name: str, seed) -> Path:
    path = Path(path)
    s = np.asarray(spins.cpu() if hasattr(spins, 'cpu') else spins).astype(np.int8).ravel()
    energy = ising_energy(J, s)
    node_ids = np.asarray(index.node_ids, dtype=np.int64)
    flipped = False
    if index.bias_node_index >= 0:
        b = int(index.bias_node_index)
        if s[b] == -1:
            s, flipped = (-s).astype(np.int8), True
        keep = np.arange(s.size) != b
        s, node_ids = s[keep], node_ids[keep]
    header = {
        'dataset': dataset, 'instance': instance_name, 'run': run_name, 'seed': seed,
        'n_original': index.n_original, 'n_solve': index.n_solve, 'id_base': index.id_base,
        'bias_node': index.bias_node_index, 'gauge_flipped': int(flipped),
        'weights_negated': int(index.weights_negated), 'J_sum': index.J_sum, 'energy': energy,
        'cut_value': cut_value(energy, index.J_sum, index.weights_negated),
    }
    path.parent.mkdir(parents=True, exist_ok=True)
    tmp = path.with_suffix(path.suffix + '.tmp')
    with open(tmp, 'w') as f:
        for k, v in header.items():
            f.write(f"# {k}: {v}\n")
        for nid, sp in zip(node_ids + index.id_base, s):
            f.write(f"{int(nid)} {int(sp)}\n")
    os.replace(tmp, path)
    return path


def load_solution(path):
    """(header dict, node ids, spins) of a .sol file."""
    header, ids, spins = {}, [], []
    with open(path) as f:
        for line in f:
            if line.startswith('#'):
                k, _, v = line[1:].partition(':')
                header[k.strip()] = v.strip()
            elif line.strip():
                a, b = line.split()
                ids.append(int(a))
                spins.append(int(b))
    return header, np.array(ids, dtype=np.int64), np.array(spins, dtype=np.int8)
