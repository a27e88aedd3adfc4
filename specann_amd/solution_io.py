"""Best spin vectors as plain-text ``.sol`` files, in the reference's file format
(utils/solution_io.py:43-122), so that files written here read back with the reference's
``load_solution`` and the other way round.

    results/<dataset>/<run>/solutions/<instance>__seed<seed>.sol

Layout of a file:

* header lines ``# <key> <value>`` (an absent value leaves the text after the key empty), in the
  order dataset, instance, run_name, seed, n_original, n_solve, id_base, bias_node, gauge_flipped,
  weights_negated, J_sum, energy, cut_value;
* the line ``# columns node_id spin``;
* one line ``<node id> <spin>`` per node, the spin with an explicit sign (``+1`` / ``-1``).

Node ids are those of the instance file (isolated nodes, dropped by the loader, are absent). With
a bias node, the whole vector is flipped so that the bias spin is +1 (a gauge choice that leaves
the energy unchanged) and the bias row is left out; ``Solution.to_solve_space`` restores the
vector over the loaded matrix's rows. The file is written under a temporary name and renamed.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any, Optional

import numpy as np

# header keys whose values are read back as int / float; every other key stays text
INT_FIELDS = frozenset(('seed', 'n_original', 'n_solve', 'id_base', 'bias_node', 'gauge_flipped',
                        'weights_negated'))
FLOAT_FIELDS = frozenset(('J_sum', 'energy', 'cut_value'))
COLUMNS_LINE = "# columns node_id spin"


def ising_energy(J, s) -> float:
    """E = -1/2 s^T J s in float64."""
    v = np.asarray(s, dtype=np.float64).ravel()
    return float(-0.5 * v @ (J @ v))


def cut_value(energy: float, J_sum: float, weights_negated: bool) -> Optional[float]:
    """Cut weight (W - E) / 2 of a MaxCut instance loaded as J = -W, W = -J_sum / 2; None when the
    weights were not negated (the instance is not a MaxCut instance)."""
    return (-0.5 * J_sum - energy) / 2.0 if weights_negated else None


def solution_path(results_dir, instance_name: str, seed) -> Path:
    """<results_dir>/solutions/<instance file name without .txt>__seed<seed>.sol"""
    stem = Path(instance_name).name
    stem = stem[:-4] if stem.endswith('.txt') else stem
    return Path(results_dir) / "solutions" / f"{stem}__seed{seed}.sol"


def _host_spins(spins) -> np.ndarray:
    if hasattr(spins, 'detach'):          # a torch tensor, possibly on the GPU
        spins = spins.detach().cpu().numpy()
    return np.asarray(spins).astype(np.int8).ravel()


def save_solution(path, spins, J, index, *, dataset: str, instance_name: str, run_name: str, seed) -> Path:
    """Write the solver's spin vector (one entry per row of J) in the instance file's numbering."""
    path = Path(path)
    s = _host_spins(spins)
    energy = ising_energy(J, s)
    ids = np.asarray(index.node_ids, dtype=np.int64)
    bias = int(index.bias_node_index)
    flipped = 0
    if bias >= 0:
        if s[bias] < 0:
            s = (-s).astype(np.int8)
            flipped = 1
        s = np.delete(s, bias)
        ids = np.delete(ids, bias)
    fields = (
        ('dataset', dataset), ('instance', instance_name), ('run_name', run_name), ('seed', seed),
        ('n_original', index.n_original), ('n_solve', index.n_solve), ('id_base', index.id_base),
        ('bias_node', bias), ('gauge_flipped', flipped), ('weights_negated', int(bool(index.weights_negated))),
        ('J_sum', index.J_sum), ('energy', energy),
        ('cut_value', cut_value(energy, index.J_sum, index.weights_negated)),
    )
    lines = [f"# {key} {'' if value is None else value}" for key, value in fields]
    lines.append(COLUMNS_LINE)
    file_ids = ids + int(index.id_base)
    lines.extend(f"{int(i)} {int(v):+d}" for i, v in zip(file_ids.tolist(), s.tolist()))
    path.parent.mkdir(parents=True, exist_ok=True)
    tmp = path.parent / f"{path.name}.tmp.{os.getpid()}"
    tmp.write_text("\n".join(lines) + "\n")
    os.replace(tmp, path)
    return path


@dataclass
class Solution:
    """A spin vector in the instance file's node numbering, with the file's header."""
    node_ids: np.ndarray                      # int64, as written in the file
    spins: np.ndarray                         # int8 +-1, aligned with node_ids
    header: dict[str, Any] = field(default_factory=dict)

    @property
    def energy(self) -> Optional[float]:
        return self.header.get('energy')

    def to_solve_space(self) -> np.ndarray:
        """The spin vector over the loaded matrix's rows: the bias spin (+1) re-inserted at its row,
        then the gauge flip undone."""
        s = self.spins.astype(np.int8, copy=True)
        bias = self.header.get('bias_node')
        if bias is not None and int(bias) >= 0:
            s = np.insert(s, int(bias), np.int8(1))
        if int(self.header.get('gauge_flipped') or 0):
            s = (-s).astype(np.int8)
        return s


def _header_value(key: str, text: str):
    if text == '':
        return None
    if key in INT_FIELDS:
        return int(float(text))
    if key in FLOAT_FIELDS:
        return float(text)
    return text


def load_solution(path) -> Solution:
    """Read a ``.sol`` file (header, then node id / spin columns)."""
    header: dict[str, Any] = {}
    ids, spins = [], []
    with open(path) as f:
        for raw in f:
            line = raw.strip()
            if not line:
                continue
            if line.startswith('#'):
                if ids:
                    continue                  # comments after the body carry nothing
                key, _, rest = line[1:].strip().partition(' ')
                if key and key != 'columns':
                    header[key] = _header_value(key, rest.strip())
                continue
            a, b = line.split()[:2]
            ids.append(int(a))
            spins.append(int(b))
    return Solution(np.asarray(ids, dtype=np.int64), np.asarray(spins, dtype=np.int8), header)
