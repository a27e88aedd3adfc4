"""Instance files -> symmetric scipy CSR couplings (the semantics of utils/data_loader.py:1-204).

Two text formats, nodes numbered from 1:
* Gset (a dataset folder whose name contains "gset", any case): "n m", then "i j w" lines; weights
  are negated so the lowest energy is the largest cut (:112-134);
* otherwise "n m", then n lines with the field h_i, then "i j J" lines; nonzero fields become
  couplings to one extra bias node, row n (:136-178).
Self-loops and out-of-range ids are skipped, blank/#/% lines and unparsable lines ignored, duplicate
entries summed (COO -> CSR), and isolated nodes dropped with a warning (:180-204). The InstanceIndex
maps the solver's rows back to file node ids.
"""
from __future__ import annotations

import warnings
from dataclasses import dataclass
from pathlib import Path
from typing import List, Optional, Tuple

import numpy as np
from scipy.sparse import coo_matrix, csr_matrix


class DataLoadError(Exception):
    """An instance file cannot be read."""


@dataclass
class InstanceIndex:
    n_original: int
    n_solve: int
    node_ids: np.ndarray
    bias_node_index: int = -1
    id_base: int = 1
    weights_negated: bool = False
    J_sum: float = 0.0


def _parse_edges(lines, n: int):
    rows, cols, vals = [], [], []
    for line in lines:
        t = line.strip()
        if not t or t[0] in '#%':
            continue
        parts = t.split()
        if len(parts) < 3:
            continue
        try:
            i, j, w = int(float(parts[0])) - 1, int(float(parts[1])) - 1, float(parts[2])
        except (ValueError, TypeError):
            continue
        if 0 <= i < n and 0 <= j < n and i != j:
            rows.append(i)
            cols.append(j)
            vals.append(w)
    return np.array(rows, dtype=np.int64), np.array(cols, dtype=np.int64), np.array(vals, dtype=np.float64)


class DataLoader:
    """Load the instances of one dataset folder (datasets/<name>)."""

    def __init__(self, dataset_name: str, datasets_path: str = "datasets"):
        self.dataset_name = dataset_name
        self.dataset_path = Path(datasets_path) / dataset_name
        if not self.dataset_path.exists():
            raise ValueError(f"Dataset path does not exist: {self.dataset_path}")
        self.dataset_format = 'gset' if 'gset' in dataset_name.lower() else 'graph_linear_bias'

    def list_instances(self) -> List[str]:
        return sorted(f.name for f in self.dataset_path.glob("*.txt"))

    def load_instance(self, instance_name: str, return_index: bool = False):
        path = self.dataset_path / instance_name
        if not path.exists():
            raise FileNotFoundError(f"Instance file not found: {path}")
        try:
            J, index = self._load(path)
        except Exception as e:
            raise DataLoadError(f"Error reading {self.dataset_format} file {path}: {e}") from e
        return (J, index) if return_index else J

    def _load(self, path: Path) -> Tuple[csr_matrix, InstanceIndex]:
        with open(path) as f:
            n, _m = map(int, f.readline().split())
            if self.dataset_format == 'gset':
                r, c, w = _parse_edges(f, n)
                w = -w
                bias_id, size = -1, n
            else:
                h = np.array([float(f.readline().strip()) for _ in range(n)], dtype=np.float64)
                r, c, w = _parse_edges(f, n)
                nz = np.flatnonzero(h != 0)
                if nz.size:
                    bias_id, size = n, n + 1
                    r = np.concatenate([r, np.full(nz.size, n, dtype=np.int64)])
                    c = np.concatenate([c, nz])
                    w = np.concatenate([w, h[nz]])
                else:
                    bias_id, size = -1, n
        rows = np.concatenate([r, c])
        cols = np.concatenate([c, r])
        vals = np.concatenate([w, w])
        J = coo_matrix((vals, (rows, cols)), shape=(size, size)).tocsr()
        return self._finalize(J, n, bias_id, self.dataset_format == 'gset')

    @staticmethod
    def _finalize(J: csr_matrix, n_original: int, bias_id: int, negated: bool) -> Tuple[csr_matrix, InstanceIndex]:
        node_ids = np.arange(J.shape[0], dtype=np.int64)
        if J.nnz:
            row_abs = np.asarray(abs(J).sum(axis=1)).ravel()
            keep = np.flatnonzero(row_abs != 0)
            if keep.size < J.shape[0]:
                warnings.warn(f"Removed {J.shape[0] - keep.size} isolated node(s) (no couplings); the matrix has "
                              f"{keep.size} of {J.shape[0]} nodes.", UserWarning)
                J = J[keep][:, keep].tocsr()
                node_ids = node_ids[keep]
        bias_index = -1
        if bias_id >= 0:
            hits = np.flatnonzero(node_ids == bias_id)
            bias_index = int(hits[0]) if hits.size else -1
        return J, InstanceIndex(n_original, int(J.shape[0]), node_ids, bias_index, 1, negated, float(J.sum()))
