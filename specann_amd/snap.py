"""SNAP signed networks -> the linear-bias instance format (the rule of datasets/snap_signed_to_txt.py:1-68).

Inputs (gzip text):
* ``edges``: soc-sign-epinions / soc-sign-Slashdot files, one directed arc "from to sign" per line
  (sign +1 or -1), '#' comment lines, node ids from 0;
* ``wikielec``: wikiElec.ElecBs3, tab-separated; a "U <id>" line names the candidate of the votes that
  follow, each "V <vote> <voter> ..." line is an arc voter -> candidate with the vote as its sign,
  neutral votes (0) skipped; user ids from 1.

Rule (snap_signed_to_txt.py:6-13, :44-59): self-loops are dropped; every arc of an unordered pair
{i, j}, in either direction, adds its sign to the pair's total, and the pair becomes one edge with
J_ij = sign(total), none when the total is 0; ids of a file whose smallest id is 0 are shifted by one
(n is the largest id after the shift), ids of a 1-based file are kept; all fields are 0; each edge is
written once, smaller row first, in the order its pair first appears in the input.

The pairs are merged with numpy over the whole arc list (np.unique on a 64-bit pair key, totals by
bincount) rather than a per-arc dictionary. Output: "n m", n lines "0", then m lines "i j J".

Usage: python -m specann_amd.snap {edges|wikielec} <input.gz> <output.txt>
"""
from __future__ import annotations

import gzip
import sys
from pathlib import Path

import numpy as np


def read_edge_arcs(path) -> np.ndarray:
    """(k, 3) int64 arcs (from, to, sign) of an edge-list file; '#' lines skipped."""
    out = []
    with gzip.open(path, "rt") as f:
        for line in f:
            if line.startswith("#"):
                continue
            parts = line.split()
            if parts:
                out.append((int(parts[0]), int(parts[1]), int(parts[2])))
    return np.asarray(out, dtype=np.int64).reshape(-1, 3)


def read_wikielec_arcs(path) -> np.ndarray:
    """(k, 3) int64 arcs (voter, candidate, vote) of a wikiElec file, neutral votes skipped."""
    out = []
    candidate = None
    with gzip.open(path, "rt", encoding="latin-1") as f:
        for line in f:
            parts = line.rstrip("\n").split("\t")
            tag = parts[0]
            if tag == "U":
                candidate = int(parts[1])
            elif tag == "V":
                vote = int(parts[1])
                if vote != 0:
                    out.append((int(parts[2]), candidate, vote))
    return np.asarray(out, dtype=np.int64).reshape(-1, 3)


def signed_edges(arcs) -> tuple[int, np.ndarray]:
    """(n, (m, 3) int64 edges i < j, J = +-1, 1-based) from arcs (from, to, sign), in first-appearance
    order of the pairs."""
    A = np.asarray(arcs, dtype=np.int64).reshape(-1, 3)
    if A.shape[0] == 0:
        raise ValueError("no arcs in the input")
    ends = A[:, :2]
    shift = 1 if int(ends.min()) == 0 else 0
    n = int(ends.max()) + shift
    keep = A[:, 0] != A[:, 1]
    lo = np.minimum(A[keep, 0], A[keep, 1]) + shift
    hi = np.maximum(A[keep, 0], A[keep, 1]) + shift
    key = lo * (n + 1) + hi
    uniq, first, inv = np.unique(key, return_index=True, return_inverse=True)
    total = np.bincount(inv.ravel(), weights=A[keep, 2], minlength=uniq.size)
    order = np.argsort(first, kind="stable")
    total = total[order]
    nz = total != 0
    i = (uniq[order] // (n + 1))[nz]
    j = (uniq[order] % (n + 1))[nz]
    J = np.where(total[nz] > 0, 1, -1).astype(np.int64)
    return n, np.stack([i, j, J], axis=1)


def write_instance(n: int, edges: np.ndarray, out_path) -> None:
    """The linear-bias text format with all fields 0."""
    with open(out_path, "w") as f:
        f.write(f"{n} {edges.shape[0]}\n")
        f.write("0\n" * n)
        if edges.shape[0]:
            f.write("\n".join(f"{a} {b} {w}" for a, b, w in edges.tolist()))
            f.write("\n")


def convert(arcs, out_path) -> tuple[int, int]:
    """Write the instance of the arcs to out_path; returns (n, number of edges)."""
    n, edges = signed_edges(arcs)
    write_instance(n, edges, out_path)
    return n, int(edges.shape[0])


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) != 3 or argv[0] not in ("edges", "wikielec"):
        print("usage: python -m specann_amd.snap {edges|wikielec} <input.gz> <output.txt>", file=sys.stderr)
        return 2
    kind, src, dst = argv
    arcs = read_wikielec_arcs(Path(src)) if kind == "wikielec" else read_edge_arcs(Path(src))
    print(convert(arcs, dst))
    return 0


if __name__ == "__main__":
    sys.exit(main())
