"""SNAP signed-network conversion (datasets/snap_signed_to_txt.py:1-68): known answers worked out by
hand from the stated rule, on small edge-list and wikiElec files, and the round trip through the
instance loader."""
import gzip
import warnings

import numpy as np
import pytest

from specann_amd import snap
from specann_amd.data_loader import DataLoader

EDGES = """# Directed graph: sample
# FromNodeId\tToNodeId\tSign
0\t1\t1
1\t0\t1
2\t3\t-1
3\t2\t1
7\t7\t1
5\t2\t-1
0\t1\t-1
3\t0\t-1
"""
# ids from 0 -> shift 1; largest id 7 (a self-loop, dropped, still sets n) -> n = 8
# pair {1,2}: +1 +1 -1 = +1; pair {3,4}: -1 +1 = 0 (no edge); pair {3,6}: -1; pair {1,4}: -1
EDGES_OUT = "8 3\n" + "0\n" * 8 + "1 2 1\n3 6 -1\n1 4 -1\n"

WIKI = "\n".join([
    "E\t1", "T\t2004-09-21 01:15:53", "U\t3\tcand3", "N\t5\tnom5",
    "V\t1\t2\t2004-09-14 16:26:00\tu2",
    "V\t-1\t4\t2004-09-14 16:26:00\tu4",
    "V\t0\t5\t2004-09-14 16:26:00\tu5",
    "", "E\t0", "U\t2\tcand2",
    "V\t-1\t3\t2004-09-14 16:26:00\tu3",
    "V\t1\t1\t2004-09-14 16:26:00\tu1",
    "V\t1\t2\t2004-09-14 16:26:00\tu2",
]) + "\n"
# arcs 2->3 (+1), 4->3 (-1), [5->3 neutral, skipped], 3->2 (-1), 1->2 (+1), [2->2 self-loop];
# ids from 1 are kept; the neutral vote yields no arc, so voter 5 does not count: n = 4
# pair {2,3}: +1 -1 = 0 (no edge); {3,4}: -1; {1,2}: +1
WIKI_OUT = "4 2\n" + "0\n" * 4 + "3 4 -1\n1 2 1\n"


def _gz(path, text):
    with gzip.open(path, "wt") as f:
        f.write(text)
    return path


def test_edges_known_answer(tmp_path):
    src = _gz(tmp_path / "soc-sign-sample.txt.gz", EDGES)
    out = tmp_path / "sample.txt"
    assert snap.main(["edges", str(src), str(out)]) == 0
    assert out.read_text() == EDGES_OUT


def test_wikielec_known_answer(tmp_path):
    src = _gz(tmp_path / "wikiElec.sample.txt.gz", WIKI)
    out = tmp_path / "wiki.txt"
    assert snap.convert(snap.read_wikielec_arcs(src), out) == (4, 2)
    assert out.read_text() == WIKI_OUT


def test_bad_usage_exit_code(capsys):
    assert snap.main(["csv", "a", "b"]) == 2
    assert "usage" in capsys.readouterr().err


def test_pair_totals_match_brute_force():
    rng = np.random.default_rng(5)
    arcs = np.stack([rng.integers(0, 40, 3000), rng.integers(0, 40, 3000), rng.choice([-1, 1], 3000)], axis=1)
    n, edges = snap.signed_edges(arcs)
    assert n == int(arcs[:, :2].max()) + 1
    total, first = {}, {}
    for t, (a, b, s) in enumerate(arcs.tolist()):
        if a != b:
            key = (min(a, b) + 1, max(a, b) + 1)
            total[key] = total.get(key, 0) + s
            first.setdefault(key, t)
    want = [(i, j, 1 if v > 0 else -1) for (i, j), v in sorted(total.items(), key=lambda kv: first[kv[0]]) if v]
    assert [tuple(e) for e in edges.tolist()] == want


def test_round_trip_through_loader(tmp_path):
    d = tmp_path / "snap_signed"
    d.mkdir()
    src = _gz(tmp_path / "in.txt.gz", EDGES)
    snap.convert(snap.read_edge_arcs(src), d / "sample.txt")
    with warnings.catch_warnings(record=True):
        warnings.simplefilter("always")
        J, idx = DataLoader("snap_signed", str(tmp_path)).load_instance("sample.txt", return_index=True)
    # nodes 5, 7, 8 have no edge and are dropped; no fields, so no bias node
    assert idx.bias_node_index == -1
    assert list(idx.node_ids + idx.id_base) == [1, 2, 3, 4, 6]
    D = J.toarray()
    assert D[0, 1] == 1 and D[2, 4] == -1 and D[0, 3] == -1 and np.count_nonzero(D) == 6
    assert (D == D.T).all()
