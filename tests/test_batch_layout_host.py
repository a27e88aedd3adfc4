"""Host build of the instance-staged batch layout's tables (CPU, torch on the CPU): every pass of every
row reads only entries whose columns lie in that pass's staged slice, the passes cover the row, and the
sign-packed entries and magnitudes reproduce the couplings."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from specann_amd import graphs
from specann_amd.batch import inst_arrays, pack


def _tables(Js, kc):
    Jp, row_off, sizes = pack(Js)
    c = np.zeros(Jp.shape[0])
    for b, J in enumerate(Js):
        c[row_off[b]:row_off[b] + sizes[b]] = abs(sp.csr_matrix(J)).sum(axis=1).max()
    vals = (Jp.data / c[np.repeat(np.arange(Jp.shape[0]), np.diff(Jp.indptr))]).astype(np.float32)
    t = inst_arrays(torch.from_numpy(Jp.indptr.astype(np.int32)), torch.from_numpy(Jp.indices.astype(np.int32)),
                    torch.from_numpy(vals), torch.from_numpy(Jp.data.astype(np.int32)), row_off, sizes, kc)
    return Jp, row_off, sizes, vals, t


@pytest.mark.parametrize("kc", [64, 128, 1 << 15])
def test_every_pass_reads_only_its_staged_columns(kc):
    Js = [graphs.erdos_renyi(300, 12 / 299, seed=1, kind="pm1"),
          graphs.barabasi_albert(200, 4, seed=2, kind="neg"),
          graphs.erdos_renyi(70, 0.2, seed=3, kind="neg"),
          graphs.barabasi_albert(129, 3, seed=4, kind="pm1")]
    Jp, row_off, sizes, vals, t = _tables(Js, kc)
    passes = t["passes"]
    assert passes == -(-max(sizes) // kc)
    n = Jp.shape[0]
    split = t["split"].numpy().reshape(passes - 1, n) if passes > 1 else np.zeros((0, n), dtype=np.int32)
    base, real, end = t["base"].numpy(), t["real"].numpy(), t["end"].numpy()
    assert np.array_equal(base, row_off[:-1]) and np.array_equal(end, row_off[1:])
    assert np.array_equal(real, row_off[:-1] + sizes)
    for b in range(len(Js)):
        for r in range(row_off[b], row_off[b + 1]):
            width = int(real[b] - base[b])
            mine = -(-width // kc)
            bounds = [int(Jp.indptr[r])] + [int(split[c - 1, r]) for c in range(1, passes)] + [int(Jp.indptr[r + 1])]
            assert bounds == sorted(bounds)
            for c in range(passes):
                cols = Jp.indices[bounds[c]:bounds[c + 1]]
                if c >= mine:                      # passes past this instance's own see no entries
                    assert cols.size == 0
                    continue
                lo = base[b] + c * kc
                hi = base[b] + min((c + 1) * kc, width)
                assert np.all((cols >= lo) & (cols < hi)), (b, r, c)


def test_sign_packed_entries_and_magnitudes():
    Js = [graphs.erdos_renyi(150, 0.1, seed=5, kind="pm1"), graphs.barabasi_albert(90, 3, seed=6, kind="neg")]
    Jp, row_off, sizes, vals, t = _tables(Js, 64)
    assert t["signed"]
    ent = t["ent"].numpy().view(np.uint32)
    assert np.array_equal(ent & 0x7FFFFFFF, Jp.indices)
    assert np.array_equal((ent >> 31).astype(bool), vals < 0)
    mag = t["mag"].numpy()
    for b in range(2):
        e = slice(Jp.indptr[row_off[b]], Jp.indptr[row_off[b] + sizes[b]])
        assert np.all(np.abs(vals[e]) == mag[b])


def test_mixed_magnitudes_keep_the_value_stream():
    J = graphs.erdos_renyi(120, 0.1, seed=7, kind="pm1").tolil()
    i, j = sp.triu(J, 1).nonzero()
    J[i[0], j[0]] = J[j[0], i[0]] = 3
    _, _, _, _, t = _tables([J.tocsr(), graphs.erdos_renyi(80, 0.1, seed=8, kind="neg")], 64)
    assert not t["signed"] and "ent" not in t
