"""Floating-point couplings on the GPU: the local search with FP32 J / c and exact FP64 fields is
bit-exact against the oracle's restatement, energies follow fp32_energies (models/spectral_annealing.py
:962-976), sweeps run end to end, integer couplings too large for int32 fields take this path, and
couplings whose dynamic range would make the fields inexact are refused with a clear error."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from oracle import specann_oracle as O
from specann_amd import graphs
from specann_amd import local_search as LS
from specann_amd.device import DeviceProblem
from specann_amd.spectral_annealing import SpectralAnnealing, unpack_bits

pytestmark = pytest.mark.gpu


def real_couplings(n, deg, seed, scale=1.5):
    """Symmetric couplings on an Erdos-Renyi pattern with weights uniform in [-scale, scale]."""
    P = graphs.erdos_renyi(n, deg / (n - 1), seed=seed, kind="neg")
    rng = np.random.default_rng(seed)
    U = sp.triu(P, 1).tocoo()
    w = rng.uniform(-scale, scale, U.nnz)
    A = sp.coo_matrix((w, (U.row, U.col)), shape=P.shape)
    return (A + A.T).tocsr()


def _bits(cands):
    return torch.from_numpy(np.stack([O.spin_bits(x) for x in cands]).view(np.int64)).cuda()


@pytest.mark.parametrize("tdtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("k", [1, 5, 70])
def test_float_local_search_bit_exact(gpu, tdtype, k):
    J = real_couplings(3000, 16, seed=3)
    print(f"float local search dtype={tdtype} k={k}", flush=True)
    prob = DeviceProblem(J, tdtype)
    assert prob.jint is None
    cands = np.stack([O.round_spins(O.normal_vector(prob.n, 9, i)) for i in range(k)])
    st = LS.flip_refinement(prob, _bits(cands), k, seed=42)
    assert not st.exact
    # the search's couplings: J / c normalised in the problem's precision, then rounded to FP32
    Jn, d, c = O.normalize(J, np.float32 if tdtype == torch.float32 else np.float64)
    ref = O.local_search(Jn, cands, 42)
    np.testing.assert_array_equal(st.spins[:, :k].cpu().numpy().T, ref.spins)
    assert st.passes == ref.passes
    np.testing.assert_allclose(st.energies_start, O.energies_float(Jn, cands.T, c), rtol=1e-12)
    np.testing.assert_allclose(st.energies_final, O.energies_float(Jn, ref.spins.T, c), rtol=1e-12)
    # the GPU's FP64 reduction of the exact fields equals the oracle's to rounding
    np.testing.assert_allclose(st.e2_final, ref.e2_final, rtol=1e-13)


def test_float_energies_of_bits(gpu):
    J = real_couplings(2000, 12, seed=5)
    prob = DeviceProblem(J, torch.float32)
    cands = np.stack([O.round_spins(O.normal_vector(prob.n, 4, i)) for i in range(9)])
    Jn, d, c = O.normalize(J, np.float32)
    got = LS.energies_of_bits(prob, _bits(cands), 9)
    np.testing.assert_allclose(got, O.energies_float(Jn, cands.T, c), rtol=1e-12)


@pytest.mark.parametrize("precision", [32, 64])
def test_float_sweep_end_to_end(gpu, tmp_path, monkeypatch, precision):
    """SpecAnn on real-valued couplings: candidates' raw energies are fp32_energies of the rounded
    vectors, the refined ones never exceed them, and solve() reports -1/2 s^T J s of its spins in FP64
    from the input couplings (:483-488)."""
    monkeypatch.chdir(tmp_path)
    J = real_couplings(1500, 14, seed=7)
    sa = SpectralAnnealing("float.txt", "synthetic", random_seed=42, precision=precision, num_alphas=12,
                           verbose=False)
    energy, spins, _ = sa.solve(J)
    s = spins.astype(np.float64)
    assert energy == pytest.approx(-0.5 * s @ (J @ s), rel=1e-12)
    prob = DeviceProblem(J, torch.float32 if precision == 32 else torch.float64)
    out = SpectralAnnealing("float.txt", "synthetic", random_seed=42, precision=precision, num_alphas=12,
                            verbose=False).run_on_device(prob)
    Jn, d, c = O.normalize(J, np.float32 if precision == 32 else np.float64)
    bits = out.cand_bits.cpu().numpy()
    cands = np.stack([unpack_bits(b, prob.n) for b in bits])
    np.testing.assert_allclose(out.raw_energies, O.energies_float(Jn, cands.T, c), rtol=1e-12)
    assert all(r <= w + 1e-9 * abs(w) for r, w in zip(out.refined_energies, out.raw_energies))
    assert out.best_energy <= min(out.raw_energies)


def test_large_integer_couplings_take_float_path(gpu):
    """Integer couplings whose largest absolute degree reaches 2^29 would overflow the int32 fields: they
    take the floating-point search, whose result is checked against the oracle's."""
    J = graphs.erdos_renyi(2000, 10 / 1999, seed=2, kind="pm1")
    rng = np.random.default_rng(1)
    U = sp.triu(J, 1).tocoo()
    w = U.data * rng.integers(1 << 10, 1 << 27, U.nnz)
    A = sp.coo_matrix((w.astype(np.float64), (U.row, U.col)), shape=J.shape)
    J = (A + A.T).tocsr()
    prob = DeviceProblem(J, torch.float32)
    assert prob.c >= 2.0 ** 29 and prob.jint is None
    cands = np.stack([O.round_spins(O.normal_vector(prob.n, 2, i)) for i in range(4)])
    st = LS.flip_refinement(prob, _bits(cands), 4, seed=7)
    Jn, d, c = O.normalize(J, np.float32)
    ref = O.local_search(Jn, cands, 7)
    np.testing.assert_array_equal(st.spins[:, :4].cpu().numpy().T, ref.spins)


def test_inexact_dynamic_range_is_refused(gpu):
    """Couplings spanning more than 2^30 relative to the largest degree would make the FP64 fields
    order-dependent: the local search refuses them with a clear error instead of a racy result."""
    J = graphs.erdos_renyi(500, 10 / 499, seed=3, kind="pm1")
    J = sp.csr_matrix(J.astype(np.float64))
    J.data[:] = 1.0
    J[0, J.indices[J.indptr[0]]] = 2.0 ** 40
    J[J.indices[J.indptr[0]], 0] = 2.0 ** 40
    prob = DeviceProblem(J, torch.float32)
    assert prob.jint is None
    cands = np.stack([O.round_spins(O.normal_vector(prob.n, 2, 0))])
    with pytest.raises(NotImplementedError, match="2\\^-30"):
        LS.flip_refinement(prob, _bits(cands), 1, seed=1)
