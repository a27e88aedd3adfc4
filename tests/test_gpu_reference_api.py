"""The reference's inner callables with their own signatures, computed on the GPU, against the oracle:
diagonal_predictor / secant_predictor (utils/predictors.py:6-56), flip_refinement(J, spins_list, ...)
(models/spectral_annealing.py:1039-1112) and fp32_energies(J, spins, normalization) (:962-976)."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from oracle import specann_oracle as O
from specann_amd import graphs
from specann_amd.device import DeviceProblem
from specann_amd.predictors import diagonal_predictor, secant_predictor
from specann_amd.spectral_annealing import flip_refinement, fp32_energies

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype,tol", [(np.float32, 2e-6), (np.float64, 1e-13)])
def test_diagonal_predictor(gpu, dtype, tol):
    J = graphs.barabasi_albert(3000, 4, seed=2, kind="pm1")
    Jn, d, c = O.normalize(J, dtype)
    y = O.normal_vector(J.shape[0], 5, 0).astype(dtype)
    y /= np.linalg.norm(y)
    got = diagonal_predictor(torch.from_numpy(y).cuda(), torch.from_numpy(d).cuda(), 0.07, dtype)
    assert got.dtype == (torch.float32 if dtype == np.float32 else torch.float64)
    ref = O.diagonal_predictor(y, d, 0.07)
    assert np.linalg.norm(got.double().cpu().numpy() - ref) <= tol
    # a zero vector falls back to a normalised normal vector (utils/predictors.py:22-25)
    z = diagonal_predictor(np.zeros_like(y), d, 0.07, dtype, seed=3)
    assert abs(float(torch.linalg.norm(z.double())) - 1.0) <= 1e-6


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_secant_predictor(gpu, dtype):
    rng = np.random.default_rng(1)
    yl = rng.standard_normal(5000).astype(dtype)
    yl /= np.linalg.norm(yl)
    yp = (yl + 0.05 * rng.standard_normal(5000)).astype(dtype)
    yp /= np.linalg.norm(yp)
    tol = 2e-6 if dtype == np.float32 else 1e-13
    for sign in (1.0, -1.0):
        for gamma in (1.0, 0.5):
            got = secant_predictor(yl, (sign * yp).astype(dtype), dtype, gamma=gamma).double().cpu().numpy()
            ref = O.secant_predictor(yl, (sign * yp).astype(dtype), gamma=gamma)
            assert np.linalg.norm(got - ref) <= tol
    assert np.array_equal(secant_predictor(yl, None, dtype).cpu().numpy(), yl)
    assert np.array_equal(secant_predictor(yl, yl.copy(), dtype).cpu().numpy(), yl)


@pytest.mark.parametrize("kind", ["pm1", "real"])
def test_flip_refinement_reference_signature(gpu, kind):
    J = graphs.erdos_renyi(4000, 12 / 3999, seed=4, kind="pm1")
    if kind == "real":
        U = sp.triu(J, 1).tocoo()
        w = U.data * np.random.default_rng(3).uniform(0.5, 1.5, U.nnz)
        A = sp.coo_matrix((w, (U.row, U.col)), shape=J.shape)
        J = (A + A.T).tocsr()
    spins_list = [O.round_spins(O.normal_vector(J.shape[0], 8, i)) for i in range(5)]
    got, passes = flip_refinement(J, [torch.from_numpy(s).cuda() for s in spins_list], seed=11)
    assert tuple(got.shape) == (J.shape[0], 5) and got.dtype == torch.int8
    Jref = O.integer_couplings(J) if kind == "pm1" else O.normalize(J, np.float32)[0]
    ref = O.local_search(Jref, np.stack(spins_list), 11)
    np.testing.assert_array_equal(got.cpu().numpy().T, ref.spins)
    assert passes == ref.passes
    # spins_list is not modified, and host arrays are taken too
    again, _ = flip_refinement(J, spins_list, seed=11)
    assert torch.equal(again, got)


def test_fp32_energies_reference_signature(gpu):
    J = graphs.erdos_renyi(3000, 10 / 2999, seed=6, kind="pm1")
    S = np.stack([O.round_spins(O.normal_vector(3000, 2, i)) for i in range(7)], axis=1)     # n x k
    e = fp32_energies(J, torch.from_numpy(S).cuda(), 1.0)
    Jint = O.integer_couplings(J)
    assert e.dtype == torch.float64 and e.shape == (7,)
    assert [float(v) for v in e.cpu()] == [float(O.energy_int64(Jint, S[:, c])) for c in range(7)]
    one = fp32_energies(J, S[:, 0], 1.0)                                                  # a single vector
    assert float(one[0]) == float(e[0])
    # the reference's call: the normalised matrix and c
    prob = DeviceProblem(J, torch.float32)
    Jn, d, c = O.normalize(J, np.float32)
    en = fp32_energies(prob, S, prob.c).cpu().numpy()
    np.testing.assert_allclose(en, O.energies_float(Jn, S, c), rtol=1e-12)
