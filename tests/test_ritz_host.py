"""The dense Ritz-vector routine of the Lanczos finisher (sa_ritz.hpp), built for the host and checked
against numpy's eigh on projected matrices of the solve's shape: tridiagonal, arrowhead after a thick
restart, and the cases the derivative formula gets wrong (a start or kept vector that is already an
eigenvector to rounding, so that theta equals a diagonal entry). The routine is shared by a team of
threads on the device; the team form runs here on host threads meeting at a barrier and must give
the same bits as one thread."""
import ctypes
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    out = tmp_path_factory.mktemp("ritz") / "libritz.so"
    cmd = [hipcc, "-O2", "-shared", "-fPIC", "-std=c++20", "-pthread", f"-I{ROOT / 'specann_amd' / 'csrc'}",
           str(ROOT / "tests" / "native" / "dense_ritz_host.cpp"), "-o", str(out)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        pytest.fail(f"host build failed:\n{res.stderr}")
    L = ctypes.CDLL(str(out))
    L.dense_ritz_host.restype = ctypes.c_int
    L.dense_ritz_team.restype = ctypes.c_int
    return L


def _dense(m, k, a, b, ar):
    P = np.diag(a.copy())
    for i in range(k):
        P[i, k] = P[k, i] = ar[i]
    for i in range(k, m - 1):
        P[i, i + 1] = P[i + 1, i] = b[i]
    return P


def _ritz(lib, m, k, a, b, ar, theta, threads=0):
    x = np.zeros(m)
    dp = ctypes.POINTER(ctypes.c_double)
    tiny = 2.220446049250313e-16 * max(np.abs(_dense(m, k, a, b, ar)).sum(axis=1).max(), 1e-300)
    args = (m, k, a.ctypes.data_as(dp), b.ctypes.data_as(dp), ar.ctypes.data_as(dp), ctypes.c_double(theta),
            ctypes.c_double(tiny), x.ctypes.data_as(dp))
    rc = lib.dense_ritz_team(threads, *args) if threads else lib.dense_ritz_host(*args)
    assert rc == 0
    return x


CASES = ["tridiagonal", "arrowhead", "kept_converged", "kept_exact", "start_converged", "clustered", "pivoting"]


def _case(case, seed, m_max=48):
    rng = np.random.default_rng(seed)
    m = int(rng.integers(8, m_max + 1))
    k = 0 if case in ("tridiagonal", "start_converged") else int(rng.integers(1, min(16, m - 2)))
    a = rng.uniform(-0.3, 0.3, m)
    b = rng.uniform(0.05, 0.2, m)
    ar = rng.uniform(-0.1, 0.1, m)
    if case == "kept_converged":
        a[0] = 0.9                      # the top kept Ritz value, already converged to 1e-12
        ar[0] = 1e-12
    if case == "kept_exact":
        a[0] = 0.9                      # a kept Ritz vector that is the eigenvector: theta == a[0]
        ar[0] = 0.0
    if case == "start_converged":
        a[0] = 0.9                      # the start vector is the eigenvector to 1e-11
        b[0] = 1e-11
    if case == "pivoting":
        # near-Toeplitz chain and arrow entries larger than the shifted diagonal: partial pivoting
        # interchanges rows at many columns, including rows that already hold multipliers
        a = rng.uniform(-0.01, 0.01, m)
        b = rng.uniform(0.5, 1.0, m)
        ar = rng.uniform(-2.0, 2.0, m)
    if case == "clustered":
        a[:k] = 0.5 + rng.uniform(-1e-9, 1e-9, k)
        ar[:k] *= 1e-6
    return m, k, a, b, ar


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("seed", range(4))
def test_dense_ritz_vector_matches_eigh(lib, case, seed):
    m, k, a, b, ar = _case(case, seed, m_max=128 if seed % 2 else 48)
    P = _dense(m, k, a, b, ar)
    w, V = np.linalg.eigh(P)
    theta, v = w[-1], V[:, -1]
    x = _ritz(lib, m, k, a, b, ar, theta)
    x = x if x @ v >= 0 else -x
    assert np.linalg.norm(x) == pytest.approx(1.0, abs=1e-14)
    gap = w[-1] - w[-2]
    assert np.linalg.norm(x - v) <= 1e-13 / min(1.0, gap) + 1e-14
    # the stopping estimate |beta s_last| needs the last component to absolute accuracy
    assert abs(abs(x[-1]) - abs(v[-1])) <= 1e-14 / min(1.0, gap) + 1e-16
    # the residual of the computed vector in the projected matrix
    assert np.linalg.norm(P @ x - theta * x) <= 1e-14


@pytest.mark.parametrize("threads", [2, 7, 64, 256])
@pytest.mark.parametrize("case", ["arrowhead", "pivoting", "kept_exact"])
def test_team_form_gives_the_serial_bits(lib, case, threads):
    for seed in range(3):
        m, k, a, b, ar = _case(case, 10 + seed, m_max=128)
        theta = np.linalg.eigvalsh(_dense(m, k, a, b, ar))[-1]
        serial = _ritz(lib, m, k, a, b, ar, theta)
        team = _ritz(lib, m, k, a, b, ar, theta, threads=threads)
        assert np.array_equal(serial, team), (case, threads, seed)
