"""The C ABI: the built library loads on the CPU host and exports every entry point the header
declares, with the binding's struct layouts (no GPU needed; no compute calls)."""
import ctypes
import re
from pathlib import Path

import pytest

from specann_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def header_functions():
    text = (ROOT / "include" / "specann.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|size_t|int64_t|void)\s+(sa_\w+)\s*\(", text, flags=re.M)))


def test_header_and_binding_agree():
    assert header_functions() == sorted(_lib.exported_names())


def test_library_exports_every_declared_symbol():
    if not _lib.LIB_PATH.exists():
        pytest.fail("libspecann.so is not built (python -m specann_amd.build)")
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [name for name in header_functions() if not hasattr(lib, name)]
    assert not missing, f"missing exports: {missing}"


def test_struct_layouts_match_library():
    lib = _lib.load()
    assert lib.sa_state_bytes() == ctypes.sizeof(_lib.SaState)
    assert lib.sa_minres_state_bytes() == ctypes.sizeof(_lib.SaMinresState)
    assert ctypes.sizeof(_lib.SaMinresState) <= ctypes.sizeof(_lib.SaState)   # shares the work-buffer layout
    assert lib.sa_work_bytes(1000) >= ctypes.sizeof(_lib.SaState)
    assert lib.sa_version() >= 100


def test_status_codes_match_header():
    text = (ROOT / "include" / "specann.h").read_text()
    for name in ("SA_OK", "SA_E_ARG", "SA_E_LAUNCH", "SA_E_BREAKDOWN", "SA_E_OOM", "SA_RUNNING", "SA_CONVERGED",
                 "SA_MAXITER", "SA_RESTART", "SA_REJECTED", "SA_BROKEN", "SA_MAX_NCV"):
        m = re.search(rf"#define {name} \(?(-?\d+)\)?", text)
        assert m, name
        assert int(m.group(1)) == getattr(_lib, name), name


def test_argument_errors_are_reported_without_a_gpu():
    lib = _lib.load()
    rc = lib.sa_csr_tiles(0, None, 2048, None, 1, None)
    assert rc == _lib.SA_E_ARG
    assert "bad arguments" in _lib.last_error()


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(_lib.SpecAnnLibraryError):
        _lib.load(tmp_path / "libspecann.so")
