"""CPU checks of the HIP library: it loads and exports every C-ABI symbol."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(sw_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    for must in ("sw_subset_check", "sw_transition_apply", "sw_leaf_create", "sw_leaf_run", "sw_leaf_destroy",
                 "sw_count_bits", "sw_exact_solve", "sw_pipeline_solve", "sw_expand"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import synchrowords_amd as sw
    if not sw.LIB_PATH.exists():
        pytest.fail("libsynchrowords_amd.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(str(sw.LIB_PATH))
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_mirror_binds_every_symbol():
    import synchrowords_amd as sw
    assert set(declared_symbols()) <= set(sw._SIGS)


def test_version_and_errors_without_gpu():
    import synchrowords_amd as sw
    lib = sw.load()
    assert lib.sw_version() >= 10000
    h = ctypes.c_void_p()
    rc = lib.sw_ctx_create(12345, ctypes.byref(h))
    assert rc != 0  # no such device here, reported as a status code
    assert lib.sw_last_error()
