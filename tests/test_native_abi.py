"""The C ABI library loads and exports every function include/beast_hip.h declares (CPU, no compute)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "beast_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(bh_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = _declared()
    for must in ("bh_assemble", "bh_classify", "bh_matrix_factor", "bh_matrix_solve", "bh_assemble_planewave"):
        assert must in names


def test_library_exports_all(bj):
    lib = ctypes.CDLL(bj._native.LIBPATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding declares exactly the same set
    assert sorted(bj._native.EXPORTED) == _declared()


def test_version_and_error_without_gpu(bj):
    lib = bj._native.load()
    assert lib.bh_version() == 100
    assert isinstance(lib.bh_last_error(), bytes)
    # argument validation fails before any device work
    assert lib.bh_ctx_create(0, None) == bj._native.BH_ERR_INVALID
