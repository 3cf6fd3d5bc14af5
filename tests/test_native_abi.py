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


def test_strategy_error_codes_without_gpu(bj):
    """bh_check_strategy: which (operator, quadstrat) pairs have a GPU path, decided on the host.
    Valid strategies the GPU path does not cover are BH_ERR_UNSUPPORTED (the caller falls back to
    its CPU assembly, as the reference's tests need for rules 12 and 13,
    test/test_sauterschwabints.jl:29-30); invalid arguments are BH_ERR_INVALID."""
    import ctypes as C
    nat = bj._native
    lib = nat.load()

    def code(kind, qkind, params):
        op = nat.bh_operator()
        op.kind = kind
        op.gamma[:] = [0.0, 1.0]
        op.alpha[:] = [0.0, -1.0]
        op.beta[:] = [0.0, 1.0]
        qs = nat.bh_quadstrat()
        qs.kind = qkind
        for i, v in enumerate(params):
            qs.params[i] = v
        return lib.bh_check_strategy(C.byref(op), C.byref(qs))

    ok, uns, inv = nat.BH_OK, nat.BH_ERR_UNSUPPORTED, nat.BH_ERR_INVALID
    SL, DL, NC, DLR, ID = (nat.OP_SINGLELAYER, nat.OP_DOUBLELAYER, nat.OP_NCROSS, nat.OP_DOUBLELAYER_ROTATED,
                           nat.OP_IDENTITY)
    WS, SA, DN, SN = nat.QS_WILTON_SAUTER, nat.QS_SAUTER, nat.QS_DOUBLENUM, nat.QS_SINGLENUM
    assert code(SL, WS, (2, 3, 6, 7, 5, 5, 4, 3)) == ok
    assert code(DL, WS, (2, 3, 9, 10, 5, 5, 4, 3)) == ok          # degrees 9 and 10 are tabulated
    assert code(SL, DN, (12, 13)) == uns                           # beyond the tables: CPU fallback
    assert b"degree 12" in lib.bh_last_error()
    assert code(SL, SA, (2, 3, 5, 5, 4, 20)) == uns                 # Gauss order above 16
    assert code(SL, WS, (0, 3, 6, 7, 5, 5, 4, 3)) == inv           # degree 0 is invalid
    assert code(DLR, DN, (6, 7)) == ok and code(DLR, SA, (2, 3, 5, 5, 4, 3)) == ok
    assert code(DLR, WS, (2, 3, 6, 7, 5, 5, 4, 3)) == uns          # no singularity extraction for n x K
    assert code(NC, SN, (6,)) == ok and code(ID, SN, (10,)) == ok
    assert code(ID, SN, (11,)) == uns and code(ID, DN, (2, 3)) == uns and code(SL, SN, (6,)) == uns
    assert code(7, DN, (2, 3)) == inv                               # unknown operator kind
