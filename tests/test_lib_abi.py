"""CPU checks of the C ABI: the library loads and exports every function the
header declares, with the signatures the ctypes layer binds."""
import ctypes
import os
import re

from tests.conftest import ROOT

HEADER = os.path.join(ROOT, "include", "cnpc.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cnpc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from cosmonpc_amd import _lib

    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert "cnpc_paint" in names and "cnpc_contract_gemm" in names
    for name in names:
        assert hasattr(lib, name), name


def test_ctypes_binding_covers_header():
    from cosmonpc_amd import _lib

    declared = set(_declared()) - {"cnpc_last_error"}
    assert declared == set(_lib.SIGNATURES), declared ^ set(_lib.SIGNATURES)
    lib = _lib.load()
    assert lib.cnpc_version() == 1
    assert lib.cnpc_last_error() is not None


def test_poly_struct_layout_matches_header():
    from cosmonpc_amd._lib import cnpc_poly

    # 4 int32 + 9 + 2 + 2 + 3 doubles
    assert ctypes.sizeof(cnpc_poly) == 16 + 16 * 8


def test_error_path_without_gpu():
    """A context cannot be created without a device; the call reports a
    status and a message instead of crashing."""
    import torch

    if torch.cuda.is_available():
        return
    from cosmonpc_amd import _lib

    lib = _lib.load()
    h = ctypes.c_void_p()
    rc = lib.cnpc_ctx_create(ctypes.byref(h), 0)
    assert rc != 0
    assert len(lib.cnpc_last_error()) > 0
