"""The C-ABI library loads and exports every symbol include/exactclosure.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import os

from conftest import ROOT


def test_library_exports_header_symbols(ec):
    names = ec._abi.declared_symbols()
    assert len(names) >= 40
    lib = ctypes.CDLL(ec._abi.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.ec_version() == 1


def test_bindings_cover_header(ec):
    names = set(ec._abi.declared_symbols())
    bound = set(ec._abi._SIGS)
    assert names == bound, (names - bound, bound - names)


def test_library_is_gfx950_only(ec):
    # the offload bundle names the target; no other GPU architecture is built
    blob = open(ec._abi.LIB_PATH, "rb").read()
    assert b"gfx950" in blob
    for other in (b"amdgcn-amd-amdhsa--gfx942", b"amdgcn-amd-amdhsa--gfx90a"):
        assert other not in blob


def test_no_gpu_error_is_reported(ec):
    # ec_create on a machine without a device returns an error code instead of crashing
    import torch

    if torch.cuda.is_available():
        return
    h = ctypes.c_void_p()
    rc = ec._abi.lib().ec_create(ctypes.byref(h), 0, None)
    assert rc != 0


def test_null_context_rejected(ec):
    lib = ec._abi.lib()
    assert lib.ec_sync(None) == ec._abi.EC_E_INVALID
    assert lib.ec_last_error(None) == b"null context"
