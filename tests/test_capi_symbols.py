"""The C-ABI library loads and exports every entry point include/gpunufft_hip.h
declares (no compute calls: CPU only)."""

import ctypes
import os
import re

from conftest import ROOT


def _declared():
    with open(os.path.join(ROOT, "include", "gpunufft_hip.h")) as f:
        src = f.read()
    return sorted(set(re.findall(r"\b(gnufft_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_all_declared_symbols():
    from gpunufft_amd import capi
    declared = _declared()
    assert len(declared) >= 20
    assert set(declared) == set(capi.EXPORTS)
    lib = ctypes.CDLL(capi.LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name
    assert capi.load().gnufft_abi_version() == 1


def test_host_only_calls():
    from gpunufft_amd import capi
    L = capi.load()
    # argument validation happens before any device work
    h = ctypes.c_void_p()
    rc = L.gnufft_plan_create(ctypes.byref(h), (ctypes.c_uint32 * 3)(0, 0, 0), 3, 8, 2.0, 0, None)
    assert rc == 1 and h.value is None
    assert b"width" in L.gnufft_last_error()
    rc = L.gnufft_plan_create(ctypes.byref(h), (ctypes.c_uint32 * 3)(8, 8, 0), 9, 8, 2.0, 0, None)
    assert rc == 5
    rc = L.gnufft_adjoint(None, None, None, 1, 2, None)
    assert rc == 1
