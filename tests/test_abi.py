"""CPU checks of the C-ABI boundary: the library loads and exports every declared entry point."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?\w+\s*\**\s*(cs_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    import cosmicscanner_amd as csa
    lib = ctypes.CDLL(csa.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_error_path_without_gpu_calls():
    import cosmicscanner_amd as csa
    L = csa.lib()
    assert L.cs_version() >= 100
    # null handle -> invalid argument, message set, no device work involved
    assert L.cs_kinship_rows(None, None) == csa.CS_ERR_INVALID_ARGUMENT
    assert b"null" in L.cs_last_error()
    assert L.cs_kinship_compact_rows(None) == csa.CS_ERR_INVALID_ARGUMENT
    assert b"cs_kinship_compact_rows" in L.cs_last_error()
    b = ctypes.c_int64()
    assert L.cs_kinship_device_bytes(None, ctypes.byref(b)) == csa.CS_ERR_INVALID_ARGUMENT


def test_dense_eigen_argument_checks_without_gpu_calls():
    """cs_sym_eigen_device validates its arguments before any device call: the rocSOLVER route
    (leaf 0) is limited to n <= 32766, the 64-bit route takes leaf sizes 1..32766."""
    import cosmicscanner_amd as csa
    L = csa.lib()
    fake = ctypes.c_void_p(16)
    assert L.cs_sym_eigen_device(None, 10, 10, None, 0) == csa.CS_ERR_INVALID_ARGUMENT
    assert L.cs_sym_eigen_device(fake, 10, 9, fake, 0) == csa.CS_ERR_INVALID_ARGUMENT
    assert L.cs_sym_eigen_device(fake, 10, 10, fake, -1) == csa.CS_ERR_INVALID_ARGUMENT
    assert L.cs_sym_eigen_device(fake, 10, 10, fake, 40000) == csa.CS_ERR_INVALID_ARGUMENT
    assert b"leaf" in L.cs_last_error()
    assert L.cs_sym_eigen_device(fake, 0, 0, fake, 0) == 0     # empty: nothing to do


def test_validate_variance_components():
    # validation.h:351-366 through the C-ABI (no device work): finite non-negative components
    # pass; the first NaN/Inf or negative one fails with the reference's report code and context
    import cosmicscanner_amd as csa
    L = csa.lib()

    def run(vals):
        a = np.asarray(vals, dtype=np.float64)
        return L.cs_validate_variance_components(a.ctypes.data_as(csa._dp), a.size, b"CosmicScanner null model")

    assert run([0.4, 0.6]) == csa.CS_OK and run([0.0, 0.0]) == csa.CS_OK
    assert run([-1e-12, 0.6]) == csa.CS_ERR_RUNTIME
    msg = L.cs_last_error().decode()
    assert "Validation : FAIL" in msg and "[VARIANCE_NEGATIVE] variance component is negative" in msg
    assert msg.rstrip().endswith("(CosmicScanner null model)")
    for bad in ([np.nan, 0.5], [0.5, np.inf]):
        assert run(bad) == csa.CS_ERR_RUNTIME
        assert "[VARIANCE_NONFINITE] variance component is NaN or Inf" in L.cs_last_error().decode()
    assert run([np.nan, -1.0]) == csa.CS_ERR_RUNTIME and "NONFINITE" in L.cs_last_error().decode()
