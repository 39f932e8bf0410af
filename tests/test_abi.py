"""CPU checks of the C-ABI boundary: the library loads and exports every declared entry point."""
import ctypes
import os
import re

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
