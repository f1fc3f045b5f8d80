"""CPU-side checks of the C ABI: the library loads and exports every declared entry point."""
import os
import re

import pytest

from apw import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "apw.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(apw_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for need in ("apw_hpsi", "apw_david_diag", "apw_david_create", "apw_fft3d", "apw_zgemm", "apw_dnevx_op",
                 "apw_ekinetic_pw_op", "apw_nonlocal_pw_op", "apw_set_real_to_recip_output_op"):
        assert need in syms


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libapw.so not built")
    lib = _lib.load()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    # every declared symbol has a ctypes prototype
    assert not [s for s in declared_symbols() if s not in _lib.PROTOTYPES]


def test_error_path_without_gpu_calls():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libapw.so not built")
    import ctypes as C
    lib = _lib.load()
    # null ctx: argument check fires before any HIP call
    rc = lib.apw_ctx_synchronize(None)
    assert rc == -1
    assert b"null" in lib.apw_last_error()
    rc = lib.apw_hpsi(None, None, 0, None, 0, 1)
    assert rc == -1
