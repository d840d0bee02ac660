"""libdcc.so loads and exports every entry point include/dcc.h declares (no GPU calls)."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "dcc.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char \*)\s*(dcc_\w+)\s*\(", txt, flags=re.M)))


def test_header_declares_entry_points():
    names = _declared()
    assert "dcc_contract_f64" in names and "dcc_r3_finalize_f64" in names and len(names) >= 20


def test_library_exports_declared_symbols():
    import ctypes
    from distr_cc import _lib as L
    lib = ctypes.CDLL(L._LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    # every declared function is also bound by the Python loader
    assert not [n for n in _declared() if n not in L.EXPORTED]


def test_abi_version_and_error_string():
    from distr_cc import _lib as L
    assert L._lib.dcc_abi_version() == 1
    assert isinstance(L._lib.dcc_last_error(), bytes)
