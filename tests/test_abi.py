"""The C-ABI library loads on a GPU-less host and exports every entry point
declared in include/untangle.h (no compute calls)."""
import ctypes
import pathlib
import re

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]


def declared():
    text = (ROOT / "include" / "untangle.h").read_text()
    return sorted(set(re.findall(r"UT_API[^;(]*?\b(ut_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "ut_ctx_create" in names and "ut_ef_dcd" in names
    assert len(names) >= 15


def test_library_exports_all_declared_symbols():
    from untangling26_amd import _lib

    if not _lib.LIB_PATH.exists():
        pytest.fail(f"{_lib.LIB_PATH} is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding declares a signature for each of them
    assert sorted(_lib.SIGNATURES) == declared()
    assert L.ut_version() == 1


def test_no_device_gives_clean_error():
    import torch

    if torch.cuda.is_available():
        pytest.skip("device present")
    from untangling26_amd import _lib

    h = ctypes.c_void_p()
    rc = _lib.lib().ut_ctx_create(0, None, 0, ctypes.byref(h))
    assert rc < 0 and not h.value


def prototypes():
    """name -> number of parameters of every UT_API prototype in the header."""
    text = (ROOT / "include" / "untangle.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"UT_API[^;(]*?\b(ut_\w+)\s*\(([^)]*)\)\s*;", text, flags=re.S):
        params = m.group(2).strip()
        out[m.group(1)] = 0 if params in ("", "void") else len(params.split(","))
    return out


def test_binding_argument_counts_match_header():
    from untangling26_amd import _lib

    protos = prototypes()
    assert sorted(protos) == declared()
    wrong = {n: (len(_lib.SIGNATURES[n][1]), k) for n, k in protos.items() if len(_lib.SIGNATURES[n][1]) != k}
    assert not wrong, wrong
