"""CPU: libwakis_fit.so loads and exports every entry point declared in include/wakis_fit.h."""

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "wakis_fit.h"


def declared():
    txt = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(wf_\w+)\s*\(", txt, re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "wf_step" in names and "wf_create" in names and len(names) >= 10


def test_library_exports_all_declared():
    from wakis_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.fail("libwakis_fit.so not built (python -m wakis_amd.build)")
    so = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared() if not hasattr(so, n)]
    assert not missing, f"missing exports: {missing}"
    assert set(_lib.EXPORTS) == set(declared()), "ctypes binding out of sync with the header"


def test_version_and_error_string_without_gpu():
    from wakis_amd import _lib
    L = _lib.lib()
    assert L.wf_version() >= 1
    # an argument error is reported without touching the GPU
    rc = L.wf_create(None, None)
    assert rc == _lib.WF_EINVAL
    assert b"null" in L.wf_last_error()
