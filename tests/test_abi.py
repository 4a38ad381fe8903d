"""C-ABI checks that need no GPU: the library loads, exports every symbol the
headers declare, and rejects bad arguments with status codes."""
import ctypes
import glob
import os
import re

import pytest

from conftest import REPO


def header_symbols():
    names = set()
    for h in glob.glob(os.path.join(REPO, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(edpp_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_entry_points():
    syms = header_symbols()
    assert "edpp_gather_i32" in syms and "edpp_split_i32" in syms and len(syms) >= 20


def test_library_exports_every_declared_symbol(edpp):
    L = edpp.lib()
    missing = [s for s in header_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_binding_matches_headers(edpp):
    # every symbol the Python binding declares exists in a header
    assert set(edpp.declared_symbols()) <= set(header_symbols())


def test_abi_version_and_strerror(edpp):
    L = edpp.lib()
    assert L.edpp_abi_version() == 1
    assert L.edpp_strerror(edpp.EINVAL) == b"invalid argument"
    assert L.edpp_strerror(edpp.ETIMEOUT) == b"device wait timed out"


def test_null_context_is_rejected(edpp):
    L = edpp.lib()
    ms = ctypes.c_double()
    assert L.edpp_gather_i32(None, None, None, None, 10, 1, ctypes.byref(ms)) == edpp.EINVAL
    assert L.edpp_scan_exclusive_i32(None, None, None, 10, 0, None, None) == edpp.EINVAL
    assert L.edpp_split_i32(None, None, None, None, None, None, 10, 16, 0, None) == edpp.EINVAL
    assert L.edpp_ctx_create(0, None, None) == edpp.EINVAL
    assert L.edpp_ctx_destroy(None) == edpp.OK


def test_context_without_gpu(edpp):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    assert edpp.lib().edpp_ctx_create(0, None, ctypes.byref(h)) == edpp.ENODEV
