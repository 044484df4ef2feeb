"""The C ABI library loads without a GPU and exports every symbol that
include/mft.h declares; the drop-in host library does the same for its own
header.  No compute calls here."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mft_[a-z0-9_]+)\s*\(", text)))


def test_libmft_exports_header_symbols():
    import mft
    lib = mft.lib()
    names = declared(os.path.join(ROOT, "include", "mft.h"))
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_status_codes_without_device_work():
    import mft
    lib = mft.lib()
    assert lib.mft_version() >= 1
    assert lib.mft_fft_supported(128) == 1
    assert lib.mft_fft_supported(100) == 0
    # invalid arguments are rejected before any device work
    assert lib.mft_tild_grid_weights(None, None, None, 0, None, None, 3, None) == mft.MFT_ERR_INVALID
    assert b"null" in lib.mft_last_error()
