"""The C-ABI library loads and exports every symbol include/rgrt.h declares.

CPU only: no compute call is made (rgrt_create is called only to check that a
missing GPU is reported as a status, not a crash)."""
import ctypes
import os
import re

import pytest

from raygun_amd import rgrt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "rgrt.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)  # drop comments
    return sorted(set(re.findall(r"\b(rgrt_[a-z_0-9]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert header_symbols() == sorted(rgrt.EXPORTS)


def test_library_exports_every_symbol():
    if not os.path.exists(rgrt.LIB_PATH):
        pytest.skip("librgrt.so not built")
    L = ctypes.CDLL(rgrt.LIB_PATH)
    for name in header_symbols():
        assert hasattr(L, name), name
    assert rgrt.version().startswith("rgrt")


def test_create_without_device_reports_status():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    if not os.path.exists(rgrt.LIB_PATH):
        pytest.skip("librgrt.so not built")
    with pytest.raises(rgrt.RgrtError) as e:
        rgrt.Context(16, 16)
    assert e.value.status == 5
