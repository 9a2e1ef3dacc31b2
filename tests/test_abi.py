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


def test_tuning_record_matches_the_header():
    """records.TUNING (what the binding passes to rgrt_set_tuning) has the
    header's rgrt_tuning fields, in order, all 32-bit unsigned."""
    from raygun_amd.records import TUNING
    text = open(os.path.join(ROOT, "include", "rgrt.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    body = re.search(r"typedef struct rgrt_tuning\s*\{(.*?)\}\s*rgrt_tuning;", text, flags=re.S).group(1)
    fields = re.findall(r"uint32_t\s+(\w+)\s*;", body)
    assert fields == list(TUNING.names)
    assert TUNING.itemsize == 4 * len(fields)
