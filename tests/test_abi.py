"""The C-ABI libraries load without a GPU and export every function include/*.h declares."""
import ctypes as C
import os
import re

import pytest

from tim_raytracing_amd import bindings as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(rtq_\w+)\s*\(", text, flags=re.M)))


def test_headers_declare_functions():
    assert len(declared("rtq.h")) >= 15
    assert len(declared("rtq_build.h")) >= 15


def test_build_library_exports_all():
    lib = C.CDLL(B.BUILD_LIB)
    missing = [f for f in declared("rtq_build.h") if not hasattr(lib, f)]
    assert not missing


@pytest.mark.skipif(not os.path.exists(B.DEVICE_LIB), reason="librtq.so not built")
def test_device_library_exports_all():
    lib = C.CDLL(B.DEVICE_LIB)
    missing = [f for f in declared("rtq.h") if not hasattr(lib, f)]
    assert not missing
    # status strings need no device: every code of the header's enum has its own name
    lib.rtq_status_string.restype = C.c_char_p
    text = open(os.path.join(ROOT, "include", "rtq.h")).read()
    enum = text[text.index("typedef enum"):text.index("} rtq_status;")]
    codes = {int(v): k for k, v in re.findall(r"\b(RTQ_(?:OK|ERR_\w+))\s*=\s*(\d+)", enum)}
    assert len(codes) >= 13 and codes[0] == "RTQ_OK"
    for value, name in codes.items():
        assert lib.rtq_status_string(value) == name.encode(), name
        assert B.STATUS_NAMES[value] == name
    assert lib.rtq_status_string(999) == b"RTQ_ERR_UNKNOWN"


def test_bindings_cover_headers():
    assert set(declared("rtq.h")) <= set(B._DEVICE_SIGS)
    assert set(declared("rtq_build.h")) <= set(B._BUILD_SIGS)
