"""Served requests: host buffers are validated before their addresses reach the library (no GPU)."""
import numpy as np
import pytest

from tim_raytracing_amd import bindings as B
from tim_raytracing_amd.device import check_request_buffers


def test_valid_buffers_pass():
    rays = np.zeros(10, B.RAY_DTYPE)
    assert check_request_buffers(B.RTQ_TRACE_CLOSEST, rays, np.zeros(10, B.HIT_DTYPE), np.zeros(10, np.uint32)) == 10
    assert check_request_buffers(B.RTQ_TRACE_CLOSEST, rays, np.zeros(12, B.HIT_DTYPE), None) == 10
    assert check_request_buffers(B.RTQ_TRACE_ANY, rays, np.zeros(10, np.uint32), None) == 10


@pytest.mark.parametrize("case", ["float_rays", "2d_rays", "strided_rays", "empty", "short_out", "any_out_hits",
                                  "closest_out_u32", "short_nid", "nid_dtype", "nid_for_any", "readonly_out",
                                  "bad_flags"])
def test_bad_buffers_raise(case):
    rays = np.zeros(10, B.RAY_DTYPE)
    out = np.zeros(10, B.HIT_DTYPE)
    nid = None
    flags = B.RTQ_TRACE_CLOSEST
    if case == "float_rays":
        rays = np.zeros((10, 8), np.float32)  # 8x too many bytes would be read as records
    elif case == "2d_rays":
        rays = np.zeros((2, 5), B.RAY_DTYPE)
    elif case == "strided_rays":
        rays = np.zeros(20, B.RAY_DTYPE)[::2]
    elif case == "empty":
        rays = np.zeros(0, B.RAY_DTYPE)
    elif case == "short_out":
        out = np.zeros(9, B.HIT_DTYPE)
    elif case == "any_out_hits":
        flags = B.RTQ_TRACE_ANY
    elif case == "closest_out_u32":
        out = np.zeros(10, np.uint32)
    elif case == "short_nid":
        nid = np.zeros(9, np.uint32)
    elif case == "nid_dtype":
        nid = np.zeros(10, np.int64)
    elif case == "nid_for_any":
        flags = B.RTQ_TRACE_ANY
        out = np.zeros(10, np.uint32)
        nid = np.zeros(10, np.uint32)
    elif case == "readonly_out":
        out.setflags(write=False)
    elif case == "bad_flags":
        flags = 7
    with pytest.raises(ValueError):
        check_request_buffers(flags, rays, out, nid)
