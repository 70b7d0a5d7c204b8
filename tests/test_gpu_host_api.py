"""The C++ host interface (include/rtq.hpp) end to end on the GPU, checked against the oracle."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.gpu


def test_cpp_host_api(tmp_path):
    print("case: C++ host API (BVHBuilder/BVHData/RayTracingPass) 320x240 vs oracle", flush=True)
    exe = tmp_path / "host_api_test"
    pkg = os.path.join(ROOT, "tim_raytracing_amd")
    orc = os.path.join(ROOT, "oracle")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), "-o", str(exe),
                    os.path.join(ROOT, "tests", "cpp", "host_api_test.cpp"),
                    os.path.join(pkg, "librtq.so"), os.path.join(pkg, "librtq_build.so"), os.path.join(orc, "liborc.so"),
                    f"-Wl,-rpath,{pkg}:{orc}"], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "ok"
