"""CPU test of upload-BVH's host half: blob decode and device layout (scene_layout.h), no GPU needed."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.mark.parametrize("top_slots", [None, 0, 32, 128])
def test_scene_layout(tmp_path, top_slots):
    """None: the library's own setting of RTQ_TOP_SLOTS; the others lay out that many top records
    breadth first."""
    exe = tmp_path / "scene_layout_test"
    pkg = os.path.join(ROOT, "tim_raytracing_amd")
    define = [] if top_slots is None else [f"-DRTQ_TOP_SLOTS={top_slots}"]
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-Wall", *define, "-Xlinker", "--no-as-needed",
                    "-fsanitize=address,undefined", "-o", str(exe), os.path.join(HERE, "cpp", "scene_layout_test.cpp"),
                    os.path.join(pkg, "librtq_build.so"), f"-Wl,-rpath,{pkg}"], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    assert r.stdout.strip() == "ok"
