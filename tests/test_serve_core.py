"""CPU test of the served form's admission pieces (no GPU needed)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_serve_core(tmp_path):
    exe = tmp_path / "serve_core_test"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Xlinker", "--no-as-needed",
                    "-fsanitize=address,undefined", "-o", str(exe), os.path.join(HERE, "cpp", "serve_core_test.cpp")],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "ok"
