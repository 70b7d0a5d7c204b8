"""Sanitizer runs of the host code on the CPU.

- AddressSanitizer + UndefinedBehaviorSanitizer builds of librtq_build.so (builder, scene generators,
  blob cache, C API) and of the oracle's liborc.so, loaded into a child pytest that runs the builder,
  cache, oracle-pin and two-level test modules against them (RTQ_BUILD_LIB / RTQ_ORACLE_LIB name the
  builds; the ASan runtime is preloaded into the child because the interpreter is not built with it).
- ThreadSanitizer: the builder's sources linked into a driver that builds a soup and a two-level scene
  on 1 and on 8 threads (identical blobs, no report), and the served form's admission pieces.
"""
import os
import subprocess
import sys

import pytest

from tim_raytracing_amd import build as Bld

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SAN = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-fno-sanitize-recover=undefined"]


def _runtime(name):
    p = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


@pytest.mark.skipif(_runtime("libasan.so") is None, reason="gcc's ASan runtime not installed")
def test_builder_and_oracle_under_asan_ubsan(tmp_path):
    build_lib = Bld.build_host(out=str(tmp_path / "librtq_build.so"), extra=SAN)
    orc_lib = str(tmp_path / "liborc.so")
    subprocess.run(["gcc", "-O1", "-g", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-pthread", *SAN,
                    "-o", orc_lib, os.path.join(ROOT, "oracle", "rtq_oracle.c"), "-lm"], check=True)
    env = dict(os.environ)
    env["RTQ_BUILD_LIB"] = build_lib
    env["RTQ_ORACLE_LIB"] = orc_lib
    env["LD_PRELOAD"] = ":".join(x for x in (_runtime("libasan.so"), env.get("LD_PRELOAD", "")) if x)
    env["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    mods = [os.path.join(HERE, m) for m in ("test_builder.py", "test_cache.py", "test_oracle_pins.py", "test_tlas.py")]
    # the SAT comparison drives two unsanitized libraries only; everything else runs
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", "-m", "not gpu",
                        "-k", "not tribox_matches_reference_sat", "--durations=5", *mods],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=1200)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-6000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-6000:]
    assert " passed" in r.stdout


@pytest.mark.skipif(_runtime("libtsan.so") is None, reason="gcc's TSan runtime not installed")
def test_threaded_builder_under_tsan(tmp_path):
    exe = tmp_path / "builder_threads_test"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-ffp-contract=off", "-fno-fast-math",
                    "-pthread", "-o", str(exe), os.path.join(HERE, "cpp", "builder_threads_test.cpp"), *Bld.HOST_SRC],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=900,
                       env=dict(os.environ, TSAN_OPTIONS="halt_on_error=1"))
    assert r.returncode == 0, r.stderr[-6000:]
    assert r.stdout.strip() == "ok" and "ThreadSanitizer" not in r.stderr


@pytest.mark.skipif(_runtime("libtsan.so") is None, reason="gcc's TSan runtime not installed")
def test_serve_core_under_tsan(tmp_path):
    exe = tmp_path / "serve_core_test"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-pthread", "-o", str(exe),
                    os.path.join(HERE, "cpp", "serve_core_test.cpp")], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, TSAN_OPTIONS="halt_on_error=1"))
    assert r.returncode == 0, r.stderr[-6000:]
    assert r.stdout.strip() == "ok" and "ThreadSanitizer" not in r.stderr
