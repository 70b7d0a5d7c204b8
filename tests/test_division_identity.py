"""CPU check of the fp64 route to the IEEE fp32 slab quotient (tests/cpp/division_identity_test.c)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_fp64_quotient_equals_ieee_fp32_division(tmp_path):
    exe = tmp_path / "division_identity_test"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-o", str(exe),
                    os.path.join(HERE, "cpp", "division_identity_test.c"), "-lm"], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok ")
