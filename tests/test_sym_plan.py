"""The work plan of the symmetric pair sum (solaris_amd/csrc/sym_plan.h),
checked on the host: every unordered pair of groups in exactly one item,
wave ranges and the item -> wave map consistent, column records within
their slots, and the epilogue's lookup of a column finding exactly the
records written for it (tests/cpp/sym_plan_check.cpp, built with g++)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_sym_plan_host_check(tmp_path):
    exe = tmp_path / "sym_plan_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(exe), os.path.join(HERE, "cpp", "sym_plan_check.cpp")],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "sym plan: ok" in r.stdout
