"""The host copy threads of the bulk-copy staging (csrc/gdp_copy_pool.h),
on the CPU: started and stopped many times with copies in between (the
staging's life follows the problems'), every copy checked.  A pool
restarted after a stop once took the previous copy's stale job and
deadlocked its next copy; this runs that sequence under a time limit."""
import os
import subprocess

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_copy_pool_restarts(tmp_path):
    exe = tmp_path / "copy_pool_driver"
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-I", os.path.join(REPO, "gpu-domain-propagator_amd", "csrc"),
                    "-o", str(exe), os.path.join(REPO, "tests", "native", "copy_pool_driver.cpp")], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout + out.stderr
