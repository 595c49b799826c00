"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer: `make -C oracle sanitize`
builds oracle/sanitize_main.cpp (every oracle entry point on small 2D/3D meshes, p = 1..4: both
operators, boundary conditions, blending, hooks, the steppers, the extended-precision operator)
with both sanitizers, any report fatal, and runs it."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None, reason="needs g++ and make")
def test_oracle_clean_under_asan_ubsan():
    r = subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    print(out[-2000:])
    assert r.returncode == 0, out[-4000:]
    assert out.count("ok: 0 failed checks") == 2
    assert "runtime error" not in out and "AddressSanitizer" not in out
