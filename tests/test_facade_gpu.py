"""The gtd::system-shaped C++ facade (include/gtd/nbody.hpp) driving the GPU:
a reference-style driver (orbit, predicate-driven evolve, Barnes-Hut, error
types) built from gregpackagescpp_amd/examples/facade_demo.cpp."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_facade_demo_runs():
    exe = os.path.join(ROOT, "gregpackagescpp_amd", "facade_demo")
    assert os.path.exists(exe), "build the package first (make -C gregpackagescpp_amd)"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "facade demo OK" in r.stdout
