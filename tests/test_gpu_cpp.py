"""Runs the C++ drop-in test programs built by __graft_entry__.build()."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("name", ["test_templates", "test_templates_clang"])
def test_cpp_program(name):
    exe = os.path.join(ROOT, "tests", "cpp", "_build", name)
    assert os.path.exists(exe), f"{exe} not built (run __graft_entry__.build())"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
