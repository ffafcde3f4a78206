"""The persistent chunked pass kernel (RSORT_AMD_PASS_KERNEL=chunk, an
alternative to the per-tile kernel kept for comparison) against the oracle,
in a process of its own because the library reads the variable once."""
from __future__ import annotations

import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu


def test_chunk_kernel_parity():
    here = os.path.dirname(os.path.abspath(__file__))
    env = dict(os.environ, RSORT_AMD_PASS_KERNEL="chunk")
    r = subprocess.run([sys.executable, "-u", os.path.join(here, "chunk_kernel_check.py")], env=env,
                       capture_output=True, text=True, timeout=240)
    print(r.stdout[-3000:], r.stderr[-3000:])
    assert r.returncode == 0 and "chunk kernel parity ok" in r.stdout
