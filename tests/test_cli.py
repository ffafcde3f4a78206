"""The benchmark / validation driver cuda-fast-sort_amd/cli/radix_gpu, with the
reference driver's options (benchmark/bench_parser.cpp:52-106): argument
handling on the CPU; on the GPU the validation matrix (every key type x value
type x mode x order, sizes 10, 100, 1000 and 2^10 .. 2^20, then 2^17 .. 2^27)
and a short benchmark run, all checked on the device."""
from __future__ import annotations

import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cuda-fast-sort_amd", "cli", "radix_gpu")


def run(*args, timeout=60):
    assert os.path.exists(EXE), "radix_gpu is not built (cuda-fast-sort_amd/build.py)"
    return subprocess.run([EXE, *args], capture_output=True, text=True, timeout=timeout)


def test_help():
    r = run("--help")
    assert r.returncode == 0
    for flag in ("--n", "--iterations", "--warmup", "--warmup_ms", "--descending", "--validation"):
        assert flag in r.stdout


@pytest.mark.parametrize("args", [["--bogus"], ["--n"], ["--n", "x"], ["--n", "0"], ["--iterations", "-1"],
                                  ["--warmup", "4294967296"], ["--warmup_ms"]])
def test_bad_arguments(args):
    r = run(*args)
    assert r.returncode == 1
    assert r.stdout.startswith("Error:")
    assert "Usage:" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("log2n", [20, 27])
def test_validation_matrix(log2n):
    """The reference's matrix up to 2^20, and up to its default size 2^27
    (sizes 10, 100, 1000, 2^17 .. 2^27), the size of the reference's own
    validation log (data/logs/validation_log.txt:6557)."""
    r = run("--validation", "--n", str(1 << log2n), timeout=600)
    print(r.stdout[-3000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    m = re.search(r"Done\.\.\. passed (\d+)/(\d+) tests\.", r.stdout)
    assert m and m.group(1) == m.group(2)
    assert int(m.group(2)) == 13 * 6 * 3 * 2 * 14
    assert "FAILED" not in r.stdout
    assert f"(10 .. 2^{log2n})" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("desc", [False, True])
def test_benchmark(desc):
    args = ["--n", str(3_000_017), "--iterations", "3", "--warmup", "1", "--warmup_ms", "0"]
    r = run(*args, *(["--descending"] if desc else []), timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "sorted yes, digit histograms yes" in r.stdout
    assert re.search(r"rate\s+[0-9.]+ GKeys/s", r.stdout)
