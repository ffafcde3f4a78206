"""Builds librsort_amd.so (gfx950) in-tree from csrc/*.hip with hipcc.

Usage: python cuda-fast-sort_amd/build.py [--jobs N]
The object files go to cuda-fast-sort_amd/_build/, the library next to this
file, so it travels with the source tree.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
VARIANTS = {
    # name: (object dir, library, extra flags)
    "release": (os.path.join(HERE, "_build"), os.path.join(HERE, "librsort_amd.so"), []),
    "timing": (os.path.join(HERE, "_build", "timing"), os.path.join(HERE, "librsort_amd_timing.so"),
               ["-DRSA_PHASE_TIMING"]),
}
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-mcode-object-version=5",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def sources() -> list[str]:
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(job: tuple[str, str, list[str]]) -> tuple[str, int, str, float]:
    src, objdir, extra = job
    obj = os.path.join(objdir, os.path.basename(src)[:-4] + ".o")
    t0 = time.time()
    cmd = [HIPCC, *HIP_FLAGS, *extra, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    return obj, r.returncode, out, time.time() - t0


def build(jobs: int | None = None, verbose: bool = True, variant: str = "release") -> str:
    objdir, lib, extra = VARIANTS[variant]
    os.makedirs(objdir, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for obj, rc, out, dt in ex.map(_compile, [(s_, objdir, extra) for s_ in srcs]):
            if out.strip() and verbose:
                print(out, file=sys.stderr)
            if rc != 0:
                raise RuntimeError(f"hipcc failed for {obj}:\n{out}")
            if verbose:
                print(f"[build] {os.path.basename(obj)} {dt:.1f}s", flush=True)
            objs.append(obj)
    tmp = lib + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs,
           "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib)
    if verbose:
        print(f"[build] {lib}", flush=True)
    if variant == "release":
        build_cli(verbose)
    return lib


CLI_PROGRAMS = ("radix_gpu", "seg_percall")


def build_cli(verbose: bool = True) -> str:
    """Host C++ programs over the C ABI: cli/radix_gpu (the benchmark /
    validation driver) and cli/seg_percall (batched segmented sort against one
    sort call per segment)."""
    for name in CLI_PROGRAMS:
        exe = os.path.join(HERE, "cli", name)
        cmd = ["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
               os.path.join(HERE, "cli", name + ".cpp"), "-o", exe + ".tmp", "-L" + HERE, "-lrsort_amd",
               "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib", "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{name} build failed:\n{r.stdout}\n{r.stderr}")
        os.replace(exe + ".tmp", exe)
        if verbose:
            print(f"[build] {exe}", flush=True)
    return os.path.join(HERE, "cli", CLI_PROGRAMS[0])


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--variant", choices=sorted(VARIANTS), default="release")
    a = ap.parse_args()
    build(a.jobs, variant=a.variant)
