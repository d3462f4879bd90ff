"""Builds libdoci_amd.so (HIP kernels for gfx950 + C host layer) and the
`doci` command-line program, in-tree.

    python -m doci_amd.build
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OUT = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libdoci_amd.so")
CLI = os.path.join(PKG, "bin", "doci")
CALLER = os.path.join(ROOT, "tests", "c", "dropin_caller")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("DOCI_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 VALU forms stay disabled for every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]
C_FLAGS = ["-O2", "-std=gnu11", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter",
           "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include"]


def _sources(sub: str, ext: str):
    d = os.path.join(CSRC, sub)
    return sorted(os.path.join(d, f) for f in os.listdir(d) if f.endswith(ext))


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build(verbose: bool = False) -> str:
    hipcc = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    gcc = shutil.which("gcc") or "cc"
    os.makedirs(OUT, exist_ok=True)
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    jobs = []
    for src in _sources("kernels", ".hip"):
        obj = os.path.join(OUT, os.path.basename(src) + ".o")
        jobs.append([hipcc, *HIP_FLAGS, "-c", src, "-o", obj])
    for src in _sources("host", ".c"):
        obj = os.path.join(OUT, os.path.basename(src) + ".o")
        jobs.append([gcc, *C_FLAGS, "-c", src, "-o", obj])
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout + r.stderr, file=sys.stderr)
    objs = [j[-1] for j in jobs]
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs,
          f"-L{ROCM}/lib", "-lamdhip64", "-lm", f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(tmp, LIB)
    _run([gcc, *C_FLAGS, "-o", CLI, os.path.join(CSRC, "cli", "doci_main.c"),
          f"-L{PKG}", "-ldoci_amd", f"-L{ROCM}/lib", "-lamdhip64", "-lm",
          "-Wl,-rpath,$ORIGIN/..", f"-Wl,-rpath,{ROCM}/lib"])
    # a builder-written caller of the drop-in ABI used by the GPU tests
    _run([gcc, *C_FLAGS, "-o", CALLER, os.path.join(ROOT, "tests", "c", "dropin_caller.c"),
          f"-L{PKG}", "-ldoci_amd", "-lm", "-Wl,-rpath,$ORIGIN/../../doci_amd", f"-Wl,-rpath,{ROCM}/lib"])
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
