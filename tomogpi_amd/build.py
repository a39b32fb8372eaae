"""Build libtomogpi_mi355x.so (HIP, gfx950) in-tree with hipcc.

    python -m tomogpi_amd.build [--jobs N] [--verbose]

Every .hip / .cpp under tomogpi_amd/csrc is compiled to an object under
tomogpi_amd/_build and linked into tomogpi_amd/libtomogpi_mi355x.so; the
command-line driver tomogpi_amd/bin/tomogpi links against it.  The build
always recompiles (it is the "does it build" check).
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libtomogpi_mi355x.so")
CLI_DIR = os.path.join(PKG, "bin")
CLI = os.path.join(CLI_DIR, "tomogpi")
ARCH = os.environ.get("TOMOGPI_ARCH", "gfx950")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

# -ffp-contract=off + correctly rounded FP32 div/sqrt: the kernels spell out
# every fused multiply-add, so their FP32 results are reproducible on the CPU.
# Packed FP32 instructions are disabled for the device (-packed-fp32-ops).
COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          "-I" + os.path.join(ROOT, "include")]
DEVICE = ["--offload-arch=" + ARCH, "-fhip-fp32-correctly-rounded-divide-sqrt",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
LIB_SOURCES = ["api.cpp", "config.cpp", "loop.cpp", "serve.cpp", "kernels_vi.hip", "kernels_rs.hip",
               "kernels_siddon.hip", "kernels_vec.hip"]
CLI_SOURCES = ["tomogpi_main.cpp"]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required)")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose and r.stderr.strip():
        print(r.stderr, file=sys.stderr)
    return r


def compile_one(src: str, verbose: bool) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src.replace(".", "_") + ".o")
    cmd = [hipcc()] + COMMON
    if src.endswith(".hip"):
        cmd += DEVICE + ["-x", "hip"]
    else:
        # host code: no device pass
        cmd += ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include")]
    cmd += ["-c", path, "-o", obj]
    _run(cmd, verbose)
    return obj


def build(jobs: int = 8, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(lambda s: compile_one(s, verbose), LIB_SOURCES))
    tmp = LIB + ".tmp"
    _run([hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs, verbose)
    os.replace(tmp, LIB)
    cli_srcs = [s for s in CLI_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    if cli_srcs:
        os.makedirs(CLI_DIR, exist_ok=True)
        cli_objs = [compile_one(s, verbose) for s in cli_srcs]
        _run([hipcc(), "-o", CLI + ".tmp"] + cli_objs +
             ["-L" + PKG, "-ltomogpi_mi355x", "-Wl,-rpath,$ORIGIN/.."], verbose)
        os.replace(CLI + ".tmp", CLI)
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args(argv)
    print(build(a.jobs, a.verbose))


if __name__ == "__main__":
    main()
