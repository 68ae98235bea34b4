"""Build libfc_dem.so (the HIP kernels + C ABI) in-tree for gfx950.

Usage: python -m fullc0ntact_amd.build [--jobs N] [--verbose]

Every .hip file under csrc/ is compiled with hipcc to an object file under
fullc0ntact_amd/_build/, then linked into fullc0ntact_amd/libfc_dem.so.  The
library is always rebuilt from source (the build is the "does it compile"
check); objects are reused only within one build invocation.  The C++ test
driver of the host API (tests/cpp/fc_particles_driver.cpp) is built against
the library afterwards.
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libfc_dem.so"
ARCH = os.environ.get("FC_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libfc_dem.so")


# Flags common to every translation unit.
#  -ffp-contract=off : no FMA contraction, so the FP32 particle arithmetic
#                      matches the CPU restatement operation for operation.
#  -packed-fp32-ops  : keep the compiler from emitting packed FP32 ops.
BASE_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument", "-Wno-unused-value",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip"))


def compile_one(src: Path, verbose: bool) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [hipcc(), *BASE_FLAGS, "-c", str(src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stdout}\n{r.stderr}")
    msgs = [ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if verbose and msgs:
        print("\n".join(msgs), flush=True)
    return obj


def build(jobs: int | None = None, verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = sources()
    if not srcs:
        raise RuntimeError(f"no .hip sources in {CSRC}")
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 2)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: compile_one(s, verbose), srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-fPIC", *map(str, objs), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


DRIVER_SRC = ROOT / "tests" / "cpp" / "fc_particles_driver.cpp"
DRIVER = DRIVER_SRC.with_suffix("")


def build_driver(verbose: bool = False) -> Path | None:
    """Host-only C++ program over include/fc_particle_system.hpp; links
    libfc_dem.so with an rpath relative to its own location."""
    if not DRIVER_SRC.exists():
        return None
    cxx = shutil.which("g++") or shutil.which("c++")
    if not cxx:
        raise RuntimeError("g++ not found: needed for the C++ host-API driver")
    rel = os.path.relpath(PKG, DRIVER.parent)
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}", str(DRIVER_SRC),
           f"-L{PKG}", "-lfc_dem", f"-Wl,-rpath,$ORIGIN/{rel}", "-o", str(DRIVER)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver build failed:\n{r.stdout}\n{r.stderr}")
    return DRIVER


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args(argv)
    lib = build(a.jobs, a.verbose)
    print(f"built {lib}")
    drv = build_driver(a.verbose)
    if drv:
        print(f"built {drv}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
