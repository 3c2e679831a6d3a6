"""Build libsphnj.so (HIP, gfx950), the `sphnj` host driver, and -- as test infrastructure --
the CPU oracle library.  Everything is built in-tree so the shared objects travel with the
source tree.

    python -m sphnj_amd.build            # build all
    python -m sphnj_amd.build --lib      # library + driver only
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libsphnj.so"
DRIVER = PKG / "bin" / "sphnj"
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "liboracle.so"

HOST_SOURCES = ["host_params.cpp", "host_particles.cpp", "host_cases.cpp"]
HIP_SOURCES = ["sphnj_ctx.hip"]
ARCH = os.environ.get("SPHNJ_OFFLOAD_ARCH", "gfx950")

CXXFLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]
HIPFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # keep the compiler from forming packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ({r.returncode})")


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and Path(c).exists():
            return c
    raise RuntimeError("hipcc not found")


def build_lib() -> Path:
    BUILD.mkdir(exist_ok=True)
    hipcc = _hipcc()
    jobs = []
    objs = []
    for s in HOST_SOURCES:
        o = BUILD / (Path(s).stem + ".o")
        jobs.append(["g++", *CXXFLAGS, "-c", str(CSRC / s), "-o", str(o)])
        objs.append(o)
    for s in HIP_SOURCES:
        o = BUILD / (Path(s).stem + ".hip.o")
        jobs.append([hipcc, *HIPFLAGS, "-c", str(CSRC / s), "-o", str(o)])
        objs.append(o)
    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        for f in [ex.submit(_run, j) for j in jobs]:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    DRIVER.parent.mkdir(exist_ok=True)
    _run(["g++", "-O2", "-std=c++17", str(CSRC / "driver_main.cpp"), "-I", str(ROOT / "include"),
          "-L", str(PKG), "-lsphnj", f"-Wl,-rpath,$ORIGIN/..", "-o", str(DRIVER)])
    return LIB


def build_oracle() -> Path:
    """The CPU oracle (test infrastructure; never linked into the product)."""
    tmp = ORACLE_LIB.with_suffix(".so.tmp")
    _run(["g++", "-O2", "-fopenmp", "-fPIC", "-shared", "-std=c++17", "-ffp-contract=off",
          str(ORACLE_DIR / "sphnj_oracle.cpp"), "-o", str(tmp)])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="library and driver only")
    ap.add_argument("--oracle", action="store_true", help="oracle only")
    a = ap.parse_args()
    if not a.oracle:
        print("built", build_lib())
    if not a.lib:
        print("built", build_oracle())


if __name__ == "__main__":
    main()
