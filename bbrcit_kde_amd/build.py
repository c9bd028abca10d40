"""Build the native parts of bbrcit_kde_amd in-tree.

* ``lib/libbbrcit_hip.so``: the HIP kernels and the C ABI declared in
  ``include/bbrcit_hip.h``, compiled by hipcc for gfx950 only.

Run as ``python -m bbrcit_kde_amd.build`` or call :func:`build_library`.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
REPO_DIR = PKG_DIR.parent
CSRC = PKG_DIR / "csrc"
LIB_DIR = PKG_DIR / "lib"
LIB_PATH = LIB_DIR / "libbbrcit_hip.so"

HIP_SOURCES = ["bk_kernels.hip", "bk_capi.hip", "bk_cvgrid.hip", "bk_mlp.hip", "bk_order.hip", "bk_sample.hip", "bk_tree.hip", "bk_tune.hip"]

# -packed-fp32-ops: keep the compiler from forming v_pk_{add,mul,fma}_f32.
HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
]


# Per-source additions.  bk_kernels.hip: MFMA results go straight to VGPRs
# (the Gram kernels read every result with the VALU; the AGPR form would add
# two v_accvgpr_read per pair term).
# bk_sample.hip: no contraction of a * b + c into FMAs anywhere in the file
# (its samples must round each product and sum as the reference's host code).
SOURCE_FLAGS = {"bk_kernels.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"], "bk_sample.hip": ["-ffp-contract=off"]}


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found; the HIP library cannot be built")
    return exe


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        sys.stderr.write(proc.stderr)
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}")


def build_library(verbose: bool = False) -> Path:
    """Compile every HIP source and link libbbrcit_hip.so (always rebuilds)."""
    LIB_DIR.mkdir(parents=True, exist_ok=True)
    build_dir = PKG_DIR / "_build"
    build_dir.mkdir(exist_ok=True)
    objs, cmds = [], []
    for src in HIP_SOURCES:
        obj = build_dir / (Path(src).stem + ".o")
        cmd = [hipcc(), *HIPCC_FLAGS, *SOURCE_FLAGS.get(src, []), "-I", str(REPO_DIR / "include"), "-c",
               str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        cmds.append(cmd)
        objs.append(str(obj))
    # the sources compile independently: one hipcc per source, at most 4 at once
    with ThreadPoolExecutor(max_workers=min(4, len(cmds))) as pool:
        for f in [pool.submit(_run, c) for c in cmds]:
            f.result()
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


CPP_DRIVER_SRC = REPO_DIR / "tests" / "cpp" / "api_driver.cc"
CPP_DRIVER_LIB = REPO_DIR / "tests" / "cpp" / "libapi_driver.so"


def build_cpp_driver(verbose: bool = False) -> Path:
    """Compile the C++ API test driver (host C++ over the C ABI) with g++."""
    gxx = shutil.which("g++") or "g++"
    tmp = CPP_DRIVER_LIB.with_suffix(".so.tmp")
    cmd = [gxx, "-std=c++14", "-O2", "-fPIC", "-shared", "-Wall", "-Wno-unused-function",
           "-I", str(PKG_DIR / "include"), "-I", str(REPO_DIR / "include"), str(CPP_DRIVER_SRC),
           "-L", str(LIB_DIR), "-lbbrcit_hip", "-Wl,-rpath,$ORIGIN/../../bbrcit_kde_amd/lib", "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd)
    os.replace(tmp, CPP_DRIVER_LIB)
    return CPP_DRIVER_LIB


GRID_SEARCH_SRC = PKG_DIR / "grid_search" / "grid_search.cc"
GRID_SEARCH_BIN = PKG_DIR / "grid_search" / "grid_search"


def build_grid_search(verbose: bool = False) -> Path:
    """Compile the grid_search tool (host C++ over the C ABI) with g++."""
    gxx = shutil.which("g++") or "g++"
    tmp = GRID_SEARCH_BIN.with_suffix(".tmp")
    cmd = [gxx, "-std=c++14", "-O2", "-Wall", "-I", str(PKG_DIR / "include"), "-I", str(REPO_DIR / "include"),
           str(GRID_SEARCH_SRC), "-L", str(LIB_DIR), "-lbbrcit_hip", "-Wl,-rpath,$ORIGIN/../lib", "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd)
    os.replace(tmp, GRID_SEARCH_BIN)
    return GRID_SEARCH_BIN


def main() -> int:
    path = build_library(verbose=True)
    print(f"built {path}")
    if CPP_DRIVER_SRC.exists():
        print(f"built {build_cpp_driver(verbose=True)}")
    print(f"built {build_grid_search(verbose=True)}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
