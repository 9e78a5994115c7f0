"""Build the gfx950 shared library ``ecogen_amd/libecogen_amd.so`` in-tree.

Compiles every HIP/C++ source of ``ecogen_amd/csrc`` with hipcc for gfx950
(one process per translation unit, in parallel) and links the shared library
that exports the C ABI of ``include/ecogen.h``.

    python -m ecogen_amd.build                  # every translation unit
    python -m ecogen_amd.build --incremental    # only out-of-date ones
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "libecogen_amd.so"
OBJ = PKG / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# FP64 parity build: no FMA contraction (the reference is compiled without FMA),
# packed FP32 instructions disabled.
COMMON = ["-O3", f"--offload-arch={ARCH}", "-std=c++17", "-fPIC", "-ffp-contract=off",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-pass-failed", "-I", str(PKG.parent / "include")]

SOURCES = ["eco_k_kapila.hip", "eco_k_multip.hip", "eco_k_euler.hip", "eco_k_kapila32.hip", "eco_k_multip32.hip",
           "eco_k_kapila3.hip", "eco_k_multip3.hip", "eco_k_kapila4.hip", "eco_k_multip4.hip",
           "eco_k_thermaleq.hip", "eco_k_eulerhom.hip", "eco_transport.hip", "eco_dispatch.cpp",
           "eco_capi.cpp", "eco_pool.cpp", "eco_api.cpp", "eco_noalloc.cpp"]


def _deps_mtime(src: str) -> float:
    # every translation unit includes headers of csrc/ and include/ only
    hdrs = [*CSRC.glob("*.h"), *CSRC.glob("*.hpp"), *CSRC.glob("*.inl"), *(PKG.parent / "include").glob("*.h*")]
    return max([(CSRC / src).stat().st_mtime, *(h.stat().st_mtime for h in hdrs)])


def _compile(src: str, incremental: bool = False) -> Path:
    obj = OBJ / (Path(src).stem + ".o")
    if incremental and obj.exists() and obj.stat().st_mtime > _deps_mtime(src):
        return obj
    cmd = [HIPCC, *COMMON, "-c", str(CSRC / src), "-o", str(obj)]
    if src.endswith(".cpp"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr[-6000:]}")
    return obj


API_CHECK_SRC = PKG.parent / "tests" / "cpp" / "api_check.cpp"
API_CHECK = PKG.parent / "tests" / "cpp" / "api_check"


def build_api_check() -> Path:
    """Host test program of the C++ class API, linked against the library."""
    cmd = ["g++", "-O2", "-std=c++17", "-I", str(PKG.parent / "include"), str(API_CHECK_SRC), str(OUT),
           f"-Wl,-rpath,{PKG}", "-Wl,-rpath,/opt/rocm/lib", "-o", str(API_CHECK)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"api_check build failed:\n{r.stderr[-3000:]}")
    return API_CHECK


def build(verbose: bool = True, incremental: bool = False) -> Path:
    """Compiles every translation unit (incremental=True: only those older
    than their source or any header) and links the library."""
    OBJ.mkdir(exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(lambda src: _compile(src, incremental), SOURCES))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, OUT)
    if API_CHECK_SRC.exists():
        build_api_check()
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv)
    sys.exit(0)
