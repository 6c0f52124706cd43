"""Build the HIP library (gfx950) and the CPU oracle in-tree.

    python -m fluidized_bed_master_amd.build           # library + oracle
    python -m fluidized_bed_master_amd.build --lib     # library only

The library is compiled file by file with hipcc for gfx950 and linked into
fluidized_bed_master_amd/libsophia.so; the FP64 oracle (test
infrastructure) is built by oracle/Makefile into oracle/liboracle.so.
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
LIB = PKG / "libsophia.so"
ORACLE_LIB = ROOT / "oracle" / "liboracle.so"

SOURCES = ["nnps.hip", "sph.hip", "dem.hip", "wcsph.hip", "wcsph2d.hip", "ppe.hip", "coarse.hip", "learned.hip", "context.cpp"]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ROCM_LIB = str(Path(os.environ.get("ROCM_PATH", "/opt/rocm")) / "lib")

COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "--offload-arch=gfx950",
    # packed FP32 (v_pk_*_f32) stays off: scalar FP32 only
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-I", str(ROOT / "include"),
]


def _compile(src: str) -> Path:
    out = BUILD / (Path(src).stem + ".o")
    cmd = [HIPCC, *COMMON, "-x", "hip", "-c", str(CSRC / src), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return out


def build_library(jobs: int = 8) -> Path:
    BUILD.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(jobs, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
           f"-L{ROCM_LIB}", "-lrocblas", f"-Wl,-rpath,{ROCM_LIB}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    r = subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    return ORACLE_LIB


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="library only")
    ap.add_argument("--oracle", action="store_true", help="oracle only")
    a = ap.parse_args(argv)
    if not a.oracle:
        print("built", build_library())
    if not a.lib:
        print("built", build_oracle())
    return 0


if __name__ == "__main__":
    sys.exit(main())
