"""Build the deeds_hip library (and command-line tools) for gfx950 in-tree.

    python -m deedsbcv_amd.build          # library + tools
    python -m deedsbcv_amd.build --jobs 8

Outputs:
    deedsbcv_amd/lib/libdeeds_hip.so      the C-ABI library (include/deeds_hip.h)
    deedsbcv_amd/bin/deedsBCV, applyBCV, applyBCVfloat, diceMulti, linearBCV   drop-in command lines
Objects go to deedsbcv_amd/_build/.  Every translation unit is compiled with
-ffp-contract=off (the reference's arithmetic has no fused multiply-adds) and
without packed-FP32 instruction selection.
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
LIB = PKG / "lib" / "libdeeds_hip.so"
BIN = PKG / "bin"
ARCH = os.environ.get("DEEDS_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-I", str(ROOT / "include"), "-Wall",
          "-Wno-unused-result", "-Wno-unused-variable"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

LIB_SOURCES = ["mind.hip", "datacost.hip", "mst.hip", "bp.hip", "warp.hip", "register.hip", "api.cpp",
               "mst_host.cpp", "stats.hip", "stats_host.cpp", "batch.cpp", "soft_features.hip", "soft_cost.hip",
               "soft_bp.hip", "soft_api.cpp", "dice.hip", "linear.hip", "linear_host.cpp"]
TOOLS = {"deedsBCV": "tool_deeds.cpp", "applyBCV": "tool_apply.cpp", "applyBCVfloat": "tool_apply.cpp",
         "diceMulti": "tool_dice.cpp", "linearBCV": "tool_linear.cpp"}


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    cmd = [HIPCC, *COMMON, *DEVICE, "-c", str(src), "-o", str(obj)]
    if src.suffix == ".hip":
        cmd[1:1] = ["-x", "hip"]
    _run(cmd)
    return obj


def build_library(jobs: int = 8) -> Path:
    BUILD.mkdir(parents=True, exist_ok=True)
    LIB.parent.mkdir(parents=True, exist_ok=True)
    srcs = [CSRC / s for s in LIB_SOURCES]
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp), "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


def build_tools() -> list[Path]:
    BIN.mkdir(parents=True, exist_ok=True)
    out = []
    for name, src in TOOLS.items():
        exe = BIN / name
        defs = ["-DDEEDS_APPLY_FLOAT=1"] if name == "applyBCVfloat" else []
        _run(["g++", "-O2", "-std=c++17", *defs, "-I", str(ROOT / "include"), str(CSRC / src), "-o", str(exe),
              "-L", str(LIB.parent), "-ldeeds_hip", f"-Wl,-rpath,$ORIGIN/../lib", "-lz", "-lpthread"])
        out.append(exe)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--no-tools", action="store_true")
    a = ap.parse_args(argv)
    lib = build_library(a.jobs)
    print("built", lib)
    if not a.no_tools and all((CSRC / s).exists() for s in set(TOOLS.values())):
        for t in build_tools():
            print("built", t)


if __name__ == "__main__":
    sys.exit(main())
