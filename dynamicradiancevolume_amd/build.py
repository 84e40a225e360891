"""Builds the native libraries in-tree.

- ``libdrv.so``        the HIP pipeline (gfx950) behind ``include/drv.h``;
- ``libdrv_inputs.so`` the procedural scene / G-buffer / RSM producer;
- ``libdrv_calib.so``  the VALU instruction-rate probe (``calibrate --valu``);
- ``oracle/liborc.so`` the CPU oracle (test infrastructure, built by its own
  Makefile).

Run ``python -m dynamicradiancevolume_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DRV_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["k_allocate.hip", "k_voxel.hip", "k_gather.hip", "k_apply.hip", "k_mlp.hip", "k_batch.hip", "k_raster.hip", "k_bruteforce.hip", "k_direct.hip", "k_specular.hip", "k_static.hip", "drv_api.cpp"]

# -ffp-contract=off: the integer decisions (cell, cascade, voxel coverage) must
# round exactly as the oracle does; hot loops use explicit fmaf.  Packed FP32
# instructions are disabled on this target.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        line for line in proc.stdout.splitlines() if "packed-fp32-ops" not in line and line.strip()
    )
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")


# Per-file flags.  k_mlp.hip: matrix-core results go to ordinary VGPRs (not
# accumulation registers), which saves one register read per result value in
# the inference kernel and keeps it within 128 registers (4 waves per SIMD).
FILE_FLAGS = {"k_mlp.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def _compile_hip(src: str) -> Path:
    obj = BUILD / (Path(src).stem + ".o")
    _run([HIPCC, *HIP_FLAGS, *FILE_FLAGS.get(src, []), "-c", str(CSRC / src), "-o", str(obj)])
    return obj


def build_hip() -> Path:
    BUILD.mkdir(exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(HIP_SOURCES))) as ex:
        objs = list(ex.map(_compile_hip, HIP_SOURCES))
    out = PKG / "libdrv.so"
    tmp = PKG / "libdrv.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, out)
    return out


def build_calib() -> Path:
    """libdrv_calib.so: the VALU instruction-rate probe of ``calibrate --valu``."""
    out = PKG / "libdrv_calib.so"
    tmp = PKG / "libdrv_calib.so.tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", str(tmp), str(CSRC / "calib" / "valu_rate.hip")])
    os.replace(tmp, out)
    return out


def build_inputs() -> Path:
    out = PKG / "libdrv_inputs.so"
    tmp = PKG / "libdrv_inputs.so.tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-o", str(tmp),
          str(CSRC / "inputs" / "drv_inputs.cpp"), "-lpthread"])
    os.replace(tmp, out)
    return out


def build_oracle() -> Path:
    oracle = ROOT / "oracle"
    _run(["make", "-B", "-s"], cwd=oracle)
    return oracle / "liborc.so"


def build_all() -> None:
    if shutil.which(HIPCC) is None and not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    build_inputs()
    build_oracle()
    build_hip()
    build_calib()


if __name__ == "__main__":
    build_all()
    print("built:", PKG / "libdrv.so", PKG / "libdrv_inputs.so", ROOT / "oracle" / "liborc.so", file=sys.stderr)
