"""Build libpcfdns.so in-tree with hipcc for gfx950.

Usage: python -m jaxfun_amd.build   (or jaxfun_amd.build.build()).
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libpcfdns.so"
ARCH = os.environ.get("PCF_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions stay disabled for all device code.
HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the HIP library cannot be built")


def build(verbose: bool = False) -> Path:
    hipcc = _hipcc()
    BUILD.mkdir(exist_ok=True)
    srcs = ("pcf_setup.cpp", "pcf_kernels.hip", "pcf_adjoint.hip", "pcf_mfma.hip", "pcf_factored.hip", "pcf_diag.hip",
            "pcf_api.hip", "tc_setup.cpp", "tc_kernels.hip", "tc_api.hip")
    cmds, objs = [], []
    for src in srcs:
        obj = BUILD / (Path(src).stem + ".o")
        cmd = [hipcc, "-c", str(CSRC / src), "-o", str(obj)]
        cmd += HIP_FLAGS if src.endswith(".hip") else ["-O3", "-std=c++17", "-fPIC"]
        if verbose:
            print(" ".join(cmd), flush=True)
        cmds.append(cmd)
        objs.append(str(obj))
    # independent translation units: compile side by side
    with ThreadPoolExecutor(max_workers=len(cmds)) as pool:
        for res in pool.map(lambda c: subprocess.run(c, capture_output=True, text=True), cmds):
            if res.returncode:
                sys.stderr.write(res.stdout + res.stderr)
                raise subprocess.CalledProcessError(res.returncode, res.args)
            elif verbose and res.stderr:
                sys.stderr.write(res.stderr)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc, "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
