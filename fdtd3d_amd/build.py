"""Builds the native parts of fdtd3d_amd in-tree for gfx950.

    python -m fdtd3d_amd.build            # libyee.so + the fdtd3d CLI binaries

Outputs (git-ignored, shipped with the tree):
    fdtd3d_amd/libyee.so        C ABI of include/yee.h (HIP kernels + host set-up)
    fdtd3d_amd/bin/fdtd3d       CLI front end, FP64 fields (reference default value type)
    fdtd3d_amd/bin/fdtd3d-f32   CLI front end, FP32 fields
    fdtd3d_amd/bin/fdtd3d-complex  CLI front end, FP64 complex fields (far field with --use-ntff)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libyee.so")
BIN = os.path.join(HERE, "bin")

ARCH = os.environ.get("YEE_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: no FMA contraction, so the FP64 kernels round exactly as the
# reference's per-operation arithmetic; packed FP32 instructions are disabled.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
]

SOURCES = ["yee_setup.cpp", "yee_kernels.hip", "yee_engine.hip", "yee_batch.hip", "yee_adjoint.hip"]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found")
    return path


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd[:3]) + " ...")
    # the host pass does not know the packed-fp32 target feature; drop that warning only
    noise = [ln for ln in res.stderr.splitlines() if ln.strip() and "packed-fp32-ops" not in ln
             and "is not a recognized feature" not in ln]
    if noise:
        sys.stderr.write("\n".join(noise) + "\n")


def build_lib() -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src + ".o")
        flags = list(HIP_FLAGS)
        if src.endswith(".cpp"):
            flags = ["-x", "hip"] + flags
        _run([hipcc()] + flags + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_cli() -> None:
    os.makedirs(BIN, exist_ok=True)
    src = os.path.join(CSRC, "fdtd3d_main.cpp")
    if not os.path.exists(src):
        return
    for name, prec, cplx in (("fdtd3d", "8", "0"), ("fdtd3d-f32", "4", "0"), ("fdtd3d-complex", "8", "1")):
        out = os.path.join(BIN, name)
        _run(["g++", "-O2", "-std=c++17", f"-DYEE_CLI_PRECISION={prec}", f"-DYEE_CLI_COMPLEX={cplx}", src, "-o", out,
              f"-L{HERE}", "-lyee", "-Wl,-rpath,$ORIGIN/..", "-lm"])


def build() -> None:
    build_lib()
    build_cli()


if __name__ == "__main__":
    build()
    print(LIB)
