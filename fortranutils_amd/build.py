"""Build liblmgx.so (HIP, gfx950) in-tree, and the Fortran interface module when flang is present."""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")

HIP_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-fPIC", "-shared", "-std=c++17",
    "-Wno-unused-value", "-Wno-unused-result",
    # packed FP32 arithmetic is kept off on these machines
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build_lib(verbose: bool = False) -> str:
    src = os.path.join(HERE, "csrc", "lmgx.hip")
    out = os.path.join(HERE, "liblmgx.so")
    tmp = out + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-o", tmp, src]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build_stream_probe(verbose: bool = False) -> str:
    """libstreamprobe.so: the benchmark's read/write-mix bandwidth probe (not part of liblmgx)."""
    src = os.path.join(HERE, "csrc", "stream_probe.hip")
    out = os.path.join(HERE, "libstreamprobe.so")
    tmp = out + ".tmp"
    r = subprocess.run([HIPCC, *HIP_FLAGS, "-o", tmp, src], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build_fortran(verbose: bool = False) -> str | None:
    """Compile the Fortran interface module (fortran/lmgx_mod.f90) when flang is present; the
    drop-in lmg_solve (fortran/lmg_solve_gpu.F) is compiled inside LESter's build, against
    LESter's include files (oracle/Makefile, target ref-gpu, does exactly that)."""
    if not os.path.exists(FLANG):
        return None
    src = os.path.join(HERE, "fortran", "lmgx_mod.f90")
    moddir = os.path.join(HERE, "fortran", "_mod")
    os.makedirs(moddir, exist_ok=True)
    out = os.path.join(moddir, "lmgx_mod.o")
    cmd = [FLANG, "-O2", "-fPIC", "-c", "-module-dir", moddir, "-o", out, src]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("flang failed:\n" + r.stdout + r.stderr)
    if verbose:
        print("built", out)
    return out


def build(verbose: bool = False) -> None:
    build_lib(verbose)
    build_stream_probe(verbose)
    build_fortran(verbose)


if __name__ == "__main__":
    build(verbose=True)
