"""Compile the HIP library in-tree for gfx950.

    python -m penguin_amd.build          # -> penguin_amd/libpenguin_amd.so

hipcc cross-compiles without a GPU.  Flags:
  --offload-arch=gfx950                 MI355X (CDNA4)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                        no packed FP32 instructions in device
                                        code (the host pass warns about the
                                        unknown feature; harmless)
  -fno-hip-fp32-correctly-rounded-divide-sqrt
                                        FP32 division and square root by the
                                        hardware approximations (~1 ulp);
                                        FP64 is unaffected
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpenguin_amd.so")
SOURCES = ["pg_kernels.hip", "pg_adjoint.hip", "pg_sph.hip", "pg_probe.hip", "pg_host.cpp"]
HEADERS = ["pg_internal.h", "pg_launch.h", "pg_device.h", "pg_adjoint_fn.h", "../../include/penguin_amd.h", "../../include/pg_params.h"]


def hipcc() -> str:
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found (ROCm toolchain required to build penguin_amd)")


def flags() -> list:
    return ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-fno-hip-fp32-correctly-rounded-divide-sqrt",
            "-Wall", "-Wno-unused-result", "-Wno-unused-variable", "-Wno-unused-function"]


def build(verbose: bool = False, out: str = LIB) -> str:
    objs = []
    tmpdir = os.path.join(HERE, "_build")
    os.makedirs(tmpdir, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(tmpdir, src + ".o")
        cmd = [hipcc()] + flags() + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"hipcc failed on {src}")
        if verbose and r.stderr.strip():
            sys.stderr.write(r.stderr)
        objs.append(obj)
    tmp_out = out + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp_out] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp_out, out)
    build_adjcheck(verbose)
    return out


ADJCHECK = os.path.join(HERE, "libpg_adjcheck.so")


def build_adjcheck(verbose: bool = False, out: str = ADJCHECK) -> str:
    """Host-only library evaluating the per-cell functions and their
    adjoints in FP64 (pg_adjcheck.cpp), for the derivative checks of the
    CPU test suite."""
    cmd = [hipcc()] + flags() + ["-shared", "-x", "hip", os.path.join(CSRC, "pg_adjcheck.cpp"), "-o", out + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed on pg_adjcheck.cpp")
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
