"""In-tree build of the native libraries (hipcc for gfx950, gcc for host-only code).

    python -m eagleimp_amd.build          # builds libeimp.so and libeimp_synth.so

The libraries land next to this file so they travel with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)

ARCH = os.environ.get("EIMP_OFFLOAD_ARCH", "gfx950")
# packed FP32 instructions are kept out of every HIP build
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

LIB = os.path.join(HERE, "libeimp.so")
SYNTH = os.path.join(HERE, "libeimp_synth.so")


def _hipcc() -> str:
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd, quiet=False):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in r.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in line)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    if out.strip() and not quiet:
        print(out)


def build_synth():
    _run(["gcc", "-std=c11", "-O3", "-fopenmp", "-fPIC", "-shared", "-Wall", "-Wextra",
          "-o", SYNTH, os.path.join(CSRC, "synth.c"), "-lm"])


# host-only phasing code: float results must equal the reference's, so no
# contraction into fused multiply-adds
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-ffp-contract=off", "-pthread"]


def build_engine():
    hip = _hipcc()
    objs = []
    for src in ("eimp_kernels.hip", "eimp_engine.cpp"):
        obj = os.path.join(CSRC, "_" + os.path.splitext(src)[0] + ".o")
        _run([hip, f"--offload-arch={ARCH}", *HIP_FLAGS, "-x", "hip", "-c",
              os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in ("eimp_phase.cpp", "eimp_qref.cpp"):
        obj = os.path.join(CSRC, "_" + os.path.splitext(src)[0] + ".o")
        _run(["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([hip, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)


def build_all():
    build_synth()
    build_engine()


if __name__ == "__main__":
    build_all()
    print("built", LIB, SYNTH, file=sys.stderr)
