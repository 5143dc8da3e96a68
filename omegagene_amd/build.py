"""Build libog_engine.so (HIP for gfx950) in-tree with hipcc.

Usage:  python -m omegagene_amd.build      (or omegagene_amd.build.build())

Every HIP source is compiled with packed FP32 disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops).  The grid-build and
integrator sources additionally use -ffp-contract=off so their float/double
expressions round like the reference host build; the pair kernel keeps FMA
contraction.  Correctly rounded float division/sqrt stay at the hipcc
default.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libog_engine.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-I" + os.path.join(ROOT, "include"), "-Wno-unused-result", "-Wno-unused-value"]
SOURCES = [
    ("og_grid.hip", ["-ffp-contract=off"]),
    ("og_pair.hip", []),
    ("og_integrate.hip", ["-ffp-contract=off"]),
    ("og_bonded.hip", ["-ffp-contract=off"]),
    ("og_engine.cpp", ["-x", "hip"]),
    ("og_celeste.cpp", ["-x", "hip", "-ffp-contract=off"]),
    ("og_server.cpp", ["-x", "hip"]),
    ("og_noalloc.hip", ["-ffp-contract=off"]),
    ("og_io.cpp", ["-x", "hip"]),
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build(verbose: bool = False) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src, extra in SOURCES:
        obj = os.path.join(objdir, src.rsplit(".", 1)[0] + ".o")
        out = _run([HIPCC] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj])
        if verbose and out.strip():
            print(out)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
