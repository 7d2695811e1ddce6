"""Build recipe for the gfx950 library libvdi_hip.so (in-tree).

Usage: python -m vdropimpact_amd.build   (or call build()).
Compiles every HIP source under csrc/ for gfx950 with hipcc and links one
shared library exporting the C ABI declared in include/vdi.h, then builds the
C++ host layer above it (host/: fluid_test, cfg_dump, tt_analysis in
_lib/bin).
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "_lib")
LIB = os.path.join(LIBDIR, "libvdi_hip.so")

SOURCES = ["mg.hip", "step.hip", "batch.hip", "nn.hip", "noalloc.hip"]
HEADERS = ["vdi_internal.hpp", "mg.hpp", "step.hpp", "nn.hpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed-FP32 instructions are disabled for all device code.
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def _objects_and_cmds():
    os.makedirs(os.path.join(LIBDIR, "obj"), exist_ok=True)
    cmds, objs = [], []
    for src in SOURCES:
        obj = os.path.join(LIBDIR, "obj", src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, *FLAGS, "-I", os.path.join(REPO, "include"), "-c",
                     os.path.join(CSRC, src), "-o", obj])
    return objs, cmds


def build(verbose: bool = False) -> str:
    """Compile every source and link libvdi_hip.so; always rebuilds."""
    objs, cmds = _objects_and_cmds()
    procs = []
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for cmd, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append((cmd, text))
        elif verbose and text.strip():
            print(text)
    if failed:
        msg = "\n".join(" ".join(c) + "\n" + t for c, t in failed)
        raise RuntimeError("HIP build failed:\n" + msg)
    tmp = LIB + ".tmp"
    link = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(link), flush=True)
    subprocess.run(link, check=True)
    os.replace(tmp, LIB)
    # the C++ host layer (configuration reader, frame loop, files, tools)
    subprocess.run(["make", "-s", "-B", "-C", os.path.join(PKG, "host")], check=True)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
