"""Builds libqsmi.so (HIP kernels + C-ABI + C++ host layer) in-tree for gfx950.

Usage: python -m g2s_amd.build   (or g2s_amd.build.build()).
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libqsmi.so")

SOURCES = ["qsmi_kernels.hip", "qsmi_fast1000.hip", "qsmi_sim.hip", "qsmi_api.hip", "qsmi_lp.hip", "qsmi_host.cpp", "qsmi_calib.cpp",
           "qsmi_bgrid.cpp"]
HEADERS = ["qsmi_fft.hpp", "qsmi_kernels.hpp", "qsmi_internal.hpp"]

FLAGS = [
    "-O3",
    "-std=c++17",
    "--offload-arch=gfx950",
    "-fPIC",
    "-shared",
    # no packed FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-Wno-unused-value",
    "-I" + os.path.join(ROOT, "include"),
    "-lpthread",
    "-lz",
]


def hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the qsmi library cannot be built")


def build(verbose=False):
    """Compile every source into g2s_amd/libqsmi.so (always recompiles)."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + FLAGS + ["-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libqsmi.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
