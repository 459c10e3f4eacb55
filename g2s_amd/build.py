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

SOURCES = ["qsmi_kernels.hip", "qsmi_scan1024.hip", "qsmi_band.hip", "qsmi_sim.hip", "qsmi_api.hip", "qsmi_lp.hip",
           "qsmi_host.cpp", "qsmi_calib.cpp", "qsmi_bgrid.cpp"]
HEADERS = ["qsmi_fft.hpp", "qsmi_kernels.hpp", "qsmi_internal.hpp", "qsmi_select.hpp", "qsmi_taps.hpp"]

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


PLUGIN_SRC = os.path.join(HERE, "g2s_plugin", "qsmi_g2s_device.cpp")
PLUGIN = os.path.join(HERE, "g2s_cuda.so")


def _g2s_include():
    """G2S's include/ directory: $G2S_INCLUDE, else the G2S tree this build
    was made beside (its include/ holding AcceleratorDevice.hpp)."""
    for cand in (os.environ.get("G2S_INCLUDE"), "/root/reference/include"):
        if cand and os.path.exists(os.path.join(cand, "AcceleratorDevice.hpp")):
            return cand
    return None


def _has_header(name):
    r = subprocess.run(["g++", "-E", "-x", "c++", "-"], input=f"#include <{name}>\n", capture_output=True, text=True)
    return r.returncode == 0


def build_plugin(verbose=False):
    """Compile the G2S AcceleratorDevice plugin (g2s_amd/g2s_cuda.so) against
    G2S's headers. Returns the path, or None with the reason printed when
    G2S's headers or the FFTW headers they include are not available."""
    inc = _g2s_include()
    if inc is None:
        print("G2S plugin not built: G2S's include/ not found (set G2S_INCLUDE)")
        return None
    if not _has_header("fftw3.h"):
        print("G2S plugin not built: fftw3.h (included by G2S's AcceleratorDevice.hpp) is not installed")
        return None
    cmd = ["g++", "-O2", "-std=c++17", "-fopenmp", "-shared", "-fPIC", "-I" + inc, "-I" + os.path.join(ROOT, "include"),
           PLUGIN_SRC, "-L" + HERE, "-lqsmi", "-Wl,-rpath," + HERE, "-o", PLUGIN]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed building the G2S plugin")
    return PLUGIN


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_plugin(verbose=True))
