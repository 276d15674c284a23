"""Builds the package's native libraries in-tree.

libgh_synth.so   -- host C++ synthetic panel / read generator (csrc/synth.cpp)
libglimpse_hip.so -- HIP/CDNA4 kernels + C-ABI (csrc/*.hip), gfx950 only

All HIP code is compiled with packed-FP32 instruction selection disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)

HIP_SOURCES = [
    "gh_capi.hip", "gh_panel.hip", "gh_condset.hip", "gh_impute.hip", "gh_phase.hip", "gh_pbwt.hip", "gh_gibbs.hip",
]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result",
]


def _run(cmd):
    print("[build]", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_synth():
    out = os.path.join(HERE, "libgh_synth.so")
    _run(["g++", "-std=c++17", "-O3", "-fPIC", "-shared", "-o", out,
          os.path.join(CSRC, "synth.cpp")])
    return out


def build_hip():
    out = os.path.join(HERE, "libglimpse_hip.so")
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    srcs = [s for s in srcs if os.path.exists(s)]
    if not srcs:
        return None
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    _run([hipcc] + HIPCC_FLAGS + ["-I", os.path.join(ROOT, "include"), "-o", out] + srcs)
    return out


def build_all():
    build_synth()
    build_hip()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
