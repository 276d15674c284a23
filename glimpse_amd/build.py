"""Builds the package's native libraries in-tree.

libgh_synth.so   -- host C++ synthetic panel / read generator (csrc/synth.cpp)
libglimpse_hip.so -- HIP/CDNA4 kernels + C-ABI (csrc/*.hip), gfx950 only, and the
                     host-only VCF/BCF I/O (csrc/gh_vcf.cpp, zlib)

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
    "gh_mlp.hip", "gh_select.hip", "gh_output.hip", "gh_noalloc.hip", "gh_concord.hip", "gh_ligate.hip",
    "gh_bam.hip", "gh_graph.hip", "gh_init.hip",
]

# host-only C++ (file I/O on zlib), built with g++ into the same library
CXX_SOURCES = ["gh_vcf.cpp", "gh_binpanel.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result",
]


def _run(cmd):
    print("[build]", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_synth():
    out = os.path.join(HERE, "libgh_synth.so")
    _run(["g++", "-std=c++17", "-O3", "-march=x86-64-v2", "-fPIC", "-shared", "-pthread", "-o", out,
          os.path.join(CSRC, "synth.cpp")])
    return out


def build_hip():
    """One object per source (compiled in parallel), then one shared library."""
    out = os.path.join(HERE, "libglimpse_hip.so")
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    jobs, objs = [], []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmd = [hipcc] + HIPCC_FLAGS + ["-I", os.path.join(ROOT, "include"), "-c", "-o", obj,
                                       os.path.join(CSRC, src)]
        print("[build]", " ".join(cmd), flush=True)
        jobs.append(subprocess.Popen(cmd))
        objs.append(obj)
    for src in CXX_SOURCES:
        obj = os.path.join(objdir, src.replace(".cpp", ".o"))
        cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-I", os.path.join(ROOT, "include"), "-c", "-o", obj,
               os.path.join(CSRC, src)]
        print("[build]", " ".join(cmd), flush=True)
        jobs.append(subprocess.Popen(cmd))
        objs.append(obj)
    bad = [src for src, j in zip(HIP_SOURCES + CXX_SOURCES, jobs) if j.wait() != 0]
    if bad:
        raise RuntimeError("compile failed on " + ", ".join(bad))
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out] + objs + ["-lz"])
    return out


def build_all():
    build_synth()
    build_hip()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
