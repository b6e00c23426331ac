"""Build libgsea_hip.so (HIP kernels + C ABI + C++ proxies) for gfx950.

Run as `python -m cudagsea_amd.build`.  The library is written in-tree next to
this file so it travels with the source tree.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libgsea_hip.so")

SOURCES = ["gsea_engine.hip", "gsea_surrogate.hip", "gsea_proxy.cpp", "gsea_io.cpp"]
CLI = os.path.join(HERE, "cudagsea")   # standalone command line (gsea_cli.cpp)

# IEEE arithmetic in source order: no contraction, no fast-math; packed FP32
# instructions disabled for the device code.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"),
                 "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES
            if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", INCLUDE, "-I", CSRC] + srcs + [
        "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                          text=True)
    out = "\n".join(l for l in proc.stdout.splitlines()
                    if "is not a recognized feature" not in l)
    if proc.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + out)
    if verbose and out.strip():
        print(out)
    os.replace(tmp, LIB)
    build_cli(verbose)
    return LIB


def build_cli(verbose=False):
    """The command line links the library (host code only, g++)."""
    gxx = shutil.which("g++") or "g++"
    tmp = CLI + ".tmp"
    cmd = [gxx, "-std=c++17", "-O2", "-Wall", "-I", INCLUDE, "-I", CSRC,
           os.path.join(CSRC, "gsea_cli.cpp"), "-L", HERE, "-lgsea_hip",
           "-Wl,-rpath,$ORIGIN", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                          text=True)
    if proc.returncode != 0:
        raise RuntimeError("g++ failed for the command line:\n" + proc.stdout)
    os.replace(tmp, CLI)
    return CLI


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
