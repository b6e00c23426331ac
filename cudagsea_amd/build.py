"""Build libgsea_hip.so (HIP kernels + C ABI + C++ proxies) for gfx950.

Run as `python -m cudagsea_amd.build`.  The library is written in-tree next to
this file so it travels with the source tree.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libgsea_hip.so")

# one object per translation unit, compiled in parallel; the kernel
# families are instantiated in their own units (launch_*.hip)
SOURCES = ["gsea_engine.hip", "launch_metric_f32.hip", "launch_metric_f64.hip",
           "launch_rank.hip", "launch_es_f32.hip", "launch_es_f64.hip",
           "gsea_surrogate.hip", "gsea_proxy.cpp", "gsea_io.cpp"]
OBJDIR = os.path.join(HERE, "_build")
CLI = os.path.join(HERE, "cudagsea")   # standalone command line (gsea_cli.cpp)

# IEEE arithmetic in source order: no contraction, no fast-math; packed FP32
# instructions disabled for the device code.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
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


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                          text=True)
    out = "\n".join(l for l in proc.stdout.splitlines()
                    if "is not a recognized feature" not in l)
    if proc.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + out)
    if verbose and out.strip():
        print(out)


def build(verbose=False, jobs=None):
    os.makedirs(OBJDIR, exist_ok=True)
    cc = hipcc()
    objs = []
    cmds = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(OBJDIR, src + ".o")
        objs.append(obj)
        cmds.append([cc] + HIPCC_FLAGS + ["-I", INCLUDE, "-I", CSRC, "-c", path,
                                          "-o", obj])
    jobs = jobs or min(len(cmds), max(1, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in [ex.submit(_run, c, verbose) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    _run([cc] + HIPCC_FLAGS + ["-shared"] + objs + ["-o", tmp], verbose)
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
