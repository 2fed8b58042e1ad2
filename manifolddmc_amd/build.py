"""Builds libmdmc.so (the HIP library behind include/mdmc.h) in-tree for gfx950.

Usage: python -m manifolddmc_amd.build   (or manifolddmc_amd.build.build())
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmdmc.so")
CLI = os.path.join(HERE, "mdmc_cli")
ROCM_LIB = "/opt/rocm/lib"
SOURCES = ["mdmc_scan.hip", "mdmc_dmc.hip", "mdmc_gridpass.hip", "mdmc_pipeline.hip", "mdmc_adjoint.hip", "mdmc_batch.hip", "mdmc_generate.hip",
           "mdmc_faces.hip", "mdmc_api.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MDMC_ARCH", "gfx950")
FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    # bit-exact float semantics: no contraction, IEEE division / square root, denormals kept
    "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    # no packed FP32 instructions (v_pk_*_f32) in generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-command-line-argument",
    "-I" + CSRC, "-I" + os.path.join(ROOT, "include"),
]


def _compile(src, verbose):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, r.stderr[-6000:]))
    if verbose and r.stderr.strip():
        print(r.stderr[-3000:], file=sys.stderr)
    return obj


def build(verbose=False):
    """Compiles every HIP source and links libmdmc.so; always rebuilds."""
    os.makedirs(BUILD, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-4000:])
    os.replace(tmp, LIB)
    # the command-line front end (host C++ over the C ABI; finds libmdmc.so next to itself)
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"), os.path.join(CSRC, "mdmc_cli.cpp"),
           "-o", CLI + ".tmp", "-L" + HERE, "-lmdmc", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link," + ROCM_LIB]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("mdmc_cli build failed:\n" + r.stderr[-4000:])
    os.replace(CLI + ".tmp", CLI)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
