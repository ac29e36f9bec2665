"""Builds libkagome_vmc.so (HIP kernels + C ABI) for gfx950 with hipcc.

Each source is compiled to an object in parallel (the translation units share
no device code), then the objects are linked into the shared library.

Usage: python kagomedsl.jl_amd/build.py   (cross-compiles without a GPU)
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "kagome_vmc", "libkagome_vmc.so")
SOURCES = ["kvmc_device.hip", "kvmc_sweep_ib.hip", "kvmc_complex.hip", "kvmc_refresh.hip", "kvmc_api.hip"]
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions stay off: plain v_fma_f32 only
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    return subprocess.run(cmd, capture_output=True, text=True)


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = [os.path.join(OBJ, os.path.splitext(s)[0] + ".o") for s in srcs]
    cmds = [[hipcc, *FLAGS, "-c", "-o", o, os.path.join(CSRC, s)] for s, o in zip(srcs, objs)]
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        results = list(ex.map(lambda c: _run(c, verbose), cmds))
    for res in results:
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError("hipcc failed building libkagome_vmc.so")
    tmp = OUT + ".tmp"
    res = _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("linking libkagome_vmc.so failed")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
