"""Build the TurboCat MI355X library and its Python extension, in-tree.

    python turbo-cat_amd/build.py          # both
    python turbo-cat_amd/build.py --lib    # libturbocat_hip.so only

Outputs
  turbo-cat_amd/lib/libturbocat_hip.so           C-ABI (include/turbocat_hip.h), gfx950
  turbo-cat_amd/turbocat/_turbocat*.so           pybind11 module over the C-ABI
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libturbocat_hip.so")
PKG = os.path.join(HERE, "turbocat")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_SOURCES = ["tc_capi.hip", "tc_quantize.hip", "tc_train.hip", "tc_predict.hip"]
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-munsafe-fp-atomics",
    # no implicit contraction: the kernels fuse exactly where they call __fmaf_rn,
    # which is where the reference's compiler fuses (bit parity of gradients,
    # split gains, training and prediction sums)
    "-ffp-contract=off",
    # keep the compiler from forming packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib(jobs: int = 8) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    procs, objs = [], []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
    rc = [p.wait() for p in procs]
    if any(rc):
        raise SystemExit(f"hipcc failed: {rc}")
    tmp = LIB + ".tmp"
    run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


def build_ext() -> str:
    import pybind11

    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(PKG, "_turbocat" + suffix)
    inc = [pybind11.get_include(), sysconfig.get_paths()["include"], os.path.join(ROOT, "include"),
           CSRC]
    cmd = ["g++", "-O3", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden",
           *[f"-I{i}" for i in inc],
           os.path.join(CSRC, "booster.cpp"), os.path.join(CSRC, "bindings.cpp"),
           "-L" + LIBDIR, "-lturbocat_hip", "-Wl,-rpath,$ORIGIN/../lib", "-o", out + ".tmp"]
    run(cmd)
    os.replace(out + ".tmp", out)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="only the HIP library")
    a = ap.parse_args(argv)
    build_lib()
    if not a.lib:
        build_ext()


if __name__ == "__main__":
    main(sys.argv[1:])
