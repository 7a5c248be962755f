"""Build libcsx.so in-tree for gfx950 (HIP kernels + host preparation + C ABI).

    python -m crystalscatter_amd.build        (or __graft_entry__.build())

Host preparation (csx_prep.cpp) and the batched fit driver (csx_fit.cpp) are
plain C++ built with g++ like the reference's engine (no FMA contraction on the
x86-64 baseline); the device file and the facade are built with hipcc
--offload-arch=gfx950.  Packed-FP32
instruction selection is disabled for all HIP code.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcsx.so")
ARCH = os.environ.get("CSX_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-I" + INC, "-I" + CSRC,
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    prep_o = os.path.join(BUILD, "csx_prep.o")
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-Wall", "-I" + INC, "-I" + CSRC, "-c",
          os.path.join(CSRC, "csx_prep.cpp"), "-o", prep_o])
    fit_o = os.path.join(BUILD, "csx_fit.o")
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-Wall", "-pthread", "-I" + INC, "-I" + CSRC, "-c",
          os.path.join(CSRC, "csx_fit.cpp"), "-o", fit_o])
    dev_o = os.path.join(BUILD, "csx_device.o")
    _run(["hipcc"] + HIP_FLAGS + ["-c", os.path.join(CSRC, "csx_device.hip"), "-o", dev_o])
    eng_o = os.path.join(BUILD, "csx_engine.o")
    _run(["hipcc"] + HIP_FLAGS + ["-c", os.path.join(CSRC, "csx_engine.cpp"), "-o", eng_o])
    extra = [os.path.join(CSRC, f) for f in sorted(os.listdir(CSRC))
             if f.startswith(("csx_mlp", "csx_post")) and f.endswith(".hip")]
    objs = [prep_o, fit_o, dev_o, eng_o]
    for src in extra:
        o = os.path.join(BUILD, os.path.basename(src).replace(".hip", ".o"))
        _run(["hipcc"] + HIP_FLAGS + ["-c", src, "-o", o])
        objs.append(o)
    tmp = LIB + ".tmp"
    _run(["hipcc", "--offload-arch=" + ARCH, "-shared", "-fPIC", "-pthread", "-o", tmp] + objs + ["-ldl"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
