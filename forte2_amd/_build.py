"""Compile the native libraries of forte2_amd in-tree.

* ``libf2jk.so``  -- HIP J/K kernels + C ABI (``include/f2jk.h``), gfx950 only.
* ``libf2ints.so`` -- host integral engine (``include/f2ints.h``), g++/OpenMP.
* ``libf2probe.so`` -- MFMA and HBM peak probes (``include/f2probe.h``), used
  by the benchmark only.

Both are plain shared objects loaded with ctypes; nothing is JIT-compiled at
import time.  ``python -m forte2_amd._build`` rebuilds both.
"""

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
INCLUDE = os.path.join(ROOT, "include")
CSRC = os.path.join(PKG, "csrc")

HIP_ARCH = "gfx950"
HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    f"--offload-arch={HIP_ARCH}",
    # packed-FP32 instructions are kept out of every kernel
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    # per-thread staging arrays stay in registers: with LDS to spare the
    # compiler would otherwise move them into shared memory
    "-mllvm",
    "-disable-promote-alloca-to-lds",
]
GXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-march=x86-64-v3"]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build failed: {' '.join(cmd)}")
    return res


def build_jk():
    out = os.path.join(PKG, "libf2jk.so")
    src = os.path.join(CSRC, "f2jk.hip")
    tmp = out + ".tmp"
    _run([_hipcc(), *HIPCC_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", src, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_probe():
    out = os.path.join(PKG, "libf2probe.so")
    src = os.path.join(CSRC, "probe.hip")
    tmp = out + ".tmp"
    # accumulators in VGPRs: in the AGPR form the compiler copies the loop-carried
    # accumulators between the register files on every round of the probe loop
    _run([_hipcc(), *HIPCC_FLAGS, "-mllvm", "-amdgpu-mfma-vgpr-form", f"-I{INCLUDE}", src, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_ints():
    out = os.path.join(PKG, "libf2ints.so")
    srcdir = os.path.join(CSRC, "ints")
    srcs = sorted(os.path.join(srcdir, f) for f in os.listdir(srcdir) if f.endswith(".cc"))
    if not srcs:
        return None
    tmp = out + ".tmp"
    _run(["g++", *GXX_FLAGS, f"-I{INCLUDE}", f"-I{srcdir}", *srcs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_all():
    return [p for p in (build_ints(), build_jk(), build_probe()) if p]


if __name__ == "__main__":
    for p in build_all():
        print(p)
