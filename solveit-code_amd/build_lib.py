"""Build recipe of libsolveit_hip.so (hand-written HIP for gfx950).

Every HIP source is compiled with hipcc for gfx950 only and linked into one
shared library next to the Python loader (solveit_hip/libsolveit_hip.so).
Packed FP32 instructions are disabled for device code
(-target-feature -packed-fp32-ops).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "solveit_hip", "libsolveit_hip.so")
OBJDIR = os.path.join(HERE, "_build")

SOURCES = ["context.hip", "direct.hip", "leapfrog.hip", "fmm.hip", "bh.hip", "adjoint.hip", "batch.hip",
           "field_io.hip", "writers.hip", "fmm_sph.hip"]

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-fconstexpr-steps=200000000",  # compile-time M2L term tables
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose=True):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJDIR, exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(OBJDIR, src.replace(".hip", ".o"))
        cmd = [hipcc] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print("+", " ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd)))
        objs.append(obj)
    bad = [s for s, p in procs if p.wait() != 0]
    if bad:
        raise RuntimeError("hipcc failed for: " + ", ".join(bad))
    tmp = OUT + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    sys.exit(0 if build() else 1)
