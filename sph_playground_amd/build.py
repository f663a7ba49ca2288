"""Build libsph_hip.so (hand-written HIP for gfx950) in-tree.

    python -m sph_playground_amd.build

The library is plain hipcc output (no torch extension): the C-ABI in
include/sph_hip.h is what every caller binds.  Packed-FP32 instructions are
disabled for all device code (-packed-fp32-ops); the neighbour test turns FP
contraction off locally so its squared distance rounds as the reference's.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsph_hip.so")
SOURCES = ["sph_kernels.hip", "sph_adjoint.hip", "sph_api.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SPH_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I" + os.path.join(ROOT, "include"),
]


def build(verbose=False, jobs=4):
    """Compile every source to an object (in parallel) and link the .so."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((src, out))
        elif verbose and out.strip():
            print(out)
    if failed:
        for src, out in failed:
            sys.stderr.write("=== %s ===\n%s\n" % (src, out[-8000:]))
        raise RuntimeError("hipcc failed for " + ", ".join(s for s, _ in failed))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_dump(verbose)
    return LIB


# The state dump (include/sph_dump.h) is host code on the HDF5 C library; it
# is built where the library is found and is not part of the GPU path.
HDF5_PREFIX = os.environ.get("SPH_HDF5_PREFIX", "/opt/conda")
DUMP_LIB = os.path.join(HERE, "libsph_dump.so")


def build_dump(verbose=False):
    inc = os.path.join(HDF5_PREFIX, "include")
    lib = os.path.join(HDF5_PREFIX, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        print("libsph_dump.so not built: no hdf5.h under " + inc, file=sys.stderr)
        return None
    tmp = DUMP_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I" + os.path.join(ROOT, "include"),
           "-I" + inc, os.path.join(CSRC, "sph_dump.cpp"), "-o", tmp, "-L" + lib, "-lhdf5",
           "-Wl,-rpath," + lib]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, DUMP_LIB)
    return DUMP_LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
