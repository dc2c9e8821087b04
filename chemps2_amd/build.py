"""Builds libchemps2_amd.so in-tree with hipcc for gfx950."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libchemps2_amd.so")
SOURCES = ["kernels.hip", "learned.hip", "layout.cpp", "bookkeeper.cpp", "heff_plan.cpp", "update_plan.cpp", "dmrg.cpp", "excited.cpp", "twodm.cpp", "context.cpp", "noalloc.hip"]
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, "_obj_" + src.replace(".", "_") + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [hipcc] + FLAGS[:-1] + ["-c"] + lang + [os.path.join(CSRC, src), "-o", obj]
        cmd.remove("-shared")
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB] + objs + [
        "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.run(cmd, check=True)
    for o in objs:
        os.remove(o)
    build_io(verbose)
    return LIB


HDF5 = "/opt/conda"
IO_LIB = os.path.join(HERE, "libchemps2_amd_io.so")


def build_io(verbose=False):
    """libchemps2_amd_io.so: MPS checkpoints in the reference's HDF5 layout, on top of the
    C ABI of libchemps2_amd (host code only). Skipped when no HDF5 is installed."""
    if not os.path.exists(os.path.join(HDF5, "include", "hdf5.h")):
        print("chemps2_amd: no HDF5 under " + HDF5 + "; the checkpoint library is not built", file=sys.stderr)
        return None
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", IO_LIB, os.path.join(CSRC, "mpsio.cpp"),
           "-I" + os.path.join(HDF5, "include"), "-L" + os.path.join(HDF5, "lib"), "-lhdf5", "-L" + HERE, "-lchemps2_amd",
           # the system's libstdc++ (which ROCm needs) before the older one beside libhdf5
           "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/usr/lib/x86_64-linux-gnu", "-Wl,-rpath," + os.path.join(HDF5, "lib")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return IO_LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
