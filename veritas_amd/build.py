"""Build the MI355X libraries in-tree (no GPU needed: hipcc cross-compiles).

    veritas_amd/lib/libvx_kernels.so  HIP kernels + the C ABI (include/veritas_amd.h)
    veritas_amd/lib/libveritas.so     C++ host classes (SolverManager, Mesh, ...)

All HIP code is built for gfx950 with floating-point contraction disabled
(the FP64 path is bit-compared with the CPU oracle) and without packed FP32.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
OBJ = os.path.join(HERE, "lib", "obj")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("VX_ARCH", "gfx950")

HIP_SOURCES = ["vx_runtime.hip", "vx_fct.hip", "vx_stage.hip", "vx_moments.hip", "vx_field.hip",
               "vx_amr.hip"]
HOST_SOURCES = ["host/veritas.cpp", "host/amr.cpp", "host/output.cpp", "host/checkpoint.cpp", "host/graph.cpp", "host/serve.cpp",
                "host/host_capi.cpp"]

HIP_FLAGS = [f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-Wall",
              "-Wno-sign-compare"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + out + "\n")
        raise RuntimeError("build step failed: " + os.path.basename(cmd[-1]))
    return out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        o = os.path.join(OBJ, src.replace(".hip", ".o"))
        objs.append(o)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", o])
    with ThreadPoolExecutor(max_workers=min(4, os.cpu_count() or 1)) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    kern = os.path.join(LIB, "libvx_kernels.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", kern])
    host = os.path.join(LIB, "libveritas.so")
    _run(["g++", *HOST_FLAGS, "-shared", *[os.path.join(CSRC, s) for s in HOST_SOURCES],
          f"-L{LIB}", "-lvx_kernels", "-lpthread", "-Wl,-rpath,$ORIGIN", "-o", host])
    return kern, host


if __name__ == "__main__":
    print(build(verbose=True))
