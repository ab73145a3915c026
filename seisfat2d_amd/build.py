"""Native build of seisfat2d_amd: hipcc for gfx950, g++ for the host library.

Everything is built in-tree so the shared objects travel with the source tree:
  seisfat2d_amd/libseisfat2d_eik.so    HIP kernels + C ABI (include/seisfat2d_eik.h)
  seisfat2d_amd/libseisfat2d_host.so   C++ Modeling / Eikonal_ISO / Eikonal_ANI classes
  seisfat2d_amd/seisfat2d_modeling     the per-shot modeling executable
  seisfat2d_amd/seisfat2d_tables       traveltime tables for migration
  seisfat2d_amd/seisfat2d_inversion    first-arrival tomography
  seisfat2d_amd/seisfat2d_migration    Kirchhoff depth migration (IDKDM, ADKDM, IDLSKDM, ADLSKDM)
"""
from __future__ import annotations

import os
import shutil
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
INCLUDE = os.path.join(ROOT, "include")

EIK_LIB = os.path.join(PKG, "libseisfat2d_eik.so")
HOST_LIB = os.path.join(PKG, "libseisfat2d_host.so")
MODELING_EXE = os.path.join(PKG, "seisfat2d_modeling")
TABLES_EXE = os.path.join(PKG, "seisfat2d_tables")
INVERSION_EXE = os.path.join(PKG, "seisfat2d_inversion")
MIGRATION_EXE = os.path.join(PKG, "seisfat2d_migration")

# Device arithmetic contract: no FMA contraction, correctly rounded fp32
# division and square root (bit parity with the oracle); no packed fp32 math.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # atomics stay single-lane: the optimizer's wave-wide rewrite of the
    # persistent sweeps' ticket claim would consume the result (and wait for
    # every memory operation in flight) right where it is issued
    "-mllvm", "-amdgpu-atomic-optimizer-strategy=None",
    "-Wall", "-Wno-unused-parameter", "-Wno-unused-command-line-argument",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter"]
# the migration driver allocates its trace and model buffers with the HIP runtime
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIP_HOST = ["-D__HIP_PLATFORM_AMD__", "-isystem", os.path.join(ROCM, "include")]
HIP_LINK = ["-L", os.path.join(ROCM, "lib"), "-lamdhip64"]


def _hipcc():
    for p in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if p and os.path.exists(p):
            return p
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + out)
    return out


def build_eik(verbose=False):
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    tmp = EIK_LIB + ".tmp"
    out = _run([_hipcc(), *HIP_FLAGS, "-shared", "-I", INCLUDE, "-o", tmp, *srcs])
    os.replace(tmp, EIK_LIB)
    if verbose and out.strip():
        print(out)
    return EIK_LIB


def build_host(verbose=False):
    srcs = sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp") and not f.endswith("_main.cpp"))
    if not srcs:
        return None
    tmp = HOST_LIB + ".tmp"
    _run(["g++", *HOST_FLAGS, *HIP_HOST, "-fopenmp", "-shared", "-I", INCLUDE, "-o", tmp, *srcs,
          "-L", PKG, "-lseisfat2d_eik", *HIP_LINK, "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, HOST_LIB)
    # one executable per driver: <name>_main.cpp -> seisfat2d_<name>
    for f in sorted(os.listdir(HOST)):
        if f.endswith("_main.cpp"):
            exe = os.path.join(PKG, "seisfat2d_" + f[:-len("_main.cpp")])
            tmp = exe + ".tmp"
            _run(["g++", *HOST_FLAGS, *HIP_HOST, "-fopenmp", "-I", INCLUDE, "-o", tmp, os.path.join(HOST, f), "-L",
                  PKG, "-lseisfat2d_host", "-lseisfat2d_eik", *HIP_LINK, "-Wl,-rpath,$ORIGIN"])
            os.replace(tmp, exe)
    return HOST_LIB


def build_all(verbose=False):
    build_eik(verbose)
    build_host(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
    print("built", EIK_LIB)
