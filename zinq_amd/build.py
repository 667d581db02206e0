"""In-tree build of the native libraries (run on the CPU host; hipcc cross-compiles for gfx950).

  libzinq_ints.so  host integral engine (g++), C ABI of include/zinq_ints.h
  libzinq_hip.so   HIP kernels + C ABI of include/zinq_hip.h + host SCF driver (hipcc, gfx950)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIP_SOURCES = ["zq_context.hip", "zq_jk.hip", "zq_transform.hip", "zq_dense.hip", "zq_eigen.hip", "zq_scf_dev.hip", "zq_group.hip",
               "zq_ci.hip"]
HOST_SOURCES = ["zq_scf.cpp", "zq_linalg.cpp"]  # plain C++ (g++), linked into libzinq_hip.so
BUILD_DIR = os.path.join(ROOT, "_build")
# Packed-FP32 VALU instructions are disabled for all device code (-packed-fp32-ops).
HIPCC_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-Wall",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-Wno-unused-command-line-argument"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_ints():
    out = os.path.join(HERE, "libzinq_ints.so")
    _run(["g++", "-O3", "-march=x86-64-v3", "-std=c++17", "-fPIC", "-shared", "-pthread", "-Wall",
          "-o", out, os.path.join(CSRC, "ints.cpp")])
    return out


def build_hip():
    out = os.path.join(HERE, "libzinq_hip.so")
    os.makedirs(BUILD_DIR, exist_ok=True)
    objs = []
    for s in HOST_SOURCES:
        o = os.path.join(BUILD_DIR, s.replace(".cpp", ".o"))
        _run(["g++", "-O3", "-march=x86-64-v3", "-std=c++17", "-fPIC", "-Wall", "-c", "-o", o, os.path.join(CSRC, s)])
        objs.append(o)
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    for s in HIP_SOURCES:
        o = os.path.join(BUILD_DIR, s.replace(".hip", ".o"))
        _run([hipcc] + HIPCC_FLAGS + ["-c", "-o", o, os.path.join(CSRC, s)])
        objs.append(o)
    _run([hipcc, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", out] + objs +
         ["-L" + os.path.join(ROCM, "lib"), "-Wl,-rpath," + os.path.join(ROCM, "lib"), "-lrocsolver", "-lrocblas"])
    return out


def build_all():
    build_ints()
    build_hip()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"ints": build_ints, "hip": build_hip, "all": build_all}[w]()
