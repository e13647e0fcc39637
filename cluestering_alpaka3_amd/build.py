"""Build recipe of the engine library (hipcc, gfx950) and of the oracle.

`python -m cluestering_alpaka3_amd.build` compiles everything in-tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = os.environ.get("CLUE_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: no multiply-add fusion, so every float operation rounds
# exactly as the reference's serial backend writes it.  Packed FP32 is disabled
# as a target feature.
HIP_FLAGS = [
    "-std=c++20", "-O3", f"--offload-arch={ARCH}", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


def run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and "hip-link" not in l)
    if out.strip():
        print(out, flush=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}")


def build_engine():
    srcs = [os.path.join(HERE, "csrc", "clue_engine.hip"), os.path.join(HERE, "csrc", "clue_server.cpp")]
    out = os.path.join(HERE, "libclue_hip.so")
    tmp = out + ".tmp"
    run([HIPCC, *HIP_FLAGS, "-pthread", "-shared", "-o", tmp, *srcs])
    os.replace(tmp, out)
    return out


def build_cpp_tests():
    """The drop-in C++ API test program (g++ only needs clue_abi.h and links the engine)."""
    src = os.path.join(ROOT, "tests", "cpp", "test_api.cpp")
    outdir = os.path.join(ROOT, "tests", "cpp", "bin")
    os.makedirs(outdir, exist_ok=True)
    out = os.path.join(outdir, "test_api")
    run(["g++", "-std=c++20", "-O2", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"), src,
         "-L", HERE, "-lclue_hip", "-Wl,-rpath,$ORIGIN/../../../cluestering_alpaka3_amd", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_bindings():
    """The reference's Python modules (CLUE_Convolutional_Kernels, CLUE_GPU_HIP) as
    pybind11 extensions over the C++ drop-in, in cluestering_alpaka3_amd/pybind/."""
    import sysconfig

    import pybind11

    outdir = os.path.join(HERE, "pybind")
    os.makedirs(outdir, exist_ok=True)
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    common = ["g++", "-std=c++20", "-O2", "-shared", "-fPIC", "-fvisibility=hidden",
              "-I", os.path.join(ROOT, "include"), "-I", pybind11.get_include(),
              "-I", sysconfig.get_paths()["include"]]
    for mod, src in (("CLUE_Convolutional_Kernels", "binding_kernels.cpp"), ("CLUE_GPU_HIP", "binding_hip.cpp")):
        out = os.path.join(outdir, mod + suffix)
        run([*common, os.path.join(HERE, "csrc", src), "-L", HERE, "-lclue_hip",
             "-Wl,-rpath,$ORIGIN/..", "-o", out + ".tmp"])
        os.replace(out + ".tmp", out)


def build_oracle():
    run(["make", "-C", os.path.join(ROOT, "oracle"), "-s"])


def build_all():
    build_oracle()
    build_engine()
    build_cpp_tests()
    build_bindings()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "engine": build_engine, "oracle": build_oracle, "cpp": build_cpp_tests,
         "bindings": build_bindings}[w]()
