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


def build_engine(jobs=None):
    """Objects compiled in parallel: the host pipeline and C ABI, the server, and the
    templated kernels once per dimension 1..10; then one shared library."""
    from concurrent.futures import ThreadPoolExecutor

    csrc = os.path.join(HERE, "csrc")
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    units = [(os.path.join(csrc, "clue_engine.hip"), "clue_engine.o", []),
             (os.path.join(csrc, "clue_server.cpp"), "clue_server.o", [])]
    # heaviest instantiations first (the block search exists up to three dimensions)
    for d in (3, 2, 1, 10, 9, 8, 7, 6, 5, 4):
        units.append((os.path.join(csrc, "clue_passes.hip"), f"clue_passes_d{d}.o", [f"-DCLUE_DIM={d}"]))
    jobs = jobs or int(os.environ.get("CLUE_BUILD_JOBS", "0")) or min(8, os.cpu_count() or 1)

    def compile_one(u):
        src, obj, extra = u
        out = os.path.join(objdir, obj)
        run([HIPCC, *HIP_FLAGS, "-pthread", *extra, "-c", "-o", out + ".tmp", src])
        os.replace(out + ".tmp", out)
        return out

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, units))
    out = os.path.join(HERE, "libclue_hip.so")
    tmp = out + ".tmp"
    run([HIPCC, f"--offload-arch={ARCH}", "-pthread", "-shared", "-o", tmp, *objs])
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
