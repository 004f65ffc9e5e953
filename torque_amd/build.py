"""Build the HIP library (libtorque_hip.so, gfx950) and the CPU oracle.

The HIP sources under torque_amd/csrc are compiled with hipcc for gfx950
only, each to an object in torque_amd/_build/, then linked into
torque_amd/libtorque_hip.so. The oracle (test infrastructure) is built by
oracle/Makefile into oracle/liboracle.so.
"""
import concurrent.futures
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "torque_amd")
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libtorque_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-command-line-argument",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


# Per-source additions. The wave-tile GEMM keeps its MFMA accumulators in
# VGPRs (AGPR forms make the compiler copy them across the k loop).
FILE_FLAGS = {
    "tq_wave_gemm.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
}


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build_hip(jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    headers = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    headers.append(os.path.join(ROOT, "include", "torque_hip.h"))
    newest_header = max(os.path.getmtime(h) for h in headers)

    def compile_one(src):
        s = os.path.join(CSRC, src)
        o = os.path.join(BUILD, src[:-4] + ".o")
        if (os.path.exists(o) and os.path.getmtime(o) > os.path.getmtime(s)
                and os.path.getmtime(o) > newest_header):
            return o
        _run([HIPCC] + HIP_FLAGS + FILE_FLAGS.get(os.path.basename(s), []) + ["-c", s, "-o", o])
        return o

    jobs = jobs or min(8, os.cpu_count() or 4)
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "liboracle.so")


CPP_TEST = os.path.join(ROOT, "tests", "cpp", "test_dropin")


def build_cpp_tests():
    """The C++ drop-in test program (host code only: g++ against the
    headers in torque_amd/include and the C-ABI library)."""
    src = CPP_TEST + ".cpp"
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"),
          "-I" + os.path.join(PKG, "include"), src, "-o", CPP_TEST,
          "-L" + PKG, "-ltorque_hip", "-Wl,-rpath,$ORIGIN/../../torque_amd"])
    return CPP_TEST


def build_all():
    lib = build_hip()
    orc = build_oracle()
    cpp = build_cpp_tests()
    return lib, orc, cpp


if __name__ == "__main__":
    for p in build_all():
        print(p)
    sys.exit(0)
