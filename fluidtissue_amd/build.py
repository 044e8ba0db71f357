"""Builds libftreinit.so (gfx950 HIP + host C++) in-tree with hipcc.

Flags that matter:
  -ffp-contract=off     no implicit fma: the FP64 crossing-point interpolation,
                        the float orientation test and the area sums round every
                        product and sum on its own, as the reference's Eigen code
                        does (src/fem/LevelSet.cpp:292-294, 332-356).
  -fhip-fp32-correctly-rounded-divide-sqrt
                        sqrtf correctly rounded, as std::sqrt in the reference's
                        CPU distance (src/fem/EuclideanDistance.cu:28).
  -Xclang -target-feature -Xclang -packed-fp32-ops
                        no packed FP32 instructions.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libftreinit.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]

SOURCES = ["ft_reinit.hip", "ft_gen.cpp", "ft_obj.cpp", "EuclideanDistance.cpp"]


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build failed: " + " ".join(cmd[:3]) + " ...")
    return out


def build_library(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + HIP_FLAGS + ["-shared", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
                                 "-o", tmp] + srcs + ["-lpthread"]
    out = _run(cmd)
    os.replace(tmp, LIB)
    if verbose and out.strip():
        print(out)
    return LIB


def build_oracle():
    """Builds the CPU checker (oracle/libftoracle.so); test infrastructure."""
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-C", odir, "-B", "libftoracle.so"])
    return os.path.join(odir, "libftoracle.so")


def build_all(verbose=False):
    build_library(verbose)
    build_oracle()


if __name__ == "__main__":
    build_all(verbose=True)
    print("built", LIB)
