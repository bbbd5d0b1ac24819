"""Build recipe of libplmregisterhip.so (hipcc, gfx950, in-tree)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libplmregisterhip.so")

SOURCES = ["bspline_kernels.hip", "lbfgs_kernels.hip", "stage_kernels.hip", "noalloc.hip", "bspline_abi.cpp",
           "bspline_lbfgs.cpp", "stage_abi.cpp"]

# -ffp-contract=off: the score path reproduces the reference's float rounding
# operation by operation; FMAs are written explicitly where they are wanted.
# The packed-FP32 feature is switched off for every kernel.  The matrix-core
# chains keep their accumulators in ordinary VGPRs (the AGPR form would add
# to the unified register count and cost the tile kernel a wave per SIMD).
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-mllvm", "-amdgpu-mfma-vgpr-form=1",
         "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    cmd = [hipcc] + FLAGS + [os.path.join(CSRC, s) for s in SOURCES] + ["-o", OUT + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    os.replace(OUT + ".tmp", OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
