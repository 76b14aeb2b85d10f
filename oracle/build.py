"""ORACLE / TEST INFRASTRUCTURE: builds oracle/librt_twin.so (the CPU restatement) with gcc.

The reference raycaster is Python over pyopencl with a stubbed ray cast
(rendering/_raycaster.py:30-36), so there is no reference C/C++ to compile into oracle/_ref;
the oracle is this restatement plus its FP64 layer.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "librt_twin.so")

# -ffp-contract=off: GCC contracts a*b+c by default in GNU C mode; the twin must evaluate exactly
# the operations written (explicit fmaf only). x86-64-v3 (AVX2+FMA) keeps fmaf a single
# instruction while still running on any current x86 server host.
CFLAGS = ["-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
          "-fPIC", "-shared", "-std=gnu11", "-Wall", "-Wno-unused-function"]


def build_oracle(verbose=False):
    src = os.path.join(HERE, "rt_twin.c")
    tmp = LIB + ".tmp"
    cmd = ["gcc", *CFLAGS, src, "-o", tmp, "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"gcc failed ({r.returncode}):\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_oracle(verbose=True))
