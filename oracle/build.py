"""Build the CPU oracle ``oracle/liboracle_sph.so`` (test infrastructure).

f32 arithmetic in the reference's operation order needs no FMA contraction
and no fast-math; x86-64 uses SSE scalar floats, so there is no excess
precision.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_sph.so")
SRCS = [os.path.join(HERE, "sph_oracle.cpp"), os.path.join(HERE, "sph2d_oracle.cpp")]


def build(verbose=False):
    tmp = LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
           "-fopenmp", *SRCS, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
