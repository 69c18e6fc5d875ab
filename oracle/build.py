"""Build recipe of the FP64 oracle: liboracle.so from treams_oracle.c (gcc, -O2).

The oracle is test infrastructure (see oracle.py).  The reference itself (Rust,
cargo) cannot be built in this image, so there is no oracle/_ref build.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def build(verbose: bool = False) -> str:
    src = os.path.join(HERE, "treams_oracle.c")
    out = os.path.join(HERE, "liboracle.so")
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-o", out, src, "-lm"]
    if verbose:
        print(" ".join(cmd))
    subprocess.check_call(cmd)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
