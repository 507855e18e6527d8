"""Builds oracle/liboracle.so from axhel_oracle.c with gcc (test infrastructure)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "axhel_oracle.c")
LIB = os.path.join(HERE, "liboracle.so")


def build(verbose=False):
    # -O3 with the x86-64-v3 ISA level (AVX2, BMI2 mulx for the 128-bit
    # products): the Native-backend flags of the reference's Release build
    # (-O3 -mcpu=native) without tying the shipped library to this host's CPU.
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-std=gnu11", "-Wall", "-Wextra", "-shared", "-fPIC",
           "-o", LIB + ".tmp", SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
