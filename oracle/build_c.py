"""Build the C oracle (oracle/c/*.c) into oracle/_build/libmgpcg.so with gcc.

TEST INFRASTRUCTURE: the shared object is loaded only by tests and by
bench.py's CPU-baseline leg.  x86-64-v3 (AVX2/FMA) keeps the binary portable
between this container and the GPU host; contraction is off so results follow
the source's operation order.
"""

from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libmgpcg.so")
SOURCES = [os.path.join(HERE, "c", "mgpcg.c")]


def build(verbose: bool = True) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared", "-std=c11",
           "-o", LIB + ".tmp"] + SOURCES + ["-lm"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"gcc failed:\n{res.stdout}\n{res.stderr}")
    os.replace(LIB + ".tmp", LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
