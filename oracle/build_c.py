"""Compile the C restatement of the oracle (oracle/cle_oracle_c.c) into
oracle/lib/liboracle.so with gcc -O3 -fopenmp (test infrastructure / CPU baseline).

The reference itself (Python + OpenCL kernels, most of them in an absent submodule)
has no C sources to compile, so there is no oracle/_ref build.
"""
import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "cle_oracle_c.c"
LIB = HERE / "lib" / "liboracle.so"


def build(verbose: bool = True) -> Path:
    LIB.parent.mkdir(exist_ok=True)
    tmp = LIB.with_suffix(".so.tmp")
    march = os.environ.get("ORACLE_MARCH", "-march=x86-64-v3")
    cmd = ["gcc", "-O3", march, "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared", "-std=c11",
           "-Wall", "-Wno-unknown-pragmas", str(SRC), "-o", str(tmp), "-lm"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{' '.join(cmd)}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
