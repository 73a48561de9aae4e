"""Build the CPU restatement ``oracle/libuph_cpu.so`` (TEST INFRASTRUCTURE:
the CPU baseline of bench.py and its tests).  gcc, -O3, OpenMP; the
instruction set is x86-64-v3 (AVX2 + FMA) so that the library built here
runs on any current x86 host."""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "cpu_fit.cpp"
LIB = HERE / "libuph_cpu.so"        # float32 pixel arithmetic (the baseline)
LIB64 = HERE / "libuph_cpu64.so"     # the same source in float64 (its check)
FLAGS = ["-O3", "-march=x86-64-v3", "-mtune=generic", "-fcx-limited-range", "-fopenmp",
         "-std=c++17", "-fPIC",
         "-shared"]


def build() -> Path:
    for lib, extra in ((LIB, []), (LIB64, ["-DUPHC_REAL=double"])):
        tmp = lib.with_suffix(f".so.tmp{os.getpid():x}")   # one per process: parallel test workers
        cmd = [os.environ.get("CXX", "g++"), *FLAGS, *extra, str(SRC), "-o", str(tmp)]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"g++ failed ({res.returncode}):\n{res.stderr}")
        os.replace(tmp, lib)
    return LIB


if __name__ == "__main__":
    print(build())
