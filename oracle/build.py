"""Builds the oracle shared libraries (test infrastructure only).

liboracle_f32.so: strict-IEEE FP32 restatement (the bit-exact checker).
liboracle_f64.so: the same algorithm in FP64 (the accuracy checker).
Run ``python oracle/build.py``.
"""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "doon_oracle.c"
FLAGS = ["-O2", "-fPIC", "-shared", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-fexcess-precision=standard",
         "-Wall", "-Wextra", "-Wno-unused-function"]


REFERENCE_SHADERS = Path("/root/reference/assets/shaders")


def build() -> list[Path]:
    outs = []
    for name, extra in (("liboracle_f32.so", []), ("liboracle_f64.so", ["-DOR_DOUBLE"])):
        out = HERE / name
        cmd = ["gcc"] + FLAGS + extra + [str(SRC), "-o", str(out), "-lm"]
        subprocess.run(cmd, check=True)
        outs.append(out)
    # The llvmpipe harness that runs the reference's own shaders (read at run
    # time from the reference tree) is only useful where that tree exists.
    if REFERENCE_SHADERS.exists() and Path("/usr/include/GL/internal/dri_interface.h").exists():
        ref = HERE / "_ref"
        ref.mkdir(exist_ok=True)
        out = ref / "libref_llvmpipe.so"
        subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-Wall", "-Wno-missing-field-initializers",
                        str(HERE / "ref_llvmpipe.c"), "-o", str(out), "-ldl"], check=True)
        outs.append(out)
    return outs


if __name__ == "__main__":
    for p in build():
        print(p)
    sys.exit(0)
