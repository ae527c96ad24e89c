"""Build recipe for the gfx950 library ``libdoon.so``.

Compiles the host engine and the HIP kernels with hipcc for gfx950 and links
them into ``doonengine_amd/libdoon.so`` (in-tree, so the built library travels
with the source tree).  Run ``python -m doonengine_amd.build``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
BUILD = PKG / "_build"
LIB = PKG / "libdoon.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Bit-exact float semantics: no contraction into FMA, IEEE division/sqrt (hipcc's
# default), no fast-math. Packed FP32 instructions are disabled.
COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
    f"-I{INCLUDE}", f"-I{CSRC}",
]
DEVICE = [
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]

SOURCES = [
    ("engine.cpp", False),
    ("rcache.cpp", False),
    ("kernels.hip", True),
    ("rcache.hip", True),
    ("edits.hip", True),
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")


def build(verbose: bool = False, defines: list[str] | None = None, out: Path | None = None) -> Path:
    """Compile every source and link libdoon.so (or `out`). Always rebuilds.
    `defines` adds -D options, for variant builds used in measurements."""
    if not Path(HIPCC).exists() and shutil.which("hipcc") is None:
        raise RuntimeError("hipcc not found")
    lib = Path(out) if out else LIB
    build_dir = BUILD if out is None else BUILD / lib.stem
    build_dir.mkdir(parents=True, exist_ok=True)
    extra = [f"-D{d}" for d in (defines or [])]
    objs = []
    for name, is_hip in SOURCES:
        src = CSRC / name
        obj = build_dir / (name + ".o")
        cmd = [HIPCC] + COMMON + extra + (DEVICE if is_hip else ["-x", "c++", "-D__HIP_PLATFORM_AMD__",
                                                                 "-I/opt/rocm/include"]) + ["-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)
        objs.append(str(obj))
    tmp = lib.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp)] + objs + ["-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd)
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
