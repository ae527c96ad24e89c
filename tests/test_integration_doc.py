"""The C examples in INTEGRATION.md compile against include/doon/doon.h (C11,
syntax and types only: nothing runs), so the documented binding stays true."""
from __future__ import annotations

import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def c_blocks():
    text = (ROOT / "INTEGRATION.md").read_text()
    return re.findall(r"```c\n(.*?)```", text, flags=re.S)


@pytest.mark.skipif(shutil.which("gcc") is None or not Path("/opt/rocm/include/hip").exists(),
                    reason="needs gcc and the HIP headers")
def test_integration_examples_compile(tmp_path):
    blocks = c_blocks()
    assert len(blocks) >= 3
    for i, code in enumerate(blocks):
        if "int main" in code:
            src = "#include <stdio.h>\n" + code
        else:
            src = ("#include <doon/doon.h>\n#include <hip/hip_runtime_api.h>\n"
                   "void snippet(DNvolume* vol, float t, size_t maxRequests)\n{\n" + code + "\n}\n")
        f = tmp_path / f"example{i}.c"
        f.write_text(src)
        r = subprocess.run(["gcc", "-std=c11", "-fsyntax-only", "-Wall", "-Wno-unused-variable",
                            "-Wno-unused-but-set-variable", "-D__HIP_PLATFORM_AMD__", f"-I{ROOT / 'include'}",
                            "-I/opt/rocm/include", str(f)], capture_output=True, text=True)
        assert r.returncode == 0, f"INTEGRATION.md C block {i} does not compile:\n{r.stderr}"
