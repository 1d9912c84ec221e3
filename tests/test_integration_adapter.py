"""The C++ adapter in INTEGRATION.md compiles against the reference's own headers: GpuSymmetricSolver
(ISolver), GpuPredictorCorrector (ITimeStepping) and GpuGravity (IGravity) override the reference's
interfaces with the C ABI's types.  Runs where the reference sources are present (not on the GPU box)."""
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_CORE = "/root/reference/core"


@pytest.mark.skipif(not os.path.isdir(REF_CORE), reason="reference sources not present")
def test_adapter_compiles_against_reference_headers():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```cpp\n(.*?)```", text, re.S)
    assert blocks, "no C++ block in INTEGRATION.md"
    src = "\n".join(blocks) + "\n"
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "adapter.cpp")
        open(path, "w").write(src)
        r = subprocess.run(["g++", "-std=c++14", "-fsyntax-only", "-msse4.1", "-w", "-I", REF_CORE,
                            "-I", os.path.join(ROOT, "include"), path], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
