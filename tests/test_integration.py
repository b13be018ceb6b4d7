"""The reference-side binding of the integration guide (its ```cpp file=...```
blocks) compiles against the reference's own headers (syntax check; skipped
where the tree is absent)."""
import os
import re
import subprocess
import tempfile

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"


def cpp_blocks():
    txt = open(os.path.join(REPO, "INTEGRATION.md")).read()
    return re.findall(r"```cpp file=(\S+)\n(.*?)```", txt, flags=re.S)


def test_integration_has_binding_files():
    names = [n for n, _ in cpp_blocks()]
    assert "src/linalg/linalg_hip.cpp" in names and "src/gw/sigma_c_hip.cpp" in names


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "include")), reason="reference tree absent")
@pytest.mark.parametrize("name,code", cpp_blocks())
def test_binding_compiles_against_reference_headers(name, code):
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, os.path.basename(name))
        open(src, "w").write(code)
        res = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-Wall",
                              "-I" + os.path.join(REF, "include"), "-I" + os.path.join(REPO, "include"),
                              src], capture_output=True, text=True)
        assert res.returncode == 0, res.stderr
