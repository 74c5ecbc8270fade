"""The drop-in binding printed in section 1 of the integration guide, compiled
against the reference's own driver: the reference's main.cpp and FishFry.hpp
(which instantiate FishFry<ParisPencil>, main.cpp:48) are compiled for gfx950
with their ParisPencil.hpp replaced by that binding stub,
through a clang file overlay (the reference tree is read, never modified or
copied).  Runs only where a reference checkout and an MPI header exist (the
build container); syntax and template instantiation only, no link, no GPU."""
import json
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
MPI_INC = "/opt/conda/include"
HIPCC = "/opt/rocm/bin/hipcc"


def stub_text():
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    m = re.search(r"binding stub:\s*```cpp\n(.*?)```", doc, re.S)
    assert m, "INTEGRATION.md no longer prints the binding stub"
    return m.group(1)


@pytest.mark.skipif(not (os.path.exists(os.path.join(REF, "main.cpp")) and os.path.exists(
    os.path.join(MPI_INC, "mpi.h")) and os.path.exists(HIPCC)), reason="needs the reference tree, mpi.h and hipcc")
def test_reference_driver_compiles_against_the_stub(tmp_path):
    stub = tmp_path / "ParisPencil_stub.hpp"
    stub.write_text(stub_text())
    overlay = tmp_path / "overlay.json"
    overlay.write_text(json.dumps({"version": 0, "case-sensitive": "true", "roots": [
        {"name": REF, "type": "directory",
         "contents": [{"name": "ParisPencil.hpp", "type": "file", "external-contents": str(stub)}]}]}))
    cmd = [HIPCC, "-x", "hip", "--offload-arch=gfx950", "-std=c++17", "-DO_HIP", "-fsyntax-only",
           "-ivfsoverlay", str(overlay), "-I", ROOT, "-I", MPI_INC, os.path.join(REF, "main.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    # the overlay is what was compiled: a stub that fails to compile makes the build fail
    stub.write_text('#error "binding stub in use"\n')
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode != 0 and "binding stub in use" in r.stderr
