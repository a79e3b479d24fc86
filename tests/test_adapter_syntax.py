"""The reference-side binding (srsoriginal_amd/adapter/OSCARSSR_Cuda_hip.cc) compiles
against the reference's own headers: its four functions keep the signatures of
include/OSCARSSR_Cuda.h:25-28 (a mismatch is a redeclaration error).  CPU only; skipped
where the reference sources are not present."""
import os
import subprocess

import pytest

from conftest import ROOT

REF_INCLUDE = "/root/reference/include"


@pytest.mark.skipif(not os.path.isdir(REF_INCLUDE), reason="reference headers not present")
def test_adapter_compiles_against_reference_headers():
    src = os.path.join(ROOT, "srsoriginal_amd", "adapter", "OSCARSSR_Cuda_hip.cc")
    r = subprocess.run(["g++", "-std=c++11", "-pedantic", "-Wall", "-Werror", "-fsyntax-only", "-I" + REF_INCLUDE,
                        "-I" + os.path.join(ROOT, "include"), src], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.skipif(not os.path.isdir(REF_INCLUDE), reason="reference headers not present")
def test_adapter_defines_every_reference_gpu_function():
    import re
    hdr = open(os.path.join(REF_INCLUDE, "OSCARSSR_Cuda.h")).read()
    want = set(re.findall(r"\b(OSCARSSR_Cuda_\w+)\s*\(", hdr))
    src = open(os.path.join(ROOT, "srsoriginal_amd", "adapter", "OSCARSSR_Cuda_hip.cc")).read()
    have = set(re.findall(r'extern "C" \w+ (OSCARSSR_Cuda_\w+)\(', src))
    assert want and want == have, (want, have)
