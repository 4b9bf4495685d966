"""The reference-side binding (corr_pc_amd/binding/corr_gpu.cpp): corels2_gpu / corels1_gpu, the
drop-ins a corr_pc maintainer compiles into corr_pc in place of corels2 / corels1 (corr.h:27-30).
It is compiled here against the reference's own data_def.h and jk.h and this repository's cpc.h, so
every reference field and engine entry point it names must exist with the types it uses.  Skipped
where the reference sources are not present."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
SRC = os.path.join(ROOT, "corr_pc_amd", "binding", "corr_gpu.cpp")


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "data_def.h")), reason="reference sources not present")
def test_binding_compiles_against_reference_headers():
    cxx = shutil.which("g++") or shutil.which("c++")
    assert cxx, "no C++ compiler"
    r = subprocess.run([cxx, "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
                        "-I" + REF, SRC], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


def test_binding_names_every_role_of_corels2():
    """The role table of corels2 (corr.cpp:226-271) is written out in full, in the reference's order."""
    src = open(SRC).read()
    body = src[src.index("long corels2_gpu("):]
    order = [body.index(f"if (di.do_{r})") for r in ("DRs", "RsRd", "RsRs", "SRd", "SD", "SRs")]
    assert order == sorted(order)
    assert "std::swap(g1, g2)" in src and "final_jk_bins(bjk, di)" in src
