"""CPU check: the Fortran shim and the conformance program compile and link
against libsincdvr.so, and the shim exports the reference's symbols."""
import os
import subprocess

import pytest

from conftest import ROOT

BIN = os.path.join(ROOT, "tests", "fortran", "bin", "conformance")
SHIM_O = os.path.join(ROOT, "tests", "fortran", "bin", "sincdvr_shim.o")


def test_shim_exports_reference_symbols():
    if not os.path.exists(SHIM_O):
        pytest.skip("shim not built (amdflang absent)")
    out = subprocess.run(["nm", SHIM_O], capture_output=True, text=True).stdout
    for sym in ("mult_ke_", "call_twoe_matel00_", "mult_reducedpot_", "sincdvr_gpu_init_"):
        assert any(line.split()[-1] == sym for line in out.splitlines() if line.strip()), sym
    assert "tinvsubmod" in out and "op_tinv" in out


def test_conformance_binary_links():
    if not os.path.exists(BIN):
        pytest.skip("conformance binary not built (amdflang absent)")
    out = subprocess.run(["ldd", BIN], capture_output=True, text=True).stdout
    assert "libsincdvr.so" in out and "not found" not in out.split("libsincdvr.so")[1].splitlines()[0]
