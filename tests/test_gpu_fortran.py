"""The Fortran drop-in: the reference's subroutine interface
(tinvsubmod::op_tinv, mult_ke, call_twoe_matel00, mult_reducedpot) through
the iso_c_binding shim, exercised by tests/fortran/conformance.f90 built with
amdflang (prebuilt by __graft_entry__.build())."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "tests", "fortran", "bin", "conformance")


def test_fortran_conformance(tmp_path):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not os.path.exists(BIN):
        pytest.skip("conformance binary not built (amdflang absent at build time)")
    print("[gpu-test] fortran conformance", flush=True)
    r = subprocess.run([BIN, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "CONFORMANCE OK" in r.stdout
    lines = (tmp_path / "twoematel.time.dat").read_text().splitlines()
    assert lines[0].split() == ["etc", "1den", "1mpi", "F.T.", "2mpi", "dot", "ALL"]
    assert len(lines) == 3 and all(len(ln.split()) == 14 for ln in lines[1:])
