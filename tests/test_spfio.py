"""Orbital file format (S/loadstuff.f90): Python writer/reader against the
Fortran runtime (amdflang) and the reference's centred chop/pad rules."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from sincdvr import spfio

HERE = os.path.dirname(os.path.abspath(__file__))


def _flang():
    for c in ("amdflang", "/opt/rocm/llvm/bin/amdflang", "flang"):
        p = shutil.which(c) or (c if os.path.exists(c) else None)
        if p:
            return p
    return None


@pytest.fixture(scope="module")
def fprog(tmp_path_factory):
    fc = _flang()
    if fc is None:
        pytest.skip("no Fortran compiler")
    d = tmp_path_factory.mktemp("spfio")
    exe = str(d / "spfio_check")
    subprocess.check_call([fc, "-O0", os.path.join(HERE, "fortran", "spfio_check.f90"), "-o", exe], cwd=str(d))
    return exe


def test_read_fortran_written_file(fprog, tmp_path):
    f = str(tmp_path / "spfs.bin")
    out = subprocess.run([fprog, "write", f], capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    dims, num, cflag = spfio.read_header(f)
    assert dims == (3, 4, 5) and num == 2 and cflag == 1
    a, k = spfio.read_spfs(f, (3, 4, 5), 2)
    i = np.arange(1, 61)
    for j in (1, 2):
        assert np.array_equal(a[:, j - 1], (i + 100 * j) - 1j * i)
    assert k == 2


def test_fortran_reads_python_written_file(fprog, tmp_path):
    f = str(tmp_path / "spfs.bin")
    i = np.arange(1, 61)
    a = np.stack([(i + 1000 * j) + 1j * (i / 2) for j in (1, 2)], axis=1)
    spfio.write_spfs(f, a, (3, 4, 5))
    out = subprocess.run([fprog, "read", f], capture_output=True, text=True)
    assert out.returncode == 0 and "spfio ok" in out.stdout, (out.returncode, out.stdout, out.stderr)


def test_round_trip_and_counts(tmp_path):
    f = str(tmp_path / "o.bin")
    rng = np.random.default_rng(1)
    a = rng.standard_normal((5 * 5 * 5, 3)) + 1j * rng.standard_normal((5 * 5 * 5, 3))
    spfio.write_spfs(f, a, (5, 5, 5))
    b, k = spfio.read_spfs(f, (5, 5, 5), 3)
    assert np.array_equal(a, b) and k == 3
    b, k = spfio.read_spfs(f, (5, 5, 5), 5)            # more requested than stored
    assert k == 3 and np.array_equal(b[:, :3], a) and not b[:, 3:].any()
    b, k = spfio.read_spfs(f, (5, 5, 5), 2)
    assert k == 2 and np.array_equal(b, a[:, :2])
    r = rng.standard_normal((125, 2))                   # real file (cflag 0)
    spfio.write_spfs(f, r, (5, 5, 5))
    assert spfio.read_header(f)[2] == 0
    b, _ = spfio.read_spfs(f, (5, 5, 5), 2)
    assert np.array_equal(b.real, r) and not b.imag.any()


def test_centred_chop_and_pad(tmp_path):
    """spfdimtype 1: a value at the grid centre stays at the centre when the
    grid grows or shrinks by an even count; odd/even mismatch drops the odd
    grid's last point (loadstuff.f90:339-357)."""
    f = str(tmp_path / "c.bin")
    n = 7
    g = np.zeros((n, n, n))
    g[3, 3, 3] = 1.0           # centre of 7
    g[6, 3, 3] = 2.0           # positive end of x
    g[0, 3, 3] = 3.0           # negative end of x
    spfio.write_spfs(f, g.reshape(-1, 1, order="F").astype(complex), (n, n, n))
    big, _ = spfio.read_spfs(f, (11, 11, 11), 1)
    B = big.reshape(11, 11, 11, order="F")
    assert B[5, 5, 5] == 1 and B[8, 5, 5] == 2 and B[2, 5, 5] == 3 and np.count_nonzero(B) == 3
    small, _ = spfio.read_spfs(f, (5, 5, 5), 1)
    S = small.reshape(5, 5, 5, order="F")
    assert S[2, 2, 2] == 1 and np.count_nonzero(S) == 1           # both ends chopped
    even, _ = spfio.read_spfs(f, (6, 6, 6), 1)
    E = even.reshape(6, 6, 6, order="F")
    # 7 -> 6: the odd grid loses its positive-end point, nothing is shifted
    assert E[3, 3, 3] == 1 and E[0, 3, 3] == 3 and np.count_nonzero(E) == 2
    with pytest.raises(ValueError):
        spfio.read_spfs(f, (6, 7, 7), 1, dimtypes=(2, 1, 1))
