"""CPU-side checks of the drop-in boundary: the HIP library loads and exports
every symbol include/sincdvr.h declares, init fails cleanly without a GPU,
the host mirror's argument handling, and the synthetic input generator."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "sincdvr.h")


def _declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(sdv_\w+)\s*\(", txt, re.M)))


def test_header_and_python_binding_agree():
    from sincdvr import _lib
    assert _declared() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    from sincdvr import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libsincdvr.so is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name


def test_init_without_gpu_is_a_clean_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from sincdvr import SdvError, SincDVR
    with pytest.raises(SdvError) as e:
        SincDVR(21, 0.4, 4)
    assert e.value.code == -5


def test_config_defaults_match_reference_params():
    """SD/sinc_params.f90: toothnsmall=40 (:32), toothnbig=240 (:33),
    spacing=0.25 (:41), numpoints=15 (:40), scalingdistance=10000 (:55),
    smoothness=5 (:56), scalingstretch=1 (:58)."""
    from sincdvr import _lib
    lib = _lib.load()
    c = _lib.SdvConfig()
    lib.sdv_config_default(ctypes.byref(c))
    assert (c.toothnsmall, c.toothnbig, c.numpoints) == (40, 240, 15)
    assert (c.spacing, c.scalingdistance, c.smoothness, c.scalingstretch) == (0.25, 10000.0, 5.0, 1.0)
    assert (c.cnormflag, c.scalingflag, c.notwoflag, c.debugflag, c.precision) == (0, 0, 0, 0, 0)


@pytest.mark.parametrize("cnorm", [False, True])
def test_gaussian_orbitals_orthonormal_and_deterministic(cnorm):
    from sincdvr import inputs
    a = inputs.gaussian_orbitals(9, 0.4, 3, 2, seed=5, cnorm=cnorm)
    b = inputs.gaussian_orbitals(9, 0.4, 3, 2, seed=5, cnorm=cnorm)
    assert np.array_equal(a, b)
    S = a.T @ a if cnorm else a.conj().T @ a
    assert np.abs(S - np.eye(3)).max() < 1e-13
    assert a.flags["F_CONTIGUOUS"] and a.shape == (729, 3)


def test_grid_points_follow_sinedvr(orc):
    from sincdvr import inputs
    _, _, pts = orc.sinedvr(21, 0.4)
    assert np.array_equal(inputs.grid_points(21, 0.4), pts)


def test_splitmix_reference_values():
    """SplitMix64 (Steele, Lea, Flood 2014) first outputs for seed 0."""
    from sincdvr.inputs import SplitMix64
    r = SplitMix64(0)
    assert [r.next_u64() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
