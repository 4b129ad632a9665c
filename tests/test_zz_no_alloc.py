"""The allocation-free op_tinv (sdv_op_tinv_noalloc_dev) against the allocating
one (sdv_op_tinv_dev), both called through the C-ABI on the same device
densities: bitwise equal outputs, on two calls in a row with the second
smaller, in both precisions and flavours, including grids whose call is cut
into several chunks of the 128 MiB scratch."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _densities(torch, n, k, seed):
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((k, n ** 3)) + 1j * rng.standard_normal((k, n ** 3))
    return torch.from_numpy(d).to("cuda")


def _run(torch, s, den, noalloc):
    from sincdvr import _lib
    lib = _lib.load()
    out = torch.full_like(den, complex(np.nan, np.nan))  # every element must be written
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    fn = lib.sdv_op_tinv_noalloc_dev if noalloc else lib.sdv_op_tinv_dev
    _lib.check(fn(s._h, ctypes.c_void_p(den.data_ptr()), ctypes.c_void_p(out.data_ptr()), den.shape[0],
                  den.shape[0], st))
    torch.cuda.synchronize()
    return out


# (grid edge, spacing, densities of the first call, of the second call).
# At 81^3 one FP64 convolution needs 17 MB of work memory, so the 128 MiB
# scratch takes 7 densities per chunk: 9 densities run as two chunks.
CASES = [(41, 0.3, 12, 5), (81, 0.25, 9, 3), (21, 0.4, 16, 1)]


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("cnorm", [False, True])
@pytest.mark.parametrize("n,dx,k1,k2", CASES)
def test_noalloc_op_tinv_matches_op_tinv(torch_dev, n, dx, k1, k2, cnorm, precision):
    from sincdvr import SincDVR
    torch = torch_dev
    s = SincDVR(n, dx, 1, cnorm=cnorm, precision=precision)
    for call, k in enumerate((k1, k2)):
        say("op_tinv noalloc n=%d allsize=%d cnorm=%s precision=%d call %d" % (n, k, cnorm, precision, call))
        den = _densities(torch, n, k, 1000 * n + 10 * call + precision)
        ref = _run(torch, s, den, noalloc=False)
        got = _run(torch, s, den, noalloc=True)
        assert not torch.isnan(torch.view_as_real(got)).any()
        assert torch.equal(got, ref), float((got - ref).abs().max())


def test_noalloc_op_tinv_size_error_and_debugflag(torch_dev):
    from sincdvr import SdvError, SincDVR, _lib
    torch = torch_dev
    lib = _lib.load()
    n = 21
    s = SincDVR(n, 0.4, 1)
    den = _densities(torch, n, 3, 7)
    out = torch.zeros_like(den)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    say("op_tinv noalloc size error")
    with pytest.raises(SdvError) as e:
        _lib.check(lib.sdv_op_tinv_noalloc_dev(s._h, ctypes.c_void_p(den.data_ptr()),
                                               ctypes.c_void_p(out.data_ptr()), 3, 2, st))
    assert "SIZE ERROR OP_TINV" in str(e.value)
    say("op_tinv noalloc debugflag 4242")
    z = SincDVR(n, 0.4, 1, debugflag=4242)
    out = torch.full_like(den, 1.0 + 1.0j)
    _lib.check(lib.sdv_op_tinv_noalloc_dev(z._h, ctypes.c_void_p(den.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                           3, 3, st))
    torch.cuda.synchronize()
    assert not out.abs().max().item()
