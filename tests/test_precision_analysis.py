"""CPU analysis behind the precision form: with FP32 butterflies the error of
a circulant Coulomb convolution is set by the butterfly arithmetic, not by the
rounding of the density or of the kernel spectrum to FP32.  Splitting either
into hi + lo FP32 parts (and convolving the parts separately) leaves the error
where it was; only FP64 arithmetic (or compensated arithmetic, which costs
more than native FP64 on a GPU whose FP64 and unpacked FP32 vector rates are
equal) removes it.  numpy's FFT computes complex64 input in single precision."""
import numpy as np
import pytest


def _circulant_spectrum(K, n, M):
    Kc = np.zeros((M, M, M))
    idx = np.r_[np.arange(0, n), np.arange(M - (n - 1), M)]
    src = np.r_[np.arange(n - 1, 2 * n - 1), np.arange(0, n - 1)]
    Kc[np.ix_(idx, idx, idx)] = K[np.ix_(src, src, src)]
    return np.fft.fftn(Kc).real


def _conv(rho, khat, n, M, single):
    ct, rt = (np.complex64, np.float32) if single else (np.complex128, np.float64)
    R = np.zeros((M, M, M), dtype=ct)
    R[:n, :n, :n] = rho.astype(ct)
    F = (np.fft.fftn(R) * khat.astype(rt)).astype(ct)
    return np.fft.ifftn(F)[:n, :n, :n].astype(np.complex128)


def test_hi_lo_split_does_not_reduce_fp32_fft_error(orc):
    from sincdvr import inputs
    n, dx = 15, 0.4
    M = 2 * n
    phi = inputs.gaussian_orbitals(n, dx, 3, 3, 123)
    khat = _circulant_spectrum(orc.coulomb_kernel(n, dx), n, M)
    k_hi = khat.astype(np.float32).astype(np.float64)
    errs = {"plain": [], "split": []}
    for a in range(3):
        for b in range(a, 3):
            rho = (phi[:, a] * np.conj(phi[:, b])).reshape(n, n, n, order="F").transpose(2, 1, 0)
            ref = _conv(rho, khat, n, M, False)
            plain = _conv(rho, khat, n, M, True)
            r_hi = rho.astype(np.complex64).astype(np.complex128)
            split = (_conv(r_hi, k_hi, n, M, True) + _conv(rho - r_hi, k_hi, n, M, True)
                     + _conv(r_hi, khat - k_hi, n, M, True))
            scale = np.abs(ref).max()
            errs["plain"].append(np.abs(plain - ref).max() / scale)
            errs["split"].append(np.abs(split - ref).max() / scale)
    p, q = np.median(errs["plain"]), np.median(errs["split"])
    assert 1e-8 < p < 1e-6                  # FP32-level error
    assert q > 0.5 * p                      # the split does not buy accuracy
