"""Writes the golden vectors in this folder from the reference's own code.

The reference ships no tests or golden data for this path.  The pieces of
the path that compile from the reference's own sources without external
libraries -- SD/SineDVR.f90, the vendored DFFTPACK complex FFT
(S/DFFTPACK/zffti.f, zfftf.f and their passes), and the convolution
circ3d_sub / myzfft3d / myzfft1d (SD/toeplitz_subs.f90, CP/ftcore.f90) --
are built from the reference sources into oracle/_ref/libref_sinc.so and
libref_conv.so by oracle/Makefile (`make -C oracle ref`, needs the reference
mounted).  This script runs them on seeded inputs and stores inputs and
outputs:

  ref_sinedvr.npz : key "sd_<n>_<spacing>" -> concat(ketot, fdtot, points)
  ref_zfftf.npz   : "in_<M>" (M,3) complex inputs, "out_<M>" zfftf outputs
  ref_op_tinv.npz : "kern_<n>" a seeded (2n-1)^3 real kernel, "den_<n>"
                    (n^3, 2) seeded complex densities, "out_<n>" what the
                    reference's own circ3d_sub / myzfft3d / DFFTPACK
                    (oracle/_ref/libref_conv.so) return for them (M = 2n)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle  # noqa: E402


def main():
    if oracle.ref_lib() is None:
        oracle.build(ref=True)
    assert oracle.ref_lib() is not None, "oracle/_ref not built"
    sd = {}
    for n, dx in [(5, 0.4), (21, 0.4), (41, 0.3), (81, 0.25)]:
        ke, fd, pts = oracle.ref_sinedvr(n, dx)
        sd["sd_%d_%s" % (n, dx)] = np.concatenate([ke, fd, pts])
    np.savez_compressed(os.path.join(HERE, "ref_sinedvr.npz"), **sd)
    ff = {}
    rng = np.random.default_rng(20261203)
    for M in (2, 7, 30, 42, 45, 64, 82, 90, 162):
        x = rng.standard_normal((M, 3)) + 1j * rng.standard_normal((M, 3))
        ff["in_%d" % M] = x
        ff["out_%d" % M] = oracle.ref_zfftf(x)
    np.savez_compressed(os.path.join(HERE, "ref_zfftf.npz"), **ff)
    ot = {}
    rng = np.random.default_rng(20261204)
    for n in (5, 7, 10):
        L = 2 * n - 1
        kern = rng.standard_normal((L, L, L))
        kern = np.asfortranarray(kern + kern[::-1, ::-1, ::-1])          # even, as a Coulomb kernel
        den = np.asfortranarray(rng.standard_normal((n ** 3, 2)) + 1j * rng.standard_normal((n ** 3, 2)))
        ot["kern_%d" % n] = kern
        ot["den_%d" % n] = den
        ot["out_%d" % n] = oracle.ref_op_tinv(n, kern, den)
    np.savez_compressed(os.path.join(HERE, "ref_op_tinv.npz"), **ot)
    print("wrote", sorted(sd), sorted(k for k in ff if k.startswith("in_")), sorted(ot))


if __name__ == "__main__":
    main()
