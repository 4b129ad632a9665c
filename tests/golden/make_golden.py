"""Writes the golden vectors in this folder from the reference's own code.

The reference ships no tests or golden data for this path.  Its only
self-contained compiled pieces on the path -- SD/SineDVR.f90 and the vendored
DFFTPACK complex FFT (S/DFFTPACK/zffti.f, zfftf.f and their passes) -- are
built from the reference sources into oracle/_ref/libref_sinc.so by
oracle/Makefile (`make -C oracle ref`, needs the reference mounted).  This
script runs them on seeded inputs and stores inputs and outputs:

  ref_sinedvr.npz : key "sd_<n>_<spacing>" -> concat(ketot, fdtot, points)
  ref_zfftf.npz   : "in_<M>" (M,3) complex inputs, "out_<M>" zfftf outputs
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
    print("wrote", sorted(sd), sorted(k for k in ff if k.startswith("in_")))


if __name__ == "__main__":
    main()
