"""ctypes wrapper over the CPU oracle (liborc.so) and the reference pieces
compiled into oracle/_ref (libref_sinc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product library.  Arrays are numpy
complex128 in Fortran order (first index fastest), exactly the layout of the
reference's DATATYPE arrays.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liborc.so")
REF_SO = os.path.join(HERE, "_ref", "libref_sinc.so")
CONV_SO = os.path.join(HERE, "_ref", "libref_conv.so")

_lib = None
_ref = None
_conv = None

_dp = ctypes.POINTER(ctypes.c_double)


def build(ref=True):
    """Compile the oracle (and oracle/_ref when the reference is mounted)."""
    subprocess.check_call(["make", "-s", "-C", HERE, "oracle"])
    if ref and os.path.isdir("/root/reference/MCTDH.SRC"):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build(ref=False)
        _lib = ctypes.CDLL(ORACLE_SO)
    return _lib


def ref_lib():
    """The reference's own SineDVR + DFFTPACK, or None if not built."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        _ref = ctypes.CDLL(REF_SO)
    return _ref


def conv_lib():
    """The reference's own convolution (circ3d_sub, myzfft3d, DFFTPACK) behind
    oracle/ref_bind.f90, or None if not built."""
    global _conv
    if _conv is None and os.path.exists(CONV_SO):
        _conv = ctypes.CDLL(CONV_SO)
        _conv.ref_op_tinv.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_int, ctypes.c_int]
        _conv.ref_set_threads.argtypes = [ctypes.c_int]
    return _conv


def ref_threed_two(n, kernel):
    """threed_two(-n:n-1)^3 (SD/sincproject.f90:110,182-187) from the kernel
    on offsets -(n-1)..n-1: offset o at index o + n, zero at offset -n."""
    t = np.zeros((2 * n, 2 * n, 2 * n), dtype=np.complex128, order="F")
    t[1:, 1:, 1:] = kernel
    return t


def ref_op_tinv(n, kernel, den, circsize=None, threads=None):
    """op_tinv through the reference's compiled circ3d_sub: den (n^3, allsize)
    -> twoereduced (n^3, allsize), M = 2n.  None if oracle/_ref is not built."""
    cl = conv_lib()
    if cl is None:
        return None
    den = np.asfortranarray(den, dtype=np.complex128)
    if den.ndim == 1:
        den = den.reshape(-1, 1, order="F")
    allsize = den.shape[1]
    circsize = allsize if circsize is None else circsize
    if circsize <= 0 or allsize % circsize:
        raise ValueError("SIZE ERROR OP_TINV")
    t = ref_threed_two(n, kernel)
    red = np.zeros_like(den, order="F")
    def call():
        # the OpenMP thread count is a per-thread setting: set it on the
        # thread that runs the reference's parallel regions
        if threads is not None:
            cl.ref_set_threads(int(threads))
        cl.ref_op_tinv(n, t.ctypes.data, den.ctypes.data, red.ctypes.data, allsize, circsize)

    _run_big_stack(call, 3 * (2 * n) ** 3 * circsize * 16)
    return red


def _run_big_stack(fn, automatic_bytes):
    """circ3d_sub's AUTOMATIC arrays ffmat, ffvec, ffprod (SD/toeplitz_subs.f90:567-568)
    live on the calling thread's stack in the amdflang build (the reference's
    notes ask for `ulimit -s unlimited`): run the call on a thread whose stack
    holds them."""
    import threading
    err = []

    def body():
        try:
            fn()
        except BaseException as e:  # pragma: no cover - re-raised below
            err.append(e)

    old = threading.stack_size()
    threading.stack_size(max(1 << 24, int(automatic_bytes * 1.25) + (64 << 20)))
    try:
        th = threading.Thread(target=body)
        th.start()
    finally:
        threading.stack_size(old)
    th.join()
    if err:
        raise err[0]


def _p(a):
    assert a.flags["F_CONTIGUOUS"] or a.ndim <= 1
    return a.ctypes.data_as(_dp)


def set_threads(n):
    lib().orc_set_threads(ctypes.c_int(n))


def num_threads():
    return lib().orc_num_threads()


def sinedvr(n, spacing):
    ke = np.zeros(2 * n - 1)
    fd = np.zeros(2 * n - 1)
    pts = np.zeros(n)
    lib().orc_sinedvr(ctypes.c_int(n), ctypes.c_double(spacing), _p(ke), _p(fd), _p(pts))
    return ke, fd, pts


def ref_sinedvr(n, spacing):
    r = ref_lib()
    if r is None:
        return None
    ke = np.zeros(2 * n - 1)
    fd = np.zeros(2 * n - 1)
    pts = np.zeros(n)
    r.sinedvr_(_p(ke), _p(fd), _p(pts), ctypes.byref(ctypes.c_int(n)), ctypes.byref(ctypes.c_double(spacing)))
    return ke, fd, pts


def ecs_point(x, theta, stretch, dist, smooth):
    out = [np.zeros(2) for _ in range(4)]
    lib().orc_ecs_point(ctypes.c_double(x), ctypes.c_double(theta), ctypes.c_double(stretch),
                        ctypes.c_double(dist), ctypes.c_double(smooth), *[_p(o) for o in out])
    return [complex(o[0], o[1]) for o in out]  # F, J, J', J''


def ecs_arrays(n, spacing, theta, stretch, dist, smooth):
    ij = np.zeros((n ** 3, 3), dtype=np.complex128, order="F")
    sd = np.zeros(n ** 3, dtype=np.complex128)
    lib().orc_ecs_arrays(ctypes.c_int(n), ctypes.c_double(spacing), ctypes.c_double(theta),
                         ctypes.c_double(stretch), ctypes.c_double(dist), ctypes.c_double(smooth),
                         _p(ij), _p(sd))
    return ij, sd


def getinverse(N, nbig, nsmall, delta, want_tr=False):
    L = 2 * N + 1
    t = np.zeros((L, L, L), order="F")
    ns = 2 * nsmall + 1
    tr = np.zeros((ns, ns, ns), order="F") if want_tr else None
    q = np.zeros((ns, ns, ns), order="F") if want_tr else None
    rc = lib().orc_getinverse(ctypes.c_int(N), ctypes.c_int(nbig), ctypes.c_int(nsmall), ctypes.c_double(delta),
                              _p(t), _p(tr) if want_tr else None, _p(q) if want_tr else None)
    if rc:
        raise ValueError(f"orc_getinverse rc={rc}")
    return (t, tr, q) if want_tr else t


def coulomb_kernel(n, spacing, nbig=240, nsmall=40):
    L = 2 * n - 1
    k = np.zeros((L, L, L), order="F")
    rc = lib().orc_coulomb_kernel(ctypes.c_int(n), ctypes.c_double(spacing), ctypes.c_int(nbig),
                                  ctypes.c_int(nsmall), _p(k))
    if rc:
        raise ValueError(f"orc_coulomb_kernel rc={rc}")
    return k


def fft1d(x):
    """Forward unnormalised FFT along axis 0 of a (M, howmany) complex array."""
    x = np.asfortranarray(x, dtype=np.complex128)
    if x.ndim == 1:
        x = x.reshape(-1, 1, order="F")
    out = np.zeros_like(x, order="F")
    lib().orc_fft1d(ctypes.c_int(x.shape[0]), ctypes.c_int(x.shape[1]), _p(x), _p(out))
    return out


def ref_zfftf(x):
    """DFFTPACK zffti+zfftf per line, as CP/ftcore.f90:140-157 calls it."""
    r = ref_lib()
    if r is None:
        return None
    x = np.asfortranarray(x, dtype=np.complex128)
    if x.ndim == 1:
        x = x.reshape(-1, 1, order="F")
    M = x.shape[0]
    out = x.copy(order="F")
    wsave = np.zeros(2 * (4 * M + 15))
    for h in range(x.shape[1]):
        line = np.ascontiguousarray(out[:, h])
        r.zffti_(ctypes.byref(ctypes.c_int(M)), _p(wsave))
        r.zfftf_(ctypes.byref(ctypes.c_int(M)), _p(line), _p(wsave))
        out[:, h] = line
    return out


def op_tinv(n, kernel, den, M=None, circsize=None):
    """den: (n^3, allsize) complex -> twoereduced (n^3, allsize)."""
    den = np.asfortranarray(den, dtype=np.complex128)
    if den.ndim == 1:
        den = den.reshape(-1, 1, order="F")
    allsize = den.shape[1]
    M = 2 * n if M is None else M
    circsize = allsize if circsize is None else circsize
    red = np.zeros_like(den, order="F")
    rc = lib().orc_op_tinv(ctypes.c_int(n), ctypes.c_int(M), _p(np.asfortranarray(kernel)), _p(den), _p(red),
                           ctypes.c_int(allsize), ctypes.c_int(circsize))
    if rc:
        raise ValueError(f"orc_op_tinv rc={rc}")
    return red


def toeplitz_direct(n, kernel, den):
    den = np.asfortranarray(den, dtype=np.complex128)
    if den.ndim == 1:
        den = den.reshape(-1, 1, order="F")
    red = np.zeros_like(den, order="F")
    lib().orc_toeplitz_direct(ctypes.c_int(n), _p(np.asfortranarray(kernel)), _p(den), _p(red),
                              ctypes.c_int(den.shape[1]))
    return red


def twoe_matel(n, kernel, phi1, phi2, cnorm=False, low=1, high=None, M=None, notwo=False):
    """Returns (twoematel(N,N,N,low:high), twoereduced(n^3,N,low:high))."""
    phi1 = np.asfortranarray(phi1, dtype=np.complex128)
    phi2 = np.asfortranarray(phi2, dtype=np.complex128)
    N = phi1.shape[1]
    high = N if high is None else high
    nnn = high - low + 1
    M = 2 * n if M is None else M
    tm = np.zeros((N, N, N, nnn), dtype=np.complex128, order="F")
    tr = np.zeros((n ** 3, N, nnn), dtype=np.complex128, order="F")
    rc = lib().orc_twoe_matel(ctypes.c_int(n), ctypes.c_int(M), _p(np.asfortranarray(kernel)), ctypes.c_int(N),
                              ctypes.c_int(low), ctypes.c_int(high), _p(phi1), _p(phi2), ctypes.c_int(int(cnorm)),
                              ctypes.c_int(int(notwo)), _p(tm), _p(tr))
    if rc:
        raise ValueError(f"orc_twoe_matel rc={rc}")
    return tm, tr


def mult_ke(n, ke, inp, invjac=None, scalediag=None):
    inp = np.asfortranarray(inp, dtype=np.complex128)
    if inp.ndim == 1:
        inp = inp.reshape(-1, 1, order="F")
    out = np.zeros_like(inp, order="F")
    lib().orc_mult_ke(ctypes.c_int(n), _p(np.ascontiguousarray(ke)),
                      _p(np.asfortranarray(invjac)) if invjac is not None else None,
                      _p(np.ascontiguousarray(scalediag)) if scalediag is not None else None,
                      _p(inp), _p(out), ctypes.c_int(inp.shape[1]))
    return out


def mask_function(n, masknumpoints):
    m = np.zeros(n)
    lib().orc_mask_function(ctypes.c_int(n), ctypes.c_int(masknumpoints), _p(m))
    return m


def mult_ke_mask(n, ke, mask, inp):
    """mult_ke with maskflag on (scalingflag 0)."""
    inp = np.asfortranarray(inp, dtype=np.complex128)
    if inp.ndim == 1:
        inp = inp.reshape(-1, 1, order="F")
    out = np.zeros_like(inp, order="F")
    lib().orc_mult_ke_mask(ctypes.c_int(n), _p(np.ascontiguousarray(ke)), _p(np.ascontiguousarray(mask)),
                           _p(inp), _p(out), ctypes.c_int(inp.shape[1]))
    return out


def mult_ke_dense(n, ke, inp, invjac=None, scalediag=None):
    inp = np.ascontiguousarray(inp, dtype=np.complex128).ravel()
    out = np.zeros_like(inp)
    lib().orc_mult_ke_dense(ctypes.c_int(n), _p(np.ascontiguousarray(ke)),
                            _p(np.asfortranarray(invjac)) if invjac is not None else None,
                            _p(np.ascontiguousarray(scalediag)) if scalediag is not None else None,
                            _p(inp), _p(out))
    return out


def mult_reducedpot(first, last, inspfs, reducedpot):
    inspfs = np.asfortranarray(inspfs, dtype=np.complex128)
    n3, N = inspfs.shape
    rp = np.asfortranarray(reducedpot, dtype=np.complex128)
    out = np.zeros((n3, last - first + 1), dtype=np.complex128, order="F")
    lib().orc_mult_reducedpot(ctypes.c_int(n3), ctypes.c_int(N), ctypes.c_int(first), ctypes.c_int(last),
                              _p(inspfs), _p(out), _p(rp))
    return out


def reducedpot(twoereduced, twoden):
    tr = np.asfortranarray(twoereduced, dtype=np.complex128)
    n3 = tr.shape[0]
    N = int(round((tr.size // n3) ** 0.5))
    g = np.asfortranarray(twoden, dtype=np.complex128)
    out = np.zeros((n3, N, N), dtype=np.complex128, order="F")
    lib().orc_reducedpot(ctypes.c_int(n3), ctypes.c_int(N), _p(tr), _p(g), _p(out))
    return out


def nuclear_pot(n, spacing, kernel, centershift, charges, realkind4=False):
    """get_3dpoisson's nuclear potential on the grid, (n^3,) reals."""
    cs = np.ascontiguousarray(np.asarray(centershift, dtype=np.int32).reshape(-1, 3))
    Z = np.ascontiguousarray(np.asarray(charges, dtype=np.float64))
    pot = np.zeros(n ** 3)
    rc = lib().orc_nuclear_pot(ctypes.c_int(n), ctypes.c_double(spacing), _p(np.asfortranarray(kernel)),
                               ctypes.c_int(len(Z)), cs.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _p(Z),
                               ctypes.c_int(int(realkind4)), _p(pot))
    if rc:
        raise ValueError("centershift parity error (SD/sincproject.f90:199-204)")
    return pot


def one_matel(n, ke, in1, in2, low=1, high=None, pot=None, cnorm=False, invjac=None, scalediag=None):
    """S/matel.f90 sparseops_matel / pot_matel: (numspf, low:high) matrix
    in1^H . op(in2(:, low:high)) with op = mult_ke (pot None) or pot * ;
    ^T instead of ^H for cnorm.  The kinetic apply is the C restatement;
    the GEMM is numpy."""
    in1 = np.asfortranarray(in1, dtype=np.complex128)
    in2 = np.asfortranarray(in2, dtype=np.complex128)
    high = in1.shape[1] if high is None else high
    cols = in2[:, low - 1:high]
    if pot is None:
        w = mult_ke(n, ke, cols, invjac, scalediag)
    else:
        w = cols * np.asarray(pot, dtype=np.float64)[:, None]
    a = in1 if cnorm else np.conj(in1)
    return np.asfortranarray(a.T @ w)


def dipoles_1d(n, spacing, ecs=None):
    """Per-axis dipole values of the cubic grid: the sine-DVR points
    (get_one_dipole, SINCDVR/init_sinc_new.f90:630-645), replaced by the ECS
    scale function F(x) when ecs = (theta, stretch, dist, smooth) is given
    (dipoles = scalefunction, :527)."""
    _, _, pts = sinedvr(n, spacing)
    if ecs is None:
        return pts.astype(np.complex128)
    return np.array([ecs_point(x, *ecs)[0] for x in pts], dtype=np.complex128)


def lenmultiply(n, w, dip, tdpot):
    """S/proputils.f90:157-182 with SINCDVR/coreproject.f90:820-884 (realflag
    0): ((0 + (w dz) Ez) + (w dy) Ey) + (w dx) Ex, a term skipped when its
    field value is zero.  dip: per-axis values (x fastest in the point index)."""
    n3 = n ** 3
    g = np.arange(n3)
    d = [np.asarray(dip)[g % n], np.asarray(dip)[(g // n) % n], np.asarray(dip)[g // (n * n)]]
    out = np.zeros_like(w)
    for axis in (2, 1, 0):
        t = complex(tdpot[axis])
        if t != 0:
            out = out + (w * d[axis][:, None]) * t
    return out


def _axis_mask(w, mask, n, op):
    """divide_mask / mult_mask of SINCDVR/coreproject.f90:1273-1319: x, then
    y, then z coordinate's mask value, one at a time."""
    H = w.shape[1]
    t = np.asfortranarray(w).reshape((n, n, n, H), order="F").copy()
    m = np.asarray(mask, dtype=np.float64)
    for d in range(3):
        shape = [1, 1, 1, 1]
        shape[d] = n
        t = op(t, m.reshape(shape))
    return t.reshape((n ** 3, H), order="F")


def velmultiply(n, w, avec, fd, invsqrtjac=None, mask=None):
    """SINCDVR/coreproject.f90:2119-2134 -> mult_general option 2
    (:1321-1379) -> mult_all0 case 2 (:2089-2092), in numpy:
        out = sum_d c_d W_d (-F_d) W_d w,   c_d = -i A_d,  F_d(p, q) = fdtot(q - p)
    along axis d, W_d = invsqrtjacobian(:, d) under ECS (SD/init_sinc_new.f90:504),
    divide_mask before and mult_mask after under maskflag, identity otherwise;
    an axis with A_d = 0 is skipped (:1339).  fd: the 2n-1 values fdtot(-(n-1)..n-1)
    of sinedvr; w: (n^3, howmany), x fastest."""
    w = np.asfortranarray(w, dtype=np.complex128)
    if w.ndim == 1:
        w = w.reshape(-1, 1, order="F")
    H = w.shape[1]
    F = np.array([[fd[q - p + n - 1] for q in range(n)] for p in range(n)])      # F(p, q)
    out = np.zeros_like(w, order="F")
    for d in range(3):
        c = -1j * complex(avec[d])
        if c == 0:
            continue
        if invsqrtjac is not None:
            t = w * np.asarray(invsqrtjac)[:, d][:, None]
        elif mask is not None:
            t = _axis_mask(w, mask, n, np.divide)
        else:
            t = w
        T = t.reshape((n, n, n, H), order="F")
        T2 = np.moveaxis(np.tensordot(T, -F, axes=([d], [0])), -1, d)           # out(., q, .) = -sum_p in(., p, .) F(p, q)
        t2 = np.asfortranarray(T2).reshape((n ** 3, H), order="F")
        if invsqrtjac is not None:
            out = out + (t2 * np.asarray(invsqrtjac)[:, d][:, None]) * c
        elif mask is not None:
            out = out + _axis_mask(t2, mask, n, np.multiply) * c
        else:
            out = out + t2 * c
    return out


def actreduced00(n, ke, inspfs, denmat, reducedpot, pot=None, invdenmat=None, timefac=1.0, projspfs=None,
                 cnorm=False, invjac=None, scalediag=None, mask=None, tdpot=None, dip=None, velflag=0, fd=None):
    """S/derivs.f90:629-852 for the sinc-DVR project (one bond point): the
    kinetic apply is the C restatement (with ECS or the edge mask), the rest
    numpy, in the reference's order.  tdpot = (Ex, Ey, Ez) with dip
    (dipoles_1d) adds the length-gauge pulse (:742-746); with velflag it is
    (Ax, Ay, Az) and adds velmultiply (:747-750) with the derivative stencil
    fd (sinedvr), sqrt(invjac) as invsqrtjacobian under ECS, or the mask."""
    phi = np.asfortranarray(inspfs, dtype=np.complex128)
    w = phi @ np.asarray(denmat)                                   # spfinvrsq = spfinvr = spfr
    if mask is not None:
        s = mult_ke_mask(n, ke, mask, w)
    else:
        s = mult_ke(n, ke, w, invjac, scalediag)
    if pot is not None:
        s = s + w * np.asarray(pot)[:, None]
    if tdpot is not None and velflag:
        s = s + velmultiply(n, w, tdpot, fd, None if invjac is None else np.sqrt(np.asarray(invjac)), mask)
    elif tdpot is not None:
        s = s + lenmultiply(n, w, dip, tdpot)
    s = s + mult_reducedpot(1, phi.shape[1], phi, reducedpot)
    if invdenmat is not None:
        s = timefac * (s @ np.asarray(invdenmat))
    if projspfs is not None:
        pj = np.asarray(projspfs)
        d = (pj if cnorm else np.conj(pj)).T @ s                    # mydot(j, i) = dot(pro_j, s_i)
        s = s - pj @ d
    return np.asfortranarray(s)
