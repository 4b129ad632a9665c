"""Host-side mirror of the SINCDVR project interface (the reference's
`mult_ke`, `op_tinv`, `call_twoe_matel00`, `mult_reducedpot` and the
`get_reducedpot0` GEMM), over the C-ABI of libsincdvr.so.

Array convention.  The reference passes column-major complex*16 arrays.  Here:
  * numpy arrays are Fortran-ordered complex128 with the reference's shapes,
    e.g. inspfs (totpoints, numspf); they go through the host-pointer entry
    points (synchronous, like the Fortran call);
  * torch tensors are complex128 on the GPU with the reversed (C-order) shape,
    e.g. inspfs (numspf, totpoints), so the storage is identical to the
    Fortran array; they go through the *_dev entry points on the current
    torch stream.
Errors raise SdvError carrying the library's message, where the reference
would print and stop (CFLST).
"""
import ctypes

import numpy as np

from . import _lib

try:
    import torch
except ImportError:  # pragma: no cover - torch is present in this image
    torch = None


def _is_torch(x):
    return torch is not None and isinstance(x, torch.Tensor)


def _dptr(t):
    assert t.is_cuda and t.is_contiguous() and t.dtype == torch.complex128, (t.device, t.dtype)
    return ctypes.c_void_p(t.data_ptr())


def _hptr(a):
    assert a.dtype == np.complex128 and (a.flags["F_CONTIGUOUS"] or a.ndim <= 1), (a.dtype, a.flags)
    return ctypes.c_void_p(a.ctypes.data)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def circulant_sizes():
    """Supported circulant half-edges m (circulant edge M = 2m)."""
    lib = _lib.load()
    buf = (ctypes.c_int * 64)()
    k = lib.sdv_circulant_sizes(buf, 64)
    return [buf[i] for i in range(min(k, 64))]


class SincDVR:
    """One sinc-DVR grid with its device-resident Coulomb kernel spectrum, KE
    stencil and ECS tables (the project state of myprojectmod / myparams)."""

    def __init__(self, numpoints, spacing, numspf, cnorm=False, scaling=None, toothnbig=240,
                 toothnsmall=40, notwoflag=0, debugflag=0, precision=0, masknumpoints=None, velflag=0):
        lib = _lib.load()
        cfg = _lib.SdvConfig()
        lib.sdv_config_default(ctypes.byref(cfg))
        cfg.numpoints = numpoints
        cfg.spacing = spacing
        cfg.numspf = numspf
        cfg.toothnbig = toothnbig
        cfg.toothnsmall = toothnsmall
        cfg.cnormflag = int(bool(cnorm))
        cfg.notwoflag = notwoflag
        cfg.debugflag = debugflag
        cfg.precision = precision
        cfg.velflag = int(velflag)   # pulse gauge of actreduced00: 0 length, else velocity
        if masknumpoints is not None:  # maskflag: kinetic apply through the edge mask (without ECS)
            cfg.maskflag = 1
            cfg.masknumpoints = int(masknumpoints)
        if scaling is not None:
            cfg.scalingflag = 1
            cfg.scalingtheta = scaling["theta"]
            cfg.scalingstretch = scaling.get("stretch", 1.0)
            cfg.scalingdistance = scaling["dist"]
            cfg.smoothness = scaling["smooth"]
        h = ctypes.c_void_p()
        _lib.check(lib.sdv_init(ctypes.byref(cfg), ctypes.byref(h)))
        self._lib = lib
        self._h = h
        self.n = numpoints
        self.totpoints = numpoints ** 3
        self.numspf = numspf
        self.cnorm = bool(cnorm)
        self.precision = precision

    def close(self):
        if getattr(self, "_h", None):
            self._lib.sdv_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_precision(self, p):
        _lib.check(self._lib.sdv_set_precision(self._h, int(p)))
        self.precision = p

    @property
    def circulant_edge(self):
        return self._lib.sdv_circulant_edge(self._h)

    def set_circulant(self, half_edge=0):
        """Rebuild the kernel spectrum for circulant edge 2*half_edge (0: the
        smallest supported half-edge >= numpoints)."""
        _lib.check(self._lib.sdv_set_circulant(self._h, int(half_edge)))

    def set_pair_batch(self, pairs):
        _lib.check(self._lib.sdv_set_pair_batch(self._h, int(pairs)))
        self.pair_batch = pairs

    def kernel_octant(self):
        """threed_two(i,j,k) for i,j,k in [0,n): (n,n,n) Fortran array."""
        out = np.zeros((self.n,) * 3, order="F")
        _lib.check(self._lib.sdv_get_kernel_octant(self._h, ctypes.c_void_p(out.ctypes.data)))
        return out

    def nuclear_potential(self, centershift, charges, realkind4=False):
        """SD/sincproject.f90:192-250: pot(totpoints) for nuclei at half-spacing
        offsets centershift (ncentres, 3) with the given charges."""
        cs = np.ascontiguousarray(np.asarray(centershift, dtype=np.int32).reshape(-1, 3))
        z = np.ascontiguousarray(np.asarray(charges, dtype=np.float64))
        pot = np.zeros(self.totpoints)
        _lib.check(self._lib.sdv_nuclear_potential(self._h, len(z), ctypes.c_void_p(cs.ctypes.data),
                                                   ctypes.c_void_p(z.ctypes.data), int(realkind4),
                                                   ctypes.c_void_p(pot.ctypes.data)))
        return pot

    def ke_stencil(self):
        out = np.zeros(2 * self.n - 1)
        _lib.check(self._lib.sdv_get_ke_stencil(self._h, ctypes.c_void_p(out.ctypes.data)))
        return out

    # ---------------------------------------------------------------- op_tinv
    def op_tinv(self, twoeden03, circsize=None):
        """SD/coreproject.f90:250-258.  numpy (totpoints, allsize) F-order or
        torch (allsize, totpoints) -> twoereduced of the same layout."""
        if _is_torch(twoeden03):
            t = twoeden03.reshape(-1, self.totpoints)
            allsize = t.shape[0]
            out = torch.empty_like(t)
            _lib.check(self._lib.sdv_op_tinv_dev(self._h, _dptr(t), _dptr(out), allsize,
                                                 circsize or allsize, _stream()))
            return out.reshape(twoeden03.shape)
        a = np.asfortranarray(twoeden03, dtype=np.complex128).reshape(self.totpoints, -1, order="F")
        allsize = a.shape[1]
        out = np.zeros_like(a, order="F")
        _lib.check(self._lib.sdv_op_tinv(self._h, _hptr(a), _hptr(out), allsize, circsize or allsize))
        return out

    # ------------------------------------------------------- call_twoe_matel00
    def call_twoe_matel00(self, lowspf, highspf, inspfs10, inspfs20=None, out=None):
        """SD/coreproject.f90:945-955.  Returns (twoematel, twoereduced).
        numpy: twoematel (N,N,N,nnn) and twoereduced (totpoints,N,nnn), F-order.
        torch: twoematel (nnn,N,N,N) and twoereduced (nnn,N,totpoints).
        Passing the same array twice (or inspfs20=None) uses pair symmetry."""
        N = self.numspf
        nnn = highspf - lowspf + 1
        same = inspfs20 is None or inspfs20 is inspfs10
        if _is_torch(inspfs10):
            p1 = inspfs10.contiguous()
            p2 = p1 if same else inspfs20.contiguous()
            if out is None:
                tm = torch.zeros((max(nnn, 0), N, N, N), dtype=torch.complex128, device=p1.device)
                tr = torch.zeros((max(nnn, 0), N, self.totpoints), dtype=torch.complex128, device=p1.device)
            else:
                tm, tr = out
            _lib.check(self._lib.sdv_call_twoe_matel00_dev(self._h, lowspf, highspf, _dptr(p1), _dptr(p2),
                                                           _dptr(tm), _dptr(tr), _stream()))
            return tm, tr
        p1 = np.asfortranarray(inspfs10, dtype=np.complex128)
        p2 = p1 if same else np.asfortranarray(inspfs20, dtype=np.complex128)
        tm = np.zeros((N, N, N, max(nnn, 0)), dtype=np.complex128, order="F")
        tr = np.zeros((self.totpoints, N, max(nnn, 0)), dtype=np.complex128, order="F")
        _lib.check(self._lib.sdv_call_twoe_matel00(self._h, lowspf, highspf, _hptr(p1), _hptr(p2), _hptr(tm),
                                                   _hptr(tr)))
        return tm, tr

    def convolutions(self, lowspf, highspf, same=True):
        """Convolutions twoereduced / call_twoe_matel00 run for this call shape."""
        r = self._lib.sdv_twoe_convolutions(self._h, lowspf, highspf, int(bool(same)))
        return _lib.check(r) if r < 0 else r

    def twoereduced(self, lowspf, highspf, inspfs10, inspfs20=None, out=None):
        """Pair potentials only (the convolution half of call_twoe_matel00)."""
        N = self.numspf
        nnn = highspf - lowspf + 1
        same = inspfs20 is None or inspfs20 is inspfs10
        if _is_torch(inspfs10):
            p1 = inspfs10.contiguous()
            p2 = p1 if same else inspfs20.contiguous()
            tr = out if out is not None else torch.zeros((max(nnn, 0), N, self.totpoints), dtype=torch.complex128,
                                                         device=p1.device)
            _lib.check(self._lib.sdv_twoereduced_dev(self._h, lowspf, highspf, _dptr(p1), _dptr(p2), _dptr(tr),
                                                     _stream()))
            return tr
        p1 = np.asfortranarray(inspfs10, dtype=np.complex128)
        p2 = p1 if same else np.asfortranarray(inspfs20, dtype=np.complex128)
        tr = np.zeros((self.totpoints, N, max(nnn, 0)), dtype=np.complex128, order="F")
        _lib.check(self._lib.sdv_twoereduced(self._h, lowspf, highspf, _hptr(p1), _hptr(p2), _hptr(tr)))
        return tr

    def profile(self, enable=True):
        _lib.check(self._lib.sdv_profile(self._h, int(enable)))

    def profile_read(self):
        """{stage: (total ms, launches)} from HIP events on the launch stream."""
        ms = (ctypes.c_double * len(_lib.STAGES))()
        cnt = (ctypes.c_int * len(_lib.STAGES))()
        _lib.check(self._lib.sdv_profile_read(self._h, ms, cnt))
        return {name: (ms[i], cnt[i]) for i, name in enumerate(_lib.STAGES)}

    # ---------------------------------------------------------------- mult_ke
    def mult_ke(self, inp):
        """SD/coreproject.f90:2104-2116."""
        if _is_torch(inp):
            t = inp.reshape(-1, self.totpoints).contiguous()
            out = torch.empty_like(t)
            _lib.check(self._lib.sdv_mult_ke_dev(self._h, _dptr(t), _dptr(out), t.shape[0], _stream()))
            return out.reshape(inp.shape)
        a = np.asfortranarray(inp, dtype=np.complex128).reshape(self.totpoints, -1, order="F")
        out = np.zeros_like(a, order="F")
        _lib.check(self._lib.sdv_mult_ke(self._h, _hptr(a), _hptr(out), a.shape[1]))
        return out

    # ------------------------------------------- one-electron matrix elements
    def sparseops_matel(self, lowspf, highspf, inspfs1, inspfs2=None):
        """S/matel.f90:241-293: opmatel(numspf, lowspf:highspf) =
        inspfs1^H . mult_ke(inspfs2(:, lowspf:highspf)) (^T for cnorm)."""
        return self._one_matel(lowspf, highspf, inspfs1, inspfs2, None)

    def pot_matel(self, lowspf, highspf, inspfs1, pot, inspfs2=None):
        """S/matel.f90:113-169: potmatel = inspfs1^H . (pot * inspfs2(:, lowspf:highspf))."""
        return self._one_matel(lowspf, highspf, inspfs1, inspfs2, pot)

    def _one_matel(self, lowspf, highspf, inspfs1, inspfs2, pot):
        N = self.numspf
        num = max(highspf - lowspf + 1, 0)
        same = inspfs2 is None
        if _is_torch(inspfs1):
            p1 = inspfs1.contiguous()
            p2 = p1 if same else inspfs2.contiguous()
            out = torch.zeros((num, N), dtype=torch.complex128, device=p1.device)
            if pot is None:
                _lib.check(self._lib.sdv_sparseops_matel_dev(self._h, lowspf, highspf, _dptr(p1), _dptr(p2),
                                                             _dptr(out), _stream()))
            else:
                pt = pot.contiguous()
                assert pt.dtype == torch.float64 and pt.is_cuda
                _lib.check(self._lib.sdv_pot_matel_dev(self._h, lowspf, highspf, _dptr(p1), _dptr(p2),
                                                       ctypes.c_void_p(pt.data_ptr()), _dptr(out), _stream()))
            return out
        p1 = np.asfortranarray(inspfs1, dtype=np.complex128)
        p2 = p1 if same else np.asfortranarray(inspfs2, dtype=np.complex128)
        out = np.zeros((N, num), dtype=np.complex128, order="F")
        if pot is None:
            _lib.check(self._lib.sdv_sparseops_matel(self._h, lowspf, highspf, _hptr(p1), _hptr(p2), _hptr(out)))
        else:
            pt = np.ascontiguousarray(pot, dtype=np.float64)
            _lib.check(self._lib.sdv_pot_matel(self._h, lowspf, highspf, _hptr(p1), _hptr(p2),
                                               ctypes.c_void_p(pt.ctypes.data), _hptr(out)))
        return out

    # ------------------------------------------------ orbital operator apply
    def actreduced00(self, inspfs, denmat, reducedpot, pot=None, invdenmat=None, timefac=None, projspfs=None,
                     tdpot=None):
        """S/derivs.f90:629-852 over all orbitals (see sdv_actreduced00):
        out = (1 - P)[timefac (T w + pot w + pulse w + mult_reducedpot) invdenmat],
        w = inspfs . denmat; P projects on projspfs (flavour dot).  The
        dentime step runs when invdenmat is given, the projection when
        projspfs is given; tdpot adds the pulse: the field (Ex, Ey, Ez) in the
        length gauge (velflag 0), the vector potential (Ax, Ay, Az) in the
        velocity gauge.  numpy (Fortran layout) or torch (reversed shape)."""
        tf = complex(1.0 if timefac is None else timefac)
        td = None
        if tdpot is not None:
            td = (ctypes.c_double * 6)(*[v for c in tdpot for v in (complex(c).real, complex(c).imag)])
        dent = invdenmat is not None
        proj = projspfs is not None
        if _is_torch(inspfs):
            p = inspfs.contiguous()
            dm = denmat.contiguous()
            out = torch.empty_like(p)
            args = [_dptr(p), _dptr(projspfs.contiguous()) if proj else None, _dptr(dm),
                    _dptr(invdenmat.contiguous()) if dent else None, _dptr(reducedpot.contiguous()),
                    ctypes.c_void_p(pot.contiguous().data_ptr()) if pot is not None else None, td]
            _lib.check(self._lib.sdv_actreduced00_dev(self._h, *args, int(dent), int(proj), tf.real, tf.imag,
                                                      _dptr(out), _stream()))
            return out
        p = np.asfortranarray(inspfs, dtype=np.complex128)
        dm = np.asfortranarray(denmat, dtype=np.complex128)
        rp = np.asfortranarray(reducedpot, dtype=np.complex128)
        pj = np.asfortranarray(projspfs, dtype=np.complex128) if proj else None
        idm = np.asfortranarray(invdenmat, dtype=np.complex128) if dent else None
        pt = np.ascontiguousarray(pot, dtype=np.float64) if pot is not None else None
        out = np.zeros_like(p, order="F")
        _lib.check(self._lib.sdv_actreduced00(self._h, _hptr(p), _hptr(pj) if proj else None, _hptr(dm),
                                              _hptr(idm) if dent else None, _hptr(rp),
                                              ctypes.c_void_p(pt.ctypes.data) if pt is not None else None, td,
                                              int(dent), int(proj), tf.real, tf.imag, _hptr(out)))
        return out

    # -------------------------------------------------------- mult_reducedpot
    def mult_reducedpot(self, firstspf, lastspf, inspfs, reducedpot):
        """SD/coreproject.f90:1200-1235."""
        nout = lastspf - firstspf + 1
        if _is_torch(inspfs):
            out = torch.zeros((max(nout, 0), self.totpoints), dtype=torch.complex128, device=inspfs.device)
            _lib.check(self._lib.sdv_mult_reducedpot_dev(self._h, firstspf, lastspf, _dptr(inspfs.contiguous()),
                                                         _dptr(out), _dptr(reducedpot.contiguous()), _stream()))
            return out
        p = np.asfortranarray(inspfs, dtype=np.complex128)
        rp = np.asfortranarray(reducedpot, dtype=np.complex128)
        out = np.zeros((self.totpoints, max(nout, 0)), dtype=np.complex128, order="F")
        _lib.check(self._lib.sdv_mult_reducedpot(self._h, firstspf, lastspf, _hptr(p), _hptr(out), _hptr(rp)))
        return out

    # -------------------------------------------------- get_reducedpot0 GEMM
    def get_reducedpot(self, twoereduced, twoden):
        """S/mean.f90:556-558: reducedpot = twoereduced . twoden."""
        N = self.numspf
        if _is_torch(twoereduced):
            out = torch.empty((N, N, self.totpoints), dtype=torch.complex128, device=twoereduced.device)
            _lib.check(self._lib.sdv_get_reducedpot_dev(self._h, _dptr(twoereduced.contiguous()),
                                                        _dptr(twoden.contiguous()), _dptr(out), _stream()))
            return out
        tr = np.asfortranarray(twoereduced, dtype=np.complex128)
        g = np.asfortranarray(twoden, dtype=np.complex128)
        out = np.zeros((self.totpoints, N, N), dtype=np.complex128, order="F")
        _lib.check(self._lib.sdv_get_reducedpot(self._h, _hptr(tr), _hptr(g), _hptr(out)))
        return out


class Batch:
    """Pair potentials of many orbital sets (each a SincDVR grid with its own
    numspf) in three persistent launches per step.  `sets` is a list of
    (SincDVR, inspfs) with inspfs a device tensor (numspf, totpoints); the
    results are device tensors (numspf, numspf, totpoints) -- for each set the
    twoereduced(totpoints, numspf, numspf) that call_twoe_matel00's
    pair-symmetric path (lowspf..highspf = 1..numspf) produces."""

    def __init__(self, sets, precision=0, schedule=None, outs=None):
        lib = _lib.load()
        self._lib = lib
        self._keep = []
        arr = (_lib.SdvBatchSet * len(sets))()
        self.outputs = []
        for i, (s, phi) in enumerate(sets):
            phi = phi.contiguous()
            assert phi.shape == (s.numspf, s.totpoints), (phi.shape, s.numspf, s.totpoints)
            out = outs[i] if outs is not None else torch.zeros((s.numspf, s.numspf, s.totpoints),
                                                               dtype=torch.complex128, device=phi.device)
            self._keep.append((s, phi))
            self.outputs.append(out)
            arr[i].ctx = s._h
            arr[i].numspf = s.numspf
            arr[i].inspfs = _dptr(phi)
            arr[i].twoereduced = _dptr(out)
        h = ctypes.c_void_p()
        _lib.check(lib.sdv_batch_create(len(sets), arr, int(precision), ctypes.byref(h)))
        self._h = h
        self.precision = precision
        self.schedule = ("grouped", 1, 4)
        if schedule is not None:
            self.set_schedule(*schedule)

    def set_schedule(self, mode="grouped", grid_mult=1, nstreams=4):
        """mode "grouped": one launch per stage per circulant size, sizes on
        `nstreams` forked streams; "fused": one persistent launch per stage,
        grid of grid_mult x compute units.  Drops a captured graph."""
        m = {"fused": _lib.BATCH_FUSED, "grouped": _lib.BATCH_GROUPED}[mode]
        _lib.check(self._lib.sdv_batch_set_schedule(self._h, m, int(grid_mult), int(nstreams)))
        self.schedule = (mode, grid_mult, nstreams)

    def run(self, stream=None):
        st = ctypes.c_void_p(stream.cuda_stream) if stream is not None else _stream()
        _lib.check(self._lib.sdv_batch_run(self._h, st))
        return self.outputs

    def capture(self, stream):
        """Record one step as a hipGraph on `stream` (a non-default
        torch.cuda.Stream); later run(stream) calls replay it."""
        _lib.check(self._lib.sdv_batch_capture(self._h, ctypes.c_void_p(stream.cuda_stream)))

    def stats(self):
        npairs, wb = ctypes.c_longlong(), ctypes.c_longlong()
        _lib.check(self._lib.sdv_batch_stats(self._h, ctypes.byref(npairs), ctypes.byref(wb)))
        return {"pairs": npairs.value, "work_bytes": wb.value}

    def profile(self, enable=True):
        _lib.check(self._lib.sdv_batch_profile(self._h, int(enable)))

    def profile_read(self):
        ms = (ctypes.c_double * len(_lib.STAGES))()
        cnt = (ctypes.c_int * len(_lib.STAGES))()
        _lib.check(self._lib.sdv_batch_profile_read(self._h, ms, cnt))
        return {name: (ms[i], cnt[i]) for i, name in enumerate(_lib.STAGES)}

    def close(self):
        if getattr(self, "_h", None):
            self._lib.sdv_batch_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
