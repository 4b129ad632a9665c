"""Serving: the request engine of libsincdvr (continuous batching, admission by
HBM footprint) and a load generator for many concurrent propagations.

Server wraps sdv_server_*: submit Coulomb (pair potentials) or kinetic apply
requests for resident orbital sets, poll, read results.  run_stream() replays
a Poisson stream of requests (request_stream) from simulated propagations
(build_propagations: each a grid edge and orbital count, orbitals perturbed
every step by seeded complex noise of norm 1e-3 and re-orthonormalised) in
real time and returns the latencies, summarised by percentiles().
"""
import ctypes
import math
import time

import numpy as np
import torch

from . import _lib, inputs
from .ops import SincDVR, _dptr

COULOMB, KE = _lib.REQ_COULOMB, _lib.REQ_KE
QUEUED, ADMITTED, DONE = 0, 1, 2


class Server:
    def __init__(self, hbm_budget, tick_bytes, precision=1):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self._lib.sdv_server_create(int(hbm_budget), int(tick_bytes), int(precision), ctypes.byref(h)))
        self._h = h
        self.precision = precision
        self._inputs = {}   # request id -> (context, input tensor) kept alive until release

    def submit(self, s: SincDVR, kind, phi):
        """phi: device tensor (numspf, totpoints), unchanged until the request is done."""
        assert phi.is_contiguous() and phi.shape == (s.numspf, s.totpoints)
        rid = ctypes.c_longlong()
        _lib.check(self._lib.sdv_server_submit(self._h, s._h, int(kind), _dptr(phi), ctypes.byref(rid)))
        self._inputs[rid.value] = (s, kind, phi)
        return rid.value

    def poll(self, wait=False):
        n = ctypes.c_int()
        _lib.check(self._lib.sdv_server_poll(self._h, int(wait), ctypes.byref(n)))
        return n.value

    def status(self, rid):
        st, lat, q = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        _lib.check(self._lib.sdv_server_result(self._h, rid, None, ctypes.byref(st), ctypes.byref(lat),
                                               ctypes.byref(q)))
        return st.value, lat.value, q.value

    def result(self, rid):
        """Copy of a done request's output: Coulomb (numspf, numspf, totpoints)
        = twoereduced(totpoints, numspf, numspf); kinetic (numspf, totpoints)."""
        s, kind, phi = self._inputs[rid]
        shape = (s.numspf, s.numspf, s.totpoints) if kind == COULOMB else (s.numspf, s.totpoints)
        out = torch.empty(shape, dtype=torch.complex128, device=phi.device)
        st = torch.cuda.current_stream()
        _lib.check(self._lib.sdv_server_copy_result(self._h, rid, _dptr(out), ctypes.c_void_p(st.cuda_stream)))
        return out

    def release(self, rid):
        _lib.check(self._lib.sdv_server_release(self._h, rid))
        self._inputs.pop(rid, None)

    def stats(self):
        v = [ctypes.c_longlong() for _ in range(5)]
        _lib.check(self._lib.sdv_server_stats(self._h, *[ctypes.byref(x) for x in v]))
        return dict(zip(("queued", "in_flight", "arena_used", "arena_peak", "ticks"), [x.value for x in v]))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.sdv_server_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------ load generator
SERVING_EDGES = (31, 41, 61, 81)


def propagations(nprop=256, seed=4000, spacing=0.3, edges=SERVING_EDGES, norb_range=(6, 16)):
    rng = inputs.SplitMix64(seed)
    out = []
    for i in range(nprop):
        n = edges[rng.randint(0, len(edges) - 1)]
        out.append(dict(n=n, spacing=spacing, norb=rng.randint(*norb_range), ngauss=3, seed=seed * 1000 + i))
    return out


def orthonormalise(phi):
    """Columns of the (numspf, totpoints) device tensor, hermnorm dot, two
    passes of classical Gram-Schmidt (the reference's S/utils.f90:275-320
    scheme) in FP64 on the device."""
    q = phi.clone()
    for k in range(q.shape[0]):
        v = q[k].clone()
        for _ in range(2):
            if k:
                c = torch.mv(q[:k].conj(), v)
                v = v - torch.mv(q[:k].transpose(0, 1), c)
            v = v / torch.sqrt(torch.vdot(v, v).real)
        q[k] = v
    return q


def perturb(phi, gen, norm=1e-3):
    """Seeded complex noise of the given 2-norm, then re-orthonormalise."""
    noise = torch.complex(torch.randn(phi.shape, dtype=torch.float64, device=phi.device, generator=gen),
                          torch.randn(phi.shape, dtype=torch.float64, device=phi.device, generator=gen))
    noise *= norm / torch.linalg.vector_norm(noise)
    return orthonormalise(phi + noise)


def build_propagations(descs, steps, device="cuda"):
    """Contexts (one per (edge, orbital count), kernel spectra shared per
    edge by the library) and `steps` perturbed orbital sets per propagation,
    all resident on the device."""
    ctxs = {}
    props = []
    for i, d in enumerate(descs):
        key = (d["n"], d["norb"])
        if key not in ctxs:
            ctxs[key] = SincDVR(d["n"], d["spacing"], d["norb"])
        cur = inputs.gaussian_orbitals_device(d["n"], d["spacing"], d["norb"], d["ngauss"], d["seed"], device)
        gen = torch.Generator(device=device)
        gen.manual_seed(d["seed"])
        sets = []
        for _ in range(steps):
            cur = perturb(cur, gen)
            sets.append(cur)
        props.append(dict(desc=d, ctx=ctxs[key], sets=sets))
    torch.cuda.synchronize()
    return ctxs, props


def request_stream(props, nreq, rate, seed=7, ke_every=2):
    """Poisson arrivals at `rate` requests/s: (time s, propagation, step, kind).
    Each propagation cycles through its perturbed steps; every ke_every-th
    request of a propagation is a kinetic apply, the rest Coulomb."""
    rng = inputs.SplitMix64(seed)
    t = 0.0
    count = [0] * len(props)
    out = []
    for _ in range(nreq):
        t += -math.log(max(rng.uniform(), 1e-300)) / rate
        p = rng.randint(0, len(props) - 1)
        k = count[p]
        count[p] += 1
        kind = KE if (k % ke_every) == ke_every - 1 else COULOMB
        out.append((t, p, k % len(props[p]["sets"]), kind))
    return out


def pair_count(props, stream):
    return sum(props[p]["ctx"].numspf * (props[p]["ctx"].numspf + 1) // 2 for _, p, _, k in stream if k == COULOMB)


def run_stream(server, props, stream, release=True):
    """Replay the arrival stream against the server in real time; returns
    per-request latencies (ms) by kind, queueing delays and wall time."""
    pending = {}
    lat = {COULOMB: [], KE: []}
    queue_ms = []
    i = 0
    t0 = time.perf_counter()
    while i < len(stream) or pending:
        now = time.perf_counter() - t0
        while i < len(stream) and stream[i][0] <= now:
            _, p, step, kind = stream[i]
            pr = props[p]
            rid = server.submit(pr["ctx"], kind, pr["sets"][step])
            pending[rid] = kind
            i += 1
        server.poll()
        for rid in list(pending):
            st, l, q = server.status(rid)
            if st == DONE:
                lat[pending.pop(rid)].append(l)
                queue_ms.append(q)
                if release:
                    server.release(rid)
        if i < len(stream) and not pending:
            time.sleep(min(2e-4, max(0.0, stream[i][0] - (time.perf_counter() - t0))))
    wall = time.perf_counter() - t0
    return lat, queue_ms, wall


def percentiles(x):
    if not x:
        return None
    a = np.sort(np.asarray(x))
    return {"p50": float(np.percentile(a, 50)), "p99": float(np.percentile(a, 99)), "max": float(a[-1]),
            "count": int(a.size)}
