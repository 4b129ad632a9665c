"""GPU tests at the workloads' own sizes, through the C-ABI:

* 81^3 grid, spacing 0.25, 16 orbitals of 6 Gaussians (136 unique / 256
  ordered pairs, circulant edge 162): FP64 twoereduced and twoematel against
  the oracle at 1e-10 (max-abs error over the largest element), the mixed
  form (FP32 butterflies) against FP64 within 1e-7 for twoematel, with the
  per-element relative error printed;
* 41^3 grid, 10 orbitals, c-norm flavour in FP64 against the oracle;
* the 64-set batch (edges 25..61 at 0.3 bohr, 4..14 orbitals): every set
  bitwise equal to its single-set result in FP64 and mixed, eager and as a
  replayed hipGraph, and the sets of edges 51 and 61 against the oracle;
* serving at edges 31, 41, 61, 81 with 6..16 orbitals: each request's result
  the same alone and in a mix, bit for bit, in both precisions.
"""
import sys

import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu

TOL64 = 1e-10


def say(*a):
    print("[gpu-workload]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _elem_err(a, ref):
    """max-abs error over the largest element, and per-element relative
    errors (median, 99th percentile, max) over elements above 1e-12 of it."""
    d = np.abs(a - ref)
    big = np.abs(ref).max()
    m = np.abs(ref) > 1e-12 * big
    rel = d[m] / np.abs(ref[m])
    return float(d.max() / big), float(np.median(rel)), float(np.percentile(rel, 99)), float(rel.max())


@pytest.mark.timeout(900)
def test_grid81_16orb_fp64_and_mixed(orc, torch_dev):
    from sincdvr import SincDVR, inputs
    cfg = inputs.CONFIGS["grid81_16orb"]
    n, dx, N = cfg["n"], cfg["spacing"], cfg["norb"]
    say("81^3 x 16 orbitals: orbitals")
    phi = inputs.gaussian_orbitals(n, dx, N, cfg["ngauss"], cfg["seed"])
    s = SincDVR(n, dx, N)
    assert s.circulant_edge == 162
    say("81^3 x 16 orbitals: FP64 call_twoe_matel00")
    tm, tr = s.call_twoe_matel00(1, N, phi, phi)
    say("81^3 x 16 orbitals: oracle (M = 162, all 256 ordered pairs)")
    kern = orc.coulomb_kernel(n, dx)
    otm, otr = orc.twoe_matel(n, kern, phi, phi, M=162)
    e_tr, e_tm = relerr(tr, otr), relerr(tm, otm)
    say("FP64 vs oracle: twoereduced %.3e, twoematel %.3e" % (e_tr, e_tm))
    assert e_tr < TOL64 and e_tm < TOL64
    del otr
    say("81^3 x 16 orbitals: mixed call_twoe_matel00")
    s.set_precision(1)
    tm32, tr32 = s.call_twoe_matel00(1, N, phi, phi)
    mx, med, p99, emax = _elem_err(tm32, tm)
    say("mixed vs FP64: twoematel max-abs/max %.3e; per element median %.3e p99 %.3e max %.3e; "
        "twoereduced %.3e; mixed vs oracle twoematel %.3e" % (mx, med, p99, emax, relerr(tr32, tr),
                                                               relerr(tm32, otm)))
    assert mx < 1e-7
    assert relerr(tr32, tr) < 1e-6


def test_grid41_10orb_cnorm_fp64(orc, torch_dev):
    from sincdvr import SincDVR, inputs
    cfg = inputs.CONFIGS["grid41_10orb"]
    n, dx, N = cfg["n"], cfg["spacing"], cfg["norb"]
    phi = inputs.gaussian_orbitals(n, dx, N, cfg["ngauss"], cfg["seed"], cnorm=True)
    say("41^3 x 10 orbitals, cnorm, FP64")
    s = SincDVR(n, dx, N, cnorm=True)
    tm, tr = s.call_twoe_matel00(1, N, phi, phi)
    otm, otr = orc.twoe_matel(n, orc.coulomb_kernel(n, dx), phi, phi, M=82, cnorm=True)
    assert relerr(tr, otr) < TOL64
    assert relerr(tm, otm) < TOL64


# ------------------------------------------------------------ 64-set batch
@pytest.fixture(scope="module")
def batch64(torch_dev):
    from sincdvr import SincDVR, inputs
    descs = inputs.batch_sets()
    sets = []
    for d in descs:
        phi = inputs.gaussian_orbitals_device(d["n"], d["spacing"], d["norb"], d["ngauss"], d["seed"])
        sets.append((SincDVR(d["n"], d["spacing"], d["norb"]), phi))
    torch_dev.cuda.synchronize()
    return descs, sets


@pytest.mark.timeout(600)
@pytest.mark.parametrize("precision", [0, 1])
def test_batch64_equals_single_set(torch_dev, batch64, precision):
    torch = torch_dev
    from sincdvr import Batch
    descs, sets = batch64
    assert len(descs) == 64 and {51, 61} <= {d["n"] for d in descs}
    say("64-set batch precision=%d: eager" % precision)
    b = Batch(sets, precision=precision)
    assert b.stats()["pairs"] == sum(d["norb"] * (d["norb"] + 1) // 2 for d in descs)
    eager = [o.clone() for o in b.run()]
    torch.cuda.synchronize()
    say("64-set batch precision=%d: graph replay" % precision)
    side = torch.cuda.Stream()
    b.capture(side)
    for _ in range(2):
        with torch.cuda.stream(side):
            for o in b.outputs:
                o.fill_(float("nan"))
        b.run(side)
        side.synchronize()
        for i, (o, e) in enumerate(zip(b.outputs, eager)):
            assert torch.equal(o, e), i
    say("64-set batch precision=%d: single-set path per set" % precision)
    for i, ((s, phi), e) in enumerate(zip(sets, eager)):
        s.set_precision(precision)
        r = s.twoereduced(1, s.numspf, phi)
        torch.cuda.synchronize()
        assert torch.equal(e, r), (i, descs[i]["n"], descs[i]["norb"], float((e - r).abs().max()))
        del r
    b.close()


@pytest.mark.timeout(600)
def test_batch64_large_edges_vs_oracle(orc, torch_dev, batch64):
    """The first set of edge 51 (circulant half-edge 54 > n) and of edge 61
    (63 > n): FP64 batch columns twoereduced(:, :, 1:2) against the oracle
    (circulant edges 108 and 126; the result does not depend on M)."""
    torch = torch_dev
    from sincdvr import Batch
    descs, sets = batch64
    pick = [next(i for i, d in enumerate(descs) if d["n"] == e) for e in (51, 61)]
    b = Batch([sets[i] for i in pick], precision=0)
    outs = b.run()
    torch.cuda.synchronize()
    for i, o in zip(pick, outs):
        d = descs[i]
        n = d["n"]
        say("batch set %d (n=%d, %d orbitals) vs oracle" % (i, n, d["norb"]))
        phi = np.asfortranarray(sets[i][1].cpu().numpy().T)
        _, otr = orc.twoe_matel(n, orc.coulomb_kernel(n, d["spacing"]), phi, phi, low=1, high=2,
                                M={51: 108, 61: 126}[n])
        got = o.cpu().numpy().transpose(2, 1, 0)[:, :, 0:2]
        assert relerr(got, otr) < TOL64, (n, relerr(got, otr))
    b.close()


# ----------------------------------------------------------------- serving
@pytest.mark.timeout(600)
@pytest.mark.parametrize("precision", [1, 0])
def test_serving_large_grids_alone_equals_mix(torch_dev, precision):
    torch = torch_dev
    from sincdvr import SincDVR, inputs
    from sincdvr.serving import COULOMB, KE, Server, perturb
    spec = [(31, 6), (41, 9), (61, 12), (81, 16)]
    world = []
    for i, (n, N) in enumerate(spec):
        cur = inputs.gaussian_orbitals_device(n, 0.3, N, 3, 4000 + i)
        gen = torch.Generator(device="cuda")
        gen.manual_seed(4000 + i)
        world.append(dict(s=SincDVR(n, 0.3, N), steps=[perturb(cur, gen), perturb(cur, gen)]))
    torch.cuda.synchronize()
    tick = 256 << 20          # a 16-orbital 81^3 request spans many ticks
    srv = Server(24 << 30, tick, precision)
    alone = {}
    for j, w in enumerate(world):
        for kind in (COULOMB, KE):
            say("serving precision=%d alone: n=%d N=%d kind=%d" % (precision, w["s"].n, w["s"].numspf, kind))
            rid = srv.submit(w["s"], kind, w["steps"][1])
            srv.poll(wait=True)
            assert srv.status(rid)[0] == 2
            alone[(j, kind)] = srv.result(rid)
            srv.release(rid)
    say("serving precision=%d: mix" % precision)
    ids = []
    for step in (0, 1):
        for j, w in enumerate(world):
            for kind in (COULOMB, KE):
                ids.append((j, kind, step, srv.submit(w["s"], kind, w["steps"][step])))
        srv.poll()
    srv.poll(wait=True)
    assert srv.stats()["queued"] == 0
    for j, kind, step, rid in ids:
        r = srv.result(rid)
        if step == 1:
            assert torch.equal(r, alone[(j, kind)]), (spec[j], kind)
        srv.release(rid)
    torch.cuda.synchronize()
    # and the served result equals the direct entry points
    w = world[3]
    w["s"].set_precision(precision)
    assert torch.equal(alone[(3, COULOMB)], w["s"].twoereduced(1, 16, w["steps"][1]))
    assert torch.equal(alone[(3, KE)], w["s"].mult_ke(w["steps"][1]))
    srv.close()
