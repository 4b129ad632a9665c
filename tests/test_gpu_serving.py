"""GPU tests of the request engine: a request's result is the same alone and
in a mix (bit for bit: every pair is computed by the same kernels whatever
else shares its tick), equals the direct entry points, matches the oracle in
FP64; admission waits for HBM (arena) space; bad requests are refused."""
import sys

import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-serving]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def world(torch_dev):
    torch = torch_dev
    from sincdvr import SincDVR, inputs
    spec = [(21, 4, 0.4), (25, 3, 0.3), (16, 5, 0.3), (21, 2, 0.4)]
    out = []
    for i, (n, N, dx) in enumerate(spec):
        phi = inputs.gaussian_orbitals(n, dx, N, 3, 900 + i)
        out.append(dict(s=SincDVR(n, dx, N), phi=phi, t=torch.from_numpy(np.ascontiguousarray(phi.T)).cuda()))
    return out


def one(server, w, kind):
    rid = server.submit(w["s"], kind, w["t"])
    server.poll(wait=True)
    st, lat, q = server.status(rid)
    assert st == 2 and lat >= 0 and q >= 0
    r = server.result(rid)
    server.release(rid)
    return r


@pytest.mark.parametrize("precision", [1, 0])
def test_alone_equals_mix(torch_dev, world, precision):
    torch = torch_dev
    from sincdvr.serving import COULOMB, KE, Server
    # a tick of 6 pairs of the largest grid: requests split over ticks and share them
    tick = 6 * 2 * 25 * 25 * 25 * (16 if precision == 0 else 8)
    srv = Server(2 << 30, tick, precision)
    say("alone, precision", precision)
    alone = one(srv, world[0], COULOMB)
    alone_ke = one(srv, world[1], KE)
    say("mix")
    ids = []
    for rep in range(3):
        for j, w in enumerate(world):
            ids.append((j, COULOMB, srv.submit(w["s"], COULOMB, w["t"])))
            ids.append((j, KE, srv.submit(w["s"], KE, w["t"])))
        srv.poll()
    srv.poll(wait=True)
    st = srv.stats()
    assert st["queued"] == 0 and st["ticks"] > 3
    for j, kind, rid in ids:
        r = srv.result(rid)
        if j == 0 and kind == COULOMB:
            assert torch.equal(r, alone)
        if j == 1 and kind == KE:
            assert torch.equal(r, alone_ke)
        w = world[j]
        w["s"].set_precision(precision)
        ref = w["s"].twoereduced(1, w["s"].numspf, w["t"]) if kind == COULOMB else w["s"].mult_ke(w["t"])
        torch.cuda.synchronize()
        assert torch.equal(r, ref), (j, kind)
        srv.release(rid)
    srv.close()


def test_fp64_server_vs_oracle(orc, torch_dev, world):
    from sincdvr.serving import COULOMB, Server
    srv = Server(1 << 30, 1 << 28, precision=0)
    w = world[0]
    r = one(srv, w, COULOMB).cpu().numpy().transpose(2, 1, 0)
    kern = orc.coulomb_kernel(21, 0.4)
    _, otr = orc.twoe_matel(21, kern, w["phi"], w["phi"])
    assert relerr(r, otr) < 1e-10
    srv.close()


def test_admission_waits_for_memory(torch_dev, world):
    from sincdvr.serving import COULOMB, DONE, QUEUED, Server
    w = world[0]
    out_bytes = w["s"].totpoints * w["s"].numspf ** 2 * 16
    tick = 1 << 26
    work = 2 * (tick + 256)                                     # the budget holds the two tick buffers
    srv = Server(work + int(out_bytes * 1.5), tick, precision=1)   # room for one output
    ids = [srv.submit(w["s"], COULOMB, w["t"]) for _ in range(3)]
    srv.poll(wait=True)
    states = [srv.status(i)[0] for i in ids]
    say("states", states, srv.stats())
    assert states == [DONE, QUEUED, QUEUED]
    srv.release(ids[0])
    srv.poll(wait=True)
    assert [srv.status(i)[0] for i in ids[1:]] == [DONE, QUEUED]
    srv.release(ids[1])
    srv.poll(wait=True)
    assert srv.status(ids[2])[0] == DONE
    st = srv.stats()
    assert st["arena_peak"] <= out_bytes * 1.5
    srv.release(ids[2])
    srv.close()


def test_bad_requests_refused(torch_dev, world):
    from sincdvr import SdvError
    from sincdvr.serving import COULOMB, Server
    w = world[0]
    with pytest.raises(SdvError):                              # budget below the two tick buffers
        Server(1 << 20, 1 << 26, precision=1)
    srv = Server(2 * ((1 << 26) + 256) + (1 << 20), 1 << 26, precision=1)   # output does not fit the rest
    with pytest.raises(SdvError):
        srv.submit(w["s"], COULOMB, w["t"])
    srv.close()
    srv = Server(1 << 30, 1024, precision=1)                   # one pair's work exceeds the tick
    with pytest.raises(SdvError):
        srv.submit(w["s"], COULOMB, w["t"])
    srv.close()


def test_poisson_stream_small(torch_dev):
    """A short real-time Poisson stream through the load generator."""
    from sincdvr import serving
    descs = serving.propagations(nprop=6, edges=(21, 25), norb_range=(3, 5))
    ctxs, props = serving.build_propagations(descs, steps=2)
    stream = serving.request_stream(props, nreq=30, rate=300.0)
    srv = serving.Server(4 << 30, 1 << 28, precision=1)
    lat, q, wall = serving.run_stream(srv, props, stream)
    say("latency", serving.percentiles(lat[serving.COULOMB]), "wall", wall)
    assert len(lat[serving.COULOMB]) + len(lat[serving.KE]) == 30
    assert srv.stats()["arena_used"] == 0
    srv.close()


def test_release_right_after_copy(torch_dev, world):
    """sdv_server_copy_result enqueues the copy on the caller's stream; a
    release right after it must not hand the output's block to the next
    request before the copy has read it.  The copy stream is held busy so the
    next request's kernels run first."""
    torch = torch_dev
    from sincdvr.serving import COULOMB, Server
    w0, w3 = world[0], world[3]                       # same grid and block size class
    srv = Server(2 * ((1 << 26) + 256) + w0["s"].totpoints * 16 * 16 + 4096, 1 << 26, precision=1)
    rid = srv.submit(w0["s"], COULOMB, w0["t"])
    srv.poll(wait=True)
    say("copy on a busy stream, then release and reuse the block")
    if hasattr(torch.cuda, "_sleep"):                # the copy waits behind this on the current stream
        torch.cuda._sleep(200_000_000)
    else:
        a = torch.randn(8192, 8192, device="cuda")
        for _ in range(8):
            a = a @ a * 1e-3
    got = srv.result(rid)
    srv.release(rid)
    rid2 = srv.submit(w3["s"], COULOMB, w3["t"])     # needs the same arena block
    srv.poll(wait=True)
    assert srv.status(rid2)[0] == 2
    torch.cuda.synchronize()
    w0["s"].set_precision(1)
    ref = w0["s"].twoereduced(1, w0["s"].numspf, w0["t"])
    torch.cuda.synchronize()
    assert torch.equal(got, ref)
    srv.release(rid2)
    srv.close()
