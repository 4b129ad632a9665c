"""GPU tests of the batched form: many orbital sets of uneven grid sizes and
orbital counts in one step (grouped or fused schedule).

The batch runs the same per-pair arithmetic as the single-set path
(sdv_twoereduced_dev), so each set's result must equal that path's result
bit for bit, in FP64 and in the mixed (FP32 butterfly) form; the FP64 result
is also checked against the CPU oracle at the plain drop-in's 1e-10."""
import sys

import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu

# (grid edge, orbitals, cnorm): edges cover circulant half-edges 9 .. 45
SETS = [(9, 3, False), (12, 2, True), (16, 4, False), (21, 4, False), (25, 3, True),
        (27, 2, False), (31, 3, False), (36, 2, True), (41, 3, False)]


def say(*a):
    print("[gpu-batch]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def make_sets(torch, spec, spacing=0.3, seed=77):
    from sincdvr import SincDVR, inputs
    out = []
    for i, (n, N, cnorm) in enumerate(spec):
        phi = inputs.gaussian_orbitals(n, spacing, N, 3, seed + i, cnorm=cnorm)
        s = SincDVR(n, spacing, N, cnorm=cnorm)
        out.append((s, torch.from_numpy(np.ascontiguousarray(phi.T)).cuda(), phi))
    return out


def single_results(torch, sets, precision):
    res = []
    for s, p, _ in sets:
        s.set_precision(precision)
        res.append(s.twoereduced(1, s.numspf, p))
    torch.cuda.synchronize()
    return res


@pytest.mark.parametrize("mode", ["grouped", "fused"])
@pytest.mark.parametrize("precision", [0, 1])
def test_batch_equals_single_set(torch_dev, precision, mode):
    torch = torch_dev
    from sincdvr import Batch
    sets = make_sets(torch, SETS)
    say("batch of %d sets precision=%d schedule %s" % (len(sets), precision, mode))
    b = Batch([(s, p) for s, p, _ in sets], precision=precision, schedule=(mode, 1, 3))
    st = b.stats()
    assert st["pairs"] == sum(N * (N + 1) // 2 for _, N, _ in SETS)
    outs = b.run()
    torch.cuda.synchronize()
    ref = single_results(torch, sets, precision)
    for (n, N, cnorm), o, r in zip(SETS, outs, ref):
        assert torch.equal(o, r), (n, N, cnorm, float((o - r).abs().max()))
    b.close()


def test_batch_fp64_vs_oracle(orc, torch_dev):
    torch = torch_dev
    from sincdvr import Batch
    spec = [(21, 4, False), (16, 3, True), (25, 2, False)]
    sets = make_sets(torch, spec, spacing=0.4, seed=5)
    say("batch vs oracle")
    b = Batch([(s, p) for s, p, _ in sets], precision=0)
    outs = b.run()
    torch.cuda.synchronize()
    for (n, N, cnorm), (s, p, phi), o in zip(spec, sets, outs):
        kern = orc.coulomb_kernel(n, 0.4)
        _, otr = orc.twoe_matel(n, kern, phi, phi, cnorm=cnorm)
        got = o.cpu().numpy().transpose(2, 1, 0)
        assert relerr(got, otr) < 1e-10, (n, N, cnorm)
    b.close()


@pytest.mark.parametrize("schedule", [("fused", 1, 0), ("fused", 3, 0), ("grouped", 1, 0), ("grouped", 1, 2)])
def test_batch_schedules_agree(torch_dev, schedule):
    torch = torch_dev
    from sincdvr import Batch
    sets = make_sets(torch, SETS[:5])
    b = Batch([(s, p) for s, p, _ in sets], precision=1)
    base = [o.clone() for o in b.run()]
    say("schedule %s" % (schedule,))
    b.set_schedule(*schedule)
    for o in b.outputs:
        o.zero_()
    outs = b.run()
    torch.cuda.synchronize()
    for o, r in zip(outs, base):
        assert torch.equal(o, r)
    b.close()


@pytest.mark.parametrize("mode", ["grouped", "fused"])
def test_batch_graph_replay(torch_dev, mode):
    """Capture one step as a hipGraph, then replay it several times with
    eager work (rescaling the orbitals in place) and host synchronisation in
    between; every replay must equal the eager result for the current input.
    The grouped schedule's sizes are parallel branches of the graph."""
    torch = torch_dev
    from sincdvr import Batch
    sets = make_sets(torch, SETS[2:6])
    b = Batch([(s, p) for s, p, _ in sets], precision=1, schedule=(mode, 1, 3))
    side = torch.cuda.Stream()
    say("capture")
    b.capture(side)
    for rep, scale in enumerate([1.0, 0.5, -2.0, 1.25]):
        with torch.cuda.stream(side):
            for _, p, _ in sets:
                if scale != 1.0:
                    p.mul_(scale)
            for o in b.outputs:
                o.fill_(float("nan"))
        b.run(side)               # graph replay on the capture stream
        side.synchronize()
        got = [o.clone() for o in b.outputs]
        eager = single_results(torch, sets, 1)
        say("replay %d" % rep)
        for g, e in zip(got, eager):
            assert torch.equal(g, e), rep
    b.close()


def test_batch_rejects_mismatched_numspf(torch_dev):
    torch = torch_dev
    from sincdvr import Batch, SdvError
    sets = make_sets(torch, SETS[:1])
    s, p, _ = sets[0]
    bad = torch.zeros((s.numspf + 1, s.totpoints), dtype=torch.complex128, device="cuda")
    s.numspf += 1  # the Python view lies about the orbital count; the library must refuse
    try:
        with pytest.raises(SdvError):
            Batch([(s, bad)])
    finally:
        s.numspf -= 1


def test_batch_outlives_context_changes(torch_dev):
    """A batch keeps its own reference on each set's kernel spectrum and its
    own pair lists: changing a context's circulant size, or freeing the
    context, leaves the batch computing (and rebuilding its schedule) with
    the spectrum it was built with, bit for bit."""
    torch = torch_dev
    from sincdvr import Batch
    sets = make_sets(torch, SETS[3:6])
    b = Batch([(s, p) for s, p, _ in sets], precision=0)
    base = [o.clone() for o in b.run()]
    torch.cuda.synchronize()
    s0 = sets[0][0]
    say("context circulant change under a batch")
    s0.set_circulant(64)                    # n = 21: default half-edge 21, now 64
    for o in b.outputs:
        o.zero_()
    outs = b.run()
    torch.cuda.synchronize()
    for o, r in zip(outs, base):
        assert torch.equal(o, r)
    say("context freed under a batch, then schedule rebuilt")
    s0.close()
    sets[1][0].close()
    b.set_schedule("fused", 1, 0)
    for o in b.outputs:
        o.zero_()
    outs = b.run()
    torch.cuda.synchronize()
    for o, r in zip(outs, base):
        assert torch.equal(o, r)
    b.close()
