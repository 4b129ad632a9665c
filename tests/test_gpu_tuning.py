"""GPU tests of the tuning knobs and the autotuner: every circulant size the
autotuner may pick gives the same linear convolution (checked against the
oracle at the FP64 tolerance), the pair batch changes nothing bit for bit,
and a tuned context and batch still produce correct results."""
import sys

import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-tune]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def w21(orc):
    from sincdvr import inputs
    n, dx = 21, 0.4
    phi = inputs.gaussian_orbitals(n, dx, 4, 3, 11)
    kern = orc.coulomb_kernel(n, dx)
    otm, otr = orc.twoe_matel(n, kern, phi, phi)
    return dict(n=n, dx=dx, phi=phi, otm=otm, otr=otr)


@pytest.mark.parametrize("m", [21, 25, 27, 32, 45])
def test_every_circulant_size_matches_oracle(orc, torch_dev, w21, m):
    from sincdvr import SincDVR
    s = SincDVR(w21["n"], w21["dx"], 4)
    say("half edge", m)
    s.set_circulant(m)
    assert s.circulant_edge == 2 * m
    tm, tr = s.call_twoe_matel00(1, 4, w21["phi"], w21["phi"])
    assert relerr(tr, w21["otr"]) < 1e-10
    assert relerr(tm, w21["otm"]) < 1e-10
    s.set_circulant(0)
    assert s.circulant_edge == 2 * 21


def test_bad_circulant_sizes_rejected(torch_dev, w21):
    from sincdvr import SdvError, SincDVR, circulant_sizes
    s = SincDVR(w21["n"], w21["dx"], 4)
    sizes = circulant_sizes()
    assert sizes == sorted(sizes) and 81 in sizes
    for bad in (16, 22, 1000):   # below n, unsupported, beyond the table
        with pytest.raises(SdvError):
            s.set_circulant(bad)
    assert s.circulant_edge == 42


def test_pair_batch_is_bitwise_neutral(torch_dev, w21):
    torch = torch_dev
    from sincdvr import SincDVR
    s = SincDVR(w21["n"], w21["dx"], 4)
    p = torch.from_numpy(np.ascontiguousarray(w21["phi"].T)).cuda()
    ref = s.twoereduced(1, 4, p)
    for pb in (1, 3, 7, 10):
        s.set_pair_batch(pb)
        got = s.twoereduced(1, 4, p)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), pb


def test_tune_context_and_cache(torch_dev, w21, tmp_path):
    torch = torch_dev
    from sincdvr import SincDVR, autotune
    s = SincDVR(w21["n"], w21["dx"], 4)
    p = torch.from_numpy(np.ascontiguousarray(w21["phi"].T)).cuda()
    cache = autotune.TuneCache(str(tmp_path / "tune.json"))
    r = autotune.tune_context(s, p, precision=0, pair_batches=(4, 10), reps=2, cache=cache)
    say("tuned", {k: r[k] for k in ("m", "pair_batch", "ms")})
    assert r["m"] in autotune.candidate_half_edges(21) and len(r["table"]) >= 3
    assert s.circulant_edge == 2 * r["m"]
    cache.save()
    again = autotune.TuneCache(str(tmp_path / "tune.json"))
    s2 = SincDVR(w21["n"], w21["dx"], 4)
    hit = autotune.tune_context(s2, p, precision=0, cache=again)
    assert hit["m"] == r["m"] and s2.circulant_edge == 2 * r["m"]
    tr = s2.twoereduced(1, 4, p).cpu().numpy().transpose(2, 1, 0)
    assert relerr(tr, w21["otr"]) < 1e-10


def test_tune_batch(torch_dev):
    torch = torch_dev
    from sincdvr import SincDVR, autotune, inputs
    spec = [(21, 3), (25, 2), (21, 2), (16, 3)]
    sets, phis = [], []
    for i, (n, N) in enumerate(spec):
        phi = inputs.gaussian_orbitals(n, 0.3, N, 3, 40 + i)
        phis.append(phi)
        sets.append((SincDVR(n, 0.3, N), torch.from_numpy(np.ascontiguousarray(phi.T)).cuda()))
    b, rep = autotune.tune_batch(sets, precision=0, reps=2)
    say("batch tuned", rep["sizes"], rep["best"])
    outs = b.run()
    torch.cuda.synchronize()
    for (s, p), o in zip(sets, outs):
        ref = s.twoereduced(1, s.numspf, p)
        torch.cuda.synchronize()
        assert torch.equal(o, ref)
    b.close()
