"""GPU tests of the orbital operator (actreduced00) and the Krylov
propagation built on it: device operator vs the oracle restatement (FP64,
1e-12), device and host entry points equal, and an Arnoldi / exponential
step with the device matvec vs the same step with the oracle operator."""
import sys

import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-orbop]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def sysm(orc, torch_dev):
    from sincdvr import SincDVR, inputs
    n, dx, N = 15, 0.4, 3
    phi = inputs.gaussian_orbitals(n, dx, N, 3, 321)
    rng = np.random.default_rng(4)
    a = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    denmat = a @ a.conj().T + N * np.eye(N)                  # Hermitian positive, as a density matrix
    invden = np.linalg.inv(denmat)
    s = SincDVR(n, dx, N)
    kern = orc.coulomb_kernel(n, dx)
    _, tr = orc.twoe_matel(n, kern, phi, phi)
    g = rng.standard_normal((N * N, N * N)) * 0.1 + 0j
    rp = orc.reducedpot(tr, g)
    pot = s.nuclear_potential([[0, 0, 0]], [2.0])
    ke, _, _ = orc.sinedvr(n, dx)
    return dict(n=n, N=N, phi=phi, denmat=denmat, invden=invden, s=s, rp=rp, pot=pot, ke=ke)


@pytest.mark.parametrize("dent,proj", [(False, False), (True, False), (True, True)])
def test_actreduced00_vs_oracle(orc, torch_dev, sysm, dent, proj):
    torch = torch_dev
    m = sysm
    say("actreduced00 dent=%s proj=%s" % (dent, proj))
    kw = dict(pot=m["pot"], invdenmat=m["invden"] if dent else None, timefac=-1j if dent else None,
              projspfs=m["phi"] if proj else None)
    got = m["s"].actreduced00(m["phi"], m["denmat"], m["rp"], **kw)
    ref = orc.actreduced00(m["n"], m["ke"], m["phi"], m["denmat"], m["rp"], pot=m["pot"],
                           invdenmat=kw["invdenmat"], timefac=-1j if dent else 1.0, projspfs=kw["projspfs"])
    assert relerr(got, ref) < 1e-12
    T = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a).T)).cuda()      # noqa: E731
    dev = m["s"].actreduced00(T(m["phi"]), T(m["denmat"]), T(m["rp"]), pot=torch.from_numpy(m["pot"]).cuda(),
                              invdenmat=T(m["invden"]) if dent else None, timefac=kw["timefac"],
                              projspfs=T(m["phi"]) if proj else None)
    torch.cuda.synchronize()
    assert np.array_equal(dev.cpu().numpy().T, got)


@pytest.mark.parametrize("ecs", [False, True])
def test_actreduced00_pulse_vs_oracle(orc, torch_dev, sysm, ecs):
    """Length-gauge pulse term of the orbital operator (field on all three
    axes, and on z only) against the oracle, with and without ECS dipoles."""
    torch = torch_dev
    from sincdvr import SincDVR
    m = sysm
    n, dx, N = m["n"], 0.4, m["N"]
    sc = dict(theta=0.35, stretch=1.1, dist=1.5, smooth=0.8) if ecs else None
    s = SincDVR(n, dx, N, scaling=sc) if ecs else m["s"]
    ij = sd = None
    if ecs:
        ij, sd = orc.ecs_arrays(n, dx, sc["theta"], sc["stretch"], sc["dist"], sc["smooth"])
    dip = orc.dipoles_1d(n, dx, (sc["theta"], sc["stretch"], sc["dist"], sc["smooth"]) if ecs else None)
    for td in [(0.02 - 0.01j, -0.03, 0.05 + 0.02j), (0.0, 0.0, 0.04)]:
        say("actreduced00 pulse ecs=%s field=%s" % (ecs, td))
        got = s.actreduced00(m["phi"], m["denmat"], m["rp"], pot=m["pot"], invdenmat=m["invden"], timefac=-1j,
                             projspfs=m["phi"], tdpot=td)
        ref = orc.actreduced00(n, m["ke"], m["phi"], m["denmat"], m["rp"], pot=m["pot"], invdenmat=m["invden"],
                               timefac=-1j, projspfs=m["phi"], invjac=ij, scalediag=sd, tdpot=td, dip=dip)
        assert relerr(got, ref) < 1e-12
        T = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a).T)).cuda()      # noqa: E731
        dev = s.actreduced00(T(m["phi"]), T(m["denmat"]), T(m["rp"]), pot=torch.from_numpy(m["pot"]).cuda(),
                             invdenmat=T(m["invden"]), timefac=-1j, projspfs=T(m["phi"]), tdpot=td)
        torch.cuda.synchronize()
        assert np.array_equal(dev.cpu().numpy().T, got)


@pytest.mark.parametrize("kind", ["plain", "ecs", "mask"])
def test_actreduced00_velocity_gauge_vs_oracle(orc, torch_dev, sysm, kind):
    """Velocity-gauge pulse (velflag 1: velmultiply, the first-derivative
    sinc matrix per axis times -i A_d) in the orbital operator against the
    oracle, plain, under ECS (invsqrtjacobian on both sides) and under the
    edge mask; the vector potential on all three axes and on z only.  The
    mixed form (FP32 matrix cores) stays within 2e-6 of the FP64 one."""
    torch = torch_dev
    from sincdvr import SincDVR
    m = sysm
    n, dx, N = m["n"], 0.4, m["N"]
    sc = dict(theta=0.35, stretch=1.1, dist=1.5, smooth=0.8) if kind == "ecs" else None
    mk = 3 if kind == "mask" else None
    s = SincDVR(n, dx, N, scaling=sc, masknumpoints=mk, velflag=1)
    ij = sd = mask = None
    if kind == "ecs":
        ij, sd = orc.ecs_arrays(n, dx, sc["theta"], sc["stretch"], sc["dist"], sc["smooth"])
    if kind == "mask":
        mask = orc.mask_function(n, mk)
    _, fd, _ = orc.sinedvr(n, dx)
    for A in [(0.02 - 0.01j, -0.03, 0.05 + 0.02j), (0.0, 0.0, 0.04)]:
        say("actreduced00 velocity gauge %s A=%s" % (kind, A))
        got = s.actreduced00(m["phi"], m["denmat"], m["rp"], pot=m["pot"], invdenmat=m["invden"], timefac=-1j,
                             projspfs=m["phi"], tdpot=A)
        ref = orc.actreduced00(n, m["ke"], m["phi"], m["denmat"], m["rp"], pot=m["pot"], invdenmat=m["invden"],
                               timefac=-1j, projspfs=m["phi"], invjac=ij, scalediag=sd, mask=mask, tdpot=A,
                               velflag=1, fd=fd)
        assert relerr(got, ref) < 1e-12
        T = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a).T)).cuda()      # noqa: E731
        dev = s.actreduced00(T(m["phi"]), T(m["denmat"]), T(m["rp"]), pot=torch.from_numpy(m["pot"]).cuda(),
                             invdenmat=T(m["invden"]), timefac=-1j, projspfs=T(m["phi"]), tdpot=A)
        torch.cuda.synchronize()
        assert np.array_equal(dev.cpu().numpy().T, got)
        s.set_precision(1)
        mixed = s.actreduced00(m["phi"], m["denmat"], m["rp"], pot=m["pot"], invdenmat=m["invden"], timefac=-1j,
                               projspfs=m["phi"], tdpot=A)
        s.set_precision(0)
        assert relerr(mixed, ref) < 2e-6


def test_krylov_step_matches_oracle(orc, torch_dev, sysm):
    """One Arnoldi / exponential step of the orbital equation with the device
    operator (vectors resident, FP64 coefficients on the host) against the same
    construction with the oracle operator in numpy."""
    torch = torch_dev
    from scipy.linalg import expm

    from sincdvr import krylov
    m = sysm
    n3, N = m["n"] ** 3, m["N"]
    T = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a).T)).cuda()      # noqa: E731
    dm, idm, rp, pot = T(m["denmat"]), T(m["invden"]), T(m["rp"]), torch.from_numpy(m["pot"]).cuda()
    proj = T(m["phi"])

    def apply_dev(v):
        return m["s"].actreduced00(v.reshape(N, n3), dm, rp, pot=pot, invdenmat=idm, timefac=-1j,
                                   projspfs=proj).reshape(-1)

    def apply_ref(v):
        x = np.asfortranarray(v.reshape(n3, N, order="F"))
        return orc.actreduced00(m["n"], m["ke"], x, m["denmat"], m["rp"], pot=m["pot"], invdenmat=m["invden"],
                                timefac=-1j, projspfs=m["phi"]).reshape(-1, order="F")

    v0 = np.asfortranarray(m["phi"] * (1 + 0.1j))
    dt, order = 0.02, 8
    say("krylov order", order)
    got, H = krylov.expmv(apply_dev, T(v0).reshape(-1), dt, order)
    # the same Arnoldi in numpy with the oracle operator
    v = v0.reshape(-1, order="F")
    beta = np.linalg.norm(v)
    V = [v / beta]
    Hr = np.zeros((order + 1, order), dtype=complex)
    for k in range(order):
        w = apply_ref(V[k])
        for _ in range(2):
            c = np.array([np.vdot(V[j], w) for j in range(k + 1)])
            Hr[: k + 1, k] += c
            for j in range(k + 1):
                w = w - c[j] * V[j]
        Hr[k + 1, k] = np.linalg.norm(w)
        V.append(w / Hr[k + 1, k])
    E = expm(dt * Hr[:order, :order])
    ref = beta * sum(E[j, 0] * V[j] for j in range(order))
    assert relerr(H, Hr) < 1e-10
    g = got.reshape(N, n3).cpu().numpy().T.reshape(-1, order="F")
    assert relerr(g, ref) < 1e-10


def test_phiv_orbital_step_matches_oracle(orc, torch_dev, sysm):
    """One interval of the reference's orbital step phi1 = phi0 + t phi(tJ) f(phi0)
    (S/expo_driver.f90:628-650: DGPHIV with v = 0, u = the derivative) with
    the device operator as J, against the same adaptive driver on the oracle
    operator; the step sizes it chooses are reported by both."""
    torch = torch_dev
    from sincdvr import krylov
    m = sysm
    n3, N = m["n"] ** 3, m["N"]
    T = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a).T)).cuda()      # noqa: E731
    dm, idm, rp, pot = T(m["denmat"]), T(m["invden"]), T(m["rp"]), torch.from_numpy(m["pot"]).cuda()
    proj = T(m["phi"])

    def apply_dev(v):
        return m["s"].actreduced00(v.reshape(N, n3), dm, rp, pot=pot, invdenmat=idm, timefac=-1j,
                                   projspfs=proj).reshape(-1)

    def apply_ref(v):
        x = np.asfortranarray(v.numpy().reshape(n3, N, order="F"))
        y = orc.actreduced00(m["n"], m["ke"], x, m["denmat"], m["rp"], pot=m["pot"], invdenmat=m["invden"],
                             timefac=-1j, projspfs=m["phi"])
        return torch.from_numpy(np.ascontiguousarray(y.reshape(-1, order="F")))

    phi0 = np.asfortranarray(m["phi"] * (1 + 0.1j))
    d0 = T(phi0).reshape(-1)
    r0 = torch.from_numpy(np.ascontiguousarray(phi0.reshape(-1, order="F")))
    t, tol, order, guess = 0.05, 1e-9, 8, 0.01
    say("phiv orbital step t=%g tol=%g m=%d" % (t, tol, order))
    wd, sd = krylov.phiv(apply_dev, t, apply_dev(d0), torch.zeros_like(d0), tol=tol, m=order, guess_step=guess)
    wr, sr = krylov.phiv(apply_ref, t, apply_ref(r0), torch.zeros_like(r0), tol=tol, m=order, guess_step=guess)
    say("device steps %d (rejected %d), oracle steps %d (rejected %d)"
        % (sd["nstep"], sd["nreject"], sr["nstep"], sr["nreject"]))
    got = wd.reshape(N, n3).cpu().numpy().T.reshape(-1, order="F")
    assert abs(sd["t_now"] - t) < 1e-14 and abs(sr["t_now"] - t) < 1e-14
    assert relerr(got, wr.numpy()) < 1e-7
