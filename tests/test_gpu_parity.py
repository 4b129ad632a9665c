"""GPU parity: the HIP path through the C-ABI against the CPU oracle on the
same seeded inputs.  FP64 form: matrix elements within 1e-10 relative (the
tolerance the task states for the plain drop-in); index bookkeeping exact.
Mixed form (FP32 butterflies): its error is measured and bounded here."""
import sys

import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu

TOL64 = 1e-10


def say(*a):
    print("[gpu-test]", *a, flush=True)
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
    cfg = inputs.CONFIGS["grid21_4orb"]
    kern = orc.coulomb_kernel(cfg["n"], cfg["spacing"])
    phi_h = inputs.gaussian_orbitals(cfg["n"], cfg["spacing"], cfg["norb"], cfg["ngauss"], cfg["seed"], cnorm=False)
    phi_c = inputs.gaussian_orbitals(cfg["n"], cfg["spacing"], cfg["norb"], cfg["ngauss"], cfg["seed"], cnorm=True)
    return dict(cfg=cfg, kern=kern, phi_h=phi_h, phi_c=phi_c)


def test_kernel_and_stencil_parity(orc, torch_dev, w21):
    from sincdvr import SincDVR
    n, dx = w21["cfg"]["n"], w21["cfg"]["spacing"]
    say("kernel octant n=%d" % n)
    s = SincDVR(n, dx, 4)
    ko = s.kernel_octant()
    ref = w21["kern"][n - 1:, n - 1:, n - 1:]
    assert relerr(ko, ref) < 1e-13
    ke, _, _ = orc.sinedvr(n, dx)
    assert relerr(s.ke_stencil(), ke) < 1e-15
    assert s.circulant_edge >= 2 * n - 1


@pytest.mark.parametrize("cnorm", [False, True])
def test_twoe_matel_grid21_fp64(orc, torch_dev, w21, cnorm):
    from sincdvr import SincDVR
    n, dx = w21["cfg"]["n"], w21["cfg"]["spacing"]
    phi = w21["phi_c"] if cnorm else w21["phi_h"]
    say("call_twoe_matel00 n=%d N=4 cnorm=%s host pointers" % (n, cnorm))
    s = SincDVR(n, dx, 4, cnorm=cnorm)
    tm, tr = s.call_twoe_matel00(1, 4, phi, phi)
    otm, otr = orc.twoe_matel(n, w21["kern"], phi, phi, cnorm=cnorm)     # reference layout M = 2n
    assert relerr(tr, otr) < TOL64
    assert relerr(tm, otm) < TOL64
    # device pointers give the same numbers
    say("call_twoe_matel00 device pointers")
    torch = torch_dev
    p = torch.from_numpy(np.ascontiguousarray(phi.T)).cuda()
    dtm, dtr = s.call_twoe_matel00(1, 4, p, p)
    torch.cuda.synchronize()
    assert np.array_equal(dtm.cpu().numpy().transpose(3, 2, 1, 0), tm)
    assert np.array_equal(dtr.cpu().numpy().transpose(2, 1, 0), tr)


@pytest.mark.parametrize("lowhigh", [(2, 3), (4, 4), (1, 1)])
@pytest.mark.parametrize("same", [True, False])
def test_twoe_matel_ranges_and_general_pairs(orc, torch_dev, w21, lowhigh, same):
    from sincdvr import SincDVR
    n, dx = w21["cfg"]["n"], w21["cfg"]["spacing"]
    phi1 = w21["phi_h"]
    phi2 = phi1 if same else np.asfortranarray(phi1[:, ::-1] * (0.5 + 0.25j))
    low, high = lowhigh
    say("call_twoe_matel00 low=%d high=%d same=%s" % (low, high, same))
    s = SincDVR(n, dx, 4)
    tm, tr = s.call_twoe_matel00(low, high, phi1, phi2)
    otm, otr = orc.twoe_matel(n, w21["kern"], phi1, phi2, low=low, high=high)
    assert tm.shape == otm.shape and tr.shape == otr.shape
    assert relerr(tr, otr) < TOL64
    assert relerr(tm, otm) < TOL64


def test_flags_notwo_and_debug(orc, torch_dev, w21):
    from sincdvr import SincDVR
    n, dx = w21["cfg"]["n"], w21["cfg"]["spacing"]
    say("notwoflag / debugflag=4242")
    tm, tr = SincDVR(n, dx, 4, notwoflag=1).call_twoe_matel00(1, 4, w21["phi_h"], w21["phi_h"])
    assert not tm.any() and not tr.any()
    den = np.asfortranarray(w21["phi_h"] * 1.0)
    assert not SincDVR(n, dx, 4, debugflag=4242).op_tinv(den).any()


def test_op_tinv_size_error(torch_dev, w21):
    from sincdvr import SdvError, SincDVR
    n, dx = w21["cfg"]["n"], w21["cfg"]["spacing"]
    s = SincDVR(n, dx, 4)
    with pytest.raises(SdvError):
        s.op_tinv(np.zeros((n ** 3, 3), dtype=complex, order="F"), circsize=2)
    with pytest.raises(SdvError):
        s.call_twoe_matel00(0, 4, w21["phi_h"], w21["phi_h"])


@pytest.mark.parametrize("n", [5, 9, 12, 16, 21, 25, 27, 32, 36, 45, 48, 54, 63, 64, 72, 81])
def test_op_tinv_all_circulant_sizes(orc, torch_dev, n):
    """Every compiled (P, Q) factorisation against the oracle at M = 2n."""
    from sincdvr import SincDVR
    dx = 0.3
    rng = np.random.default_rng(n)
    den = np.asfortranarray(rng.standard_normal((n ** 3, 2)) + 1j * rng.standard_normal((n ** 3, 2)))
    say("op_tinv n=%d" % n)
    s = SincDVR(n, dx, 2)
    kern = orc.coulomb_kernel(n, dx)
    ref = orc.op_tinv(n, kern, den)
    assert relerr(s.op_tinv(den), ref) < 1e-12
    s.set_precision(1)
    e32 = relerr(s.op_tinv(den), ref)
    say("op_tinv n=%d fp32 rel err %.3e" % (n, e32))
    assert e32 < 2e-5


@pytest.mark.parametrize("n", [70, 77])
def test_op_tinv_partial_grid_large_circulant(orc, torch_dev, n):
    """Grid edges below the circulant half-edge at the two largest sizes
    (70 -> m = 72, 77 -> m = 81): the FP64 y/z kernel keeps the even y
    channel's output of the rows that fit in the LDS beyond the plane and
    parks the other rows in the plane's own HBM slots."""
    from sincdvr import SincDVR
    dx = 0.3
    rng = np.random.default_rng(n)
    den = np.asfortranarray(rng.standard_normal((n ** 3, 2)) + 1j * rng.standard_normal((n ** 3, 2)))
    say("op_tinv n=%d (m > n)" % n)
    s = SincDVR(n, dx, 2)
    assert s.circulant_edge == {70: 144, 77: 162}[n]
    ref = orc.op_tinv(n, orc.coulomb_kernel(n, dx), den)
    assert relerr(s.op_tinv(den), ref) < 1e-12
    s.set_precision(1)
    assert relerr(s.op_tinv(den), ref) < 2e-5


@pytest.mark.parametrize("n", [41, 21, 7, 81])
@pytest.mark.parametrize("ecs", [False, True])
def test_mult_ke(orc, torch_dev, ecs, n):
    """FP64 matrix-core kinetic apply vs the oracle at 1e-12; the FP32
    matrix-core (mixed) form within 2e-6 of it.  Edges 7, 21, 41, 81 cover
    partial q tiles, partial line tiles and the largest grid."""
    from sincdvr import SincDVR, inputs
    dx = 0.3
    phi = inputs.gaussian_orbitals(n, dx, 3, 4, seed=11)
    sc = inputs.ECS_41 if ecs else None
    say("mult_ke n=%d ecs=%s" % (n, ecs))
    s = SincDVR(n, dx, 3, scaling=sc)
    out = s.mult_ke(phi)
    ke, _, _ = orc.sinedvr(n, dx)
    ij = sd = None
    if ecs:
        ij, sd = orc.ecs_arrays(n, dx, sc["theta"], sc["stretch"], sc["dist"], sc["smooth"])
    ref = orc.mult_ke(n, ke, phi, ij, sd)
    assert relerr(out, ref) < 1e-12
    s.set_precision(1)
    out32 = s.mult_ke(phi)
    e32 = relerr(out32, ref)
    say("mult_ke fp32 matrix cores rel err %.3e" % e32)
    assert e32 < 2e-6


def test_mean_field(orc, torch_dev, w21):
    from sincdvr import SincDVR
    n, dx, N = w21["cfg"]["n"], w21["cfg"]["spacing"], 4
    rng = np.random.default_rng(9)
    tr = np.asfortranarray(rng.standard_normal((n ** 3, N, N)) + 1j * rng.standard_normal((n ** 3, N, N)))
    g = np.asfortranarray(rng.standard_normal((N, N, N, N)) + 1j * rng.standard_normal((N, N, N, N)))
    say("get_reducedpot / mult_reducedpot")
    s = SincDVR(n, dx, N)
    rp = s.get_reducedpot(tr, g)
    assert relerr(rp, orc.reducedpot(tr, g)) < 1e-13
    out = s.mult_reducedpot(2, 4, w21["phi_h"], np.asfortranarray(rp[:, :, 1:4]))
    assert relerr(out, orc.mult_reducedpot(2, 4, w21["phi_h"], rp[:, :, 1:4])) < 1e-13


@pytest.mark.parametrize("n,N", [(41, 10), (9, 17), (81, 16)])
def test_mean_field_matrix_cores(orc, torch_dev, n, N):
    """get_reducedpot0's GEMM on the FP64 matrix cores: against the oracle
    (41^3 with 10 orbitals; 9^3 with 17 orbitals, N^2 = 289 columns, two
    column groups with a ragged one) or, at 81^3 with 16 orbitals, on 4096
    sampled grid rows (including the ragged last row tile) against numpy."""
    torch = torch_dev
    from sincdvr import SincDVR
    say("get_reducedpot n=%d N=%d" % (n, N))
    rng = np.random.default_rng(1000 * n + N)
    g = np.asfortranarray(rng.standard_normal((N, N, N, N)) + 1j * rng.standard_normal((N, N, N, N)))
    s = SincDVR(n, 0.3, N)
    if n < 81:
        tr = np.asfortranarray(rng.standard_normal((n ** 3, N, N)) + 1j * rng.standard_normal((n ** 3, N, N)))
        assert relerr(s.get_reducedpot(tr, g), orc.reducedpot(tr, g)) < 1e-13
        return
    gen = torch.Generator(device="cuda")
    gen.manual_seed(n)
    trd = torch.complex(torch.randn((N, N, n ** 3), dtype=torch.float64, device="cuda", generator=gen),
                        torch.randn((N, N, n ** 3), dtype=torch.float64, device="cuda", generator=gen))
    gd = torch.from_numpy(np.ascontiguousarray(g.reshape(N * N, N * N, order="F").T)).cuda()
    rp = s.get_reducedpot(trd, gd)          # (N, N, n^3): reducedpot(n^3, N^2) in Fortran order
    torch.cuda.synchronize()
    rows = np.unique(np.concatenate([rng.integers(0, n ** 3, 4000), np.arange(n ** 3 - 96, n ** 3)]))
    A = trd.reshape(N * N, n ** 3)[:, rows].cpu().numpy().T           # (rows, N^2)
    ref = A @ g.reshape(N * N, N * N, order="F")
    got = rp.reshape(N * N, n ** 3)[:, rows].cpu().numpy().T
    assert relerr(got, ref) < 1e-13


def test_twoe_matel_grid41(orc, torch_dev):
    """41^3 grid, 10 orbitals (55 unique / 100 ordered pairs), hermnorm."""
    from sincdvr import SincDVR, inputs
    cfg = inputs.CONFIGS["grid41_10orb"]
    n, dx, N = cfg["n"], cfg["spacing"], cfg["norb"]
    phi = inputs.gaussian_orbitals(n, dx, N, cfg["ngauss"], cfg["seed"])
    say("call_twoe_matel00 n=41 N=10")
    s = SincDVR(n, dx, N)
    tm, tr = s.call_twoe_matel00(1, N, phi, phi)
    kern = orc.coulomb_kernel(n, dx)
    otm, otr = orc.twoe_matel(n, kern, phi, phi, M=81)
    assert relerr(tr, otr) < TOL64
    assert relerr(tm, otm) < TOL64


@pytest.mark.parametrize("cnorm", [False, True])
def test_twoe_matel_mixed_charge_carrier(orc, torch_dev, cnorm):
    """Mixed form (FP32 butterflies) at 41^3 with 10 orbitals: the two-electron
    matrix elements stay within 1e-7 (max-abs error over the largest element)
    of the oracle.  When bra and ket are the same array, each diagonal pair
    density convolves as rho_aa - g in FP32 with K*g (g a unit-charge Gaussian)
    added in FP64; with distinct arrays of equal values the carrier is off and
    the plain FP32 error remains.  The carrier must lower the error of the
    diagonal pair potentials, the only ones it changes.  (For cnorm the
    largest matrix-element error is the part that is linear in the density,
    from the FP32-rounded twiddles and spectrum, which no carrier removes, so
    the matrix elements are only held to the bound.)"""
    from sincdvr import SincDVR, inputs
    cfg = inputs.CONFIGS["grid41_10orb"]
    n, dx, N = cfg["n"], cfg["spacing"], cfg["norb"]
    phi = inputs.gaussian_orbitals(n, dx, N, cfg["ngauss"], cfg["seed"], cnorm=cnorm)
    kern = orc.coulomb_kernel(n, dx)
    otm, otr = orc.twoe_matel(n, kern, phi, phi, M=81, cnorm=cnorm)
    s = SincDVR(n, dx, N, cnorm=cnorm, precision=1)
    tm, tr = s.call_twoe_matel00(1, N, phi, phi)
    tm2, tr2 = s.call_twoe_matel00(1, N, phi, phi.copy())
    e_carrier, e_plain = relerr(tm, otm), relerr(tm2, otm)
    # twoereduced(:, a, a): the potentials of the diagonal pairs
    dg = np.arange(N)
    den = np.abs(otr).max()
    d_carrier = np.abs(tr[:, dg, dg] - otr[:, dg, dg]).max() / den
    d_plain = np.abs(tr2[:, dg, dg] - otr[:, dg, dg]).max() / den
    say("mixed n=41 N=10 cnorm=%d: twoematel err %.3e with carrier, %.3e without; diagonal pair potentials "
        "%.3e, %.3e; twoereduced %.3e" % (cnorm, e_carrier, e_plain, d_carrier, d_plain, relerr(tr, otr)))
    assert e_carrier < 1e-7
    if not cnorm:
        assert e_carrier < e_plain
    assert d_carrier < 0.8 * d_plain
    assert relerr(tr, otr) < 2e-6


@pytest.mark.parametrize("cnorm", [False, True])
def test_convolution_count_and_packed_diagonals(orc, torch_dev, cnorm):
    """Each unordered pair is convolved once; for hermnorm two diagonal pairs
    (real densities) share one convolution, the odd one out runs alone.  The
    packed diagonal potentials are real and match the oracle."""
    from sincdvr import SincDVR, inputs
    n, dx = 9, 0.5
    for N, low, high in [(5, 1, 5), (7, 3, 7)]:
        s = SincDVR(n, dx, N, cnorm=cnorm)
        nd = high - low + 1                                     # diagonal pairs in range
        offdiag = N * nd - nd - (nd * (nd - 1)) // 2            # unordered off-diagonal pairs
        want = offdiag + ((nd + 1) // 2 if not cnorm else nd)
        assert s.convolutions(low, high, same=True) == want
        assert s.convolutions(low, high, same=False) == N * nd
        phi = inputs.gaussian_orbitals(n, dx, N, 2, seed=91, cnorm=cnorm)
        say("packed diagonals N=%d low=%d high=%d cnorm=%d" % (N, low, high, cnorm))
        tm, tr = s.call_twoe_matel00(low, high, phi, phi)
        otm, otr = orc.twoe_matel(n, orc.coulomb_kernel(n, dx), phi, phi, low=low, high=high, cnorm=cnorm)
        assert relerr(tr, otr) < TOL64 and relerr(tm, otm) < TOL64
        if not cnorm:  # packed in order (low, low+1), (low+2, low+3), ...; an odd last one runs alone
            for b in range(low, low + 2 * (nd // 2)):
                assert np.all(tr[:, b - 1, b - low].imag == 0.0)


def test_twoereduced_entry_and_profiling(torch_dev, w21):
    """sdv_twoereduced_dev delivers the same pair potentials as
    call_twoe_matel00, and the event profiler counts one launch per stage."""
    from sincdvr import SincDVR
    torch = torch_dev
    n, dx = w21["cfg"]["n"], w21["cfg"]["spacing"]
    say("twoereduced + profile")
    s = SincDVR(n, dx, 4)
    p = torch.from_numpy(np.ascontiguousarray(w21["phi_h"].T)).cuda()
    _, tr = s.call_twoe_matel00(1, 4, p, p)
    s.profile(True)
    tr2 = s.twoereduced(1, 4, p)
    prof = s.profile_read()
    s.profile(False)
    assert torch.equal(tr, tr2)
    assert prof["xfwd"][1] == prof["yz"][1] == prof["xinv"][1] >= 1
    assert prof["yz"][0] > 0


@pytest.mark.parametrize("lowhigh", [(1, 16), (3, 9)])
def test_twoe_matel_16_orbitals_small_grid(orc, torch_dev, lowhigh):
    """16 orbitals (136 unique pairs: the 81^3 workload's reduction tiling,
    including the symmetric upper-triangle split-K sizing) on a 9^3 grid."""
    from sincdvr import SincDVR, inputs
    n, dx, N = 9, 0.5, 16
    phi = inputs.gaussian_orbitals(n, dx, N, 2, seed=77)
    low, high = lowhigh
    say("call_twoe_matel00 n=9 N=16 low=%d high=%d" % (low, high))
    s = SincDVR(n, dx, N)
    tm, tr = s.call_twoe_matel00(low, high, phi, phi)
    kern = orc.coulomb_kernel(n, dx)
    otm, otr = orc.twoe_matel(n, kern, phi, phi, low=low, high=high)
    assert relerr(tr, otr) < TOL64
    assert relerr(tm, otm) < TOL64


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("cnorm", [False, True])
def test_single_orbital_and_range_edges(orc, torch_dev, cnorm, precision):
    """One orbital: a single diagonal pair, which for hermnorm has no partner
    to share a convolution with and runs alone.  An empty spf range (high <
    low: an empty loop at SD/coreproject.f90:1042) computes nothing and leaves
    the output as it was; a range outside 1..numspf and a grid edge beyond
    the largest circulant half-edge (81) are errors."""
    from sincdvr import SdvError, SincDVR, inputs
    n, dx = 21, 0.4
    phi = inputs.gaussian_orbitals(n, dx, 1, 3, seed=91, cnorm=cnorm)
    say("single orbital n=21 cnorm=%d precision=%d" % (cnorm, precision))
    s = SincDVR(n, dx, 1, cnorm=cnorm, precision=precision)
    tm, tr = s.call_twoe_matel00(1, 1, phi, phi)
    otm, otr = orc.twoe_matel(n, orc.coulomb_kernel(n, dx), phi, phi, cnorm=cnorm)
    tol = TOL64 if precision == 0 else 2e-6
    assert relerr(tr, otr) < tol
    assert relerr(tm, otm) < tol
    if not cnorm:  # |phi|^2 convolved with a real kernel: real up to rounding
        assert np.abs(tr.imag).max() <= tol * np.abs(tr).max()
    p = torch_dev.from_numpy(np.ascontiguousarray(phi.T)).cuda()
    keep = torch_dev.full((1, 1, n ** 3), 7.0 + 3.0j, dtype=torch_dev.complex128, device="cuda")
    s.twoereduced(2, 1, p, out=keep)
    torch_dev.cuda.synchronize()
    assert bool((keep == 7.0 + 3.0j).all())
    for low, high in ((0, 1), (1, 2)):
        with pytest.raises(SdvError) as e:
            s.twoereduced(low, high, p, out=keep)
        assert e.value.code == -1
    with pytest.raises(SdvError) as e:
        SincDVR(82, dx, 1, precision=precision)
    assert e.value.code == -3


@pytest.mark.parametrize("realkind4", [False, True])
def test_nuclear_potential(orc, torch_dev, realkind4):
    """Two nuclei of charge 1 at +-0.6 bohr on z (centershift +-4) on the
    41^3, 0.3-bohr grid, and one charge-2 nucleus at the origin on 21^3."""
    from sincdvr import SdvError, SincDVR
    for n, dx, cs, zc in [(41, 0.3, [[0, 0, 4], [0, 0, -4]], [1.0, 1.0]), (21, 0.4, [[0, 0, 0]], [2.0])]:
        say("nuclear potential n=%d realkind4=%s" % (n, realkind4))
        s = SincDVR(n, dx, 1)
        kern = orc.coulomb_kernel(n, dx)
        ref = orc.nuclear_pot(n, dx, kern, cs, zc, realkind4=realkind4)
        assert relerr(s.nuclear_potential(cs, zc, realkind4=realkind4), ref) < 1e-13
    with pytest.raises(SdvError):
        s.nuclear_potential([[1, 0, 0]], [1.0])


@pytest.mark.parametrize("cnorm", [False, True])
@pytest.mark.parametrize("ecs", [False, True])
def test_one_electron_matels(orc, torch_dev, cnorm, ecs):
    """sparseops_matel / pot_matel: FP64 vs the oracle (1e-12), sub-range,
    bra != ket, device and host entry points; hermnorm matrices Hermitian and
    cnorm ones symmetric for bra == ket without ECS."""
    torch = torch_dev
    from sincdvr import SincDVR, inputs
    n, dx, N = 21, 0.4, 5
    phi = inputs.gaussian_orbitals(n, dx, N, 3, 77, cnorm=cnorm)
    phi2 = np.asfortranarray(phi[:, ::-1] * (0.3 - 0.7j))
    sc = dict(theta=0.3, stretch=1.0, dist=2.5, smooth=1.5) if ecs else None
    s = SincDVR(n, dx, N, cnorm=cnorm, scaling=sc)
    pot = s.nuclear_potential([[0, 0, 0]], [2.0])
    ke, _, _ = orc.sinedvr(n, dx)
    ij = sd = None
    if ecs:
        ij, sd = orc.ecs_arrays(n, dx, sc["theta"], sc["stretch"], sc["dist"], sc["smooth"])
    say("one-electron matels cnorm=%s ecs=%s" % (cnorm, ecs))
    for (lo, hi, p2) in [(1, N, phi), (2, 4, phi2)]:
        op = s.sparseops_matel(lo, hi, phi, p2)
        ref = orc.one_matel(n, ke, phi, p2, lo, hi, cnorm=cnorm, invjac=ij, scalediag=sd)
        assert relerr(op, ref) < 1e-12
        pm = s.pot_matel(lo, hi, phi, pot, p2)
        pref = orc.one_matel(n, ke, phi, p2, lo, hi, pot=pot, cnorm=cnorm)
        assert relerr(pm, pref) < 1e-12
    t1 = torch.from_numpy(np.ascontiguousarray(phi.T)).cuda()
    opd = s.sparseops_matel(1, N, t1).cpu().numpy().T
    pmd = s.pot_matel(1, N, t1, torch.from_numpy(pot).cuda()).cpu().numpy().T
    op = s.sparseops_matel(1, N, phi)
    assert np.array_equal(opd, op)
    assert np.array_equal(pmd, s.pot_matel(1, N, phi, pot))
    if not ecs:
        sym = op.T if cnorm else op.conj().T
        assert relerr(op, sym) < 1e-12
        pm = s.pot_matel(1, N, phi, pot)
        assert relerr(pm, pm.T if cnorm else pm.conj().T) < 1e-12


@pytest.mark.parametrize("k", [0, 3, 10])
def test_mult_ke_mask(orc, torch_dev, k):
    """maskflag kinetic apply (no ECS) vs the oracle, FP64 and FP32 matrix
    cores, and through sparseops_matel."""
    from sincdvr import SincDVR, inputs
    n, dx = 21, 0.4
    phi = inputs.gaussian_orbitals(n, dx, 3, 3, seed=8)
    say("mult_ke mask k=%d" % k)
    s = SincDVR(n, dx, 3, masknumpoints=k)
    ke, _, _ = orc.sinedvr(n, dx)
    m = orc.mask_function(n, k)
    ref = orc.mult_ke_mask(n, ke, m, phi)
    assert relerr(s.mult_ke(phi), ref) < 1e-12
    op = s.sparseops_matel(1, 3, phi)
    assert relerr(op, np.conj(phi).T @ ref) < 1e-12
    s.set_precision(1)
    assert relerr(s.mult_ke(phi), ref) < 2e-6

