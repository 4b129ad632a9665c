"""Pins the CPU oracle (oracle/sincdvr_oracle.c) without trusting it:
closed forms, brute force on tiny grids, invariants of the defining linear
systems, an independent FFT library (numpy's pocketfft), the reference's own
SineDVR / DFFTPACK compiled into oracle/_ref, and the committed fixtures that
the reference's compiled pieces produced (tests/golden/)."""
import os

import numpy as np
import pytest

from conftest import relerr

GOLD = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------- a1 sineDVR
def test_sinedvr_closed_form(orc):
    for n, dx in [(5, 0.4), (21, 0.4), (41, 0.3)]:
        ke, fd, pts = orc.sinedvr(n, dx)
        i = np.arange(1 - n, n)
        with np.errstate(divide="ignore"):
            ref = np.where(i == 0, np.pi ** 2 / 3, 2.0 * (-1.0) ** i / np.where(i == 0, 1, i) ** 2) / 2 / dx ** 2
        assert relerr(ke, ref) < 1e-15
        assert np.all(ke == ke[::-1])                        # symmetric Toeplitz stencil
        assert abs(pts[0] + pts[-1]) < 1e-14 and abs(pts[1] - pts[0] - dx) < 1e-14
        # symbol of the stencil is k^2/2 on the band (sum over all offsets,
        # infinite-grid limit approximated by a long stencil)
    ke, _, _ = orc.sinedvr(2001, 0.3)
    m = np.arange(-2000, 2001)
    for k in (0.5, 2.0, 5.0):
        sym = np.sum(ke * np.cos(k * m * 0.3))
        assert abs(sym - k * k / 2) < 2e-3 * max(1, k * k)


def test_sinedvr_matches_reference_build(orc):
    r = orc.ref_sinedvr(21, 0.4)
    if r is None:
        pytest.skip("oracle/_ref not built (reference sources not mounted)")
    for a, b in zip(orc.sinedvr(21, 0.4), r):
        assert np.array_equal(a, b)


def test_sinedvr_golden(orc):
    g = np.load(os.path.join(GOLD, "ref_sinedvr.npz"))
    for key in g.files:
        n, dx = key.split("_")[1:]
        ke, fd, pts = orc.sinedvr(int(n), float(dx))
        assert np.array_equal(np.concatenate([ke, fd, pts]), g[key])


# ------------------------------------------------------------ a8/a9 1-D FFT
@pytest.mark.parametrize("M", [1, 2, 7, 42, 45, 64, 82, 90, 162])
def test_fft_vs_numpy(orc, M):
    rng = np.random.default_rng(M)
    x = rng.standard_normal((M, 3)) + 1j * rng.standard_normal((M, 3))
    assert relerr(orc.fft1d(x), np.fft.fft(x, axis=0)) < 2e-15 * max(1, np.log2(M))


def test_fft_golden_dfftpack(orc):
    """Vectors produced by the reference's DFFTPACK zffti+zfftf (compiled from
    S/DFFTPACK into oracle/_ref by tests/golden/make_golden.py)."""
    g = np.load(os.path.join(GOLD, "ref_zfftf.npz"))
    for key in [k for k in g.files if k.startswith("in_")]:
        M = int(key.split("_")[1])
        x = g[key]
        assert relerr(orc.fft1d(x), g["out_%d" % M]) < 2e-15 * max(1, np.log2(M))


# ----------------------------------------------------------- a3 GETINVERSE
def _t1d(ns, d):
    i = np.arange(ns)
    D = i[:, None] - i[None, :]
    with np.errstate(divide="ignore"):
        T = np.where(D == 0, np.pi ** 2 / (6 * d * d), (-1.0) ** D / (d * d * np.where(D == 0, 1, D) ** 2))
    return T


@pytest.fixture(scope="module")
def tooth_small(orc):
    # N = 5, nbig = 60, nsmall = 8 keeps this fast while exercising every branch
    return orc.getinverse(5, 60, 8, 0.4, want_tr=True)


def test_getinverse_solves_linear_system(tooth_small):
    _, TR, Q = tooth_small
    T = _t1d(TR.shape[0], 0.4)
    res = np.einsum("ai,ijk->ajk", T, Q) + np.einsum("bj,ijk->ibk", T, Q) + np.einsum("ck,ijk->ijc", T, Q)
    assert relerr(res, TR) < 1e-13


def test_getinverse_tr_brute_force(tooth_small):
    """TR(i,j,k) = -sum_p T(i-p)R(p,j,k)+T(j-p)R(i,p,k)+T(k-p)R(i,j,p) (+2 pi at 0)."""
    _, TR, _ = tooth_small
    d, nbig, ns = 0.4, 60, 8

    def R(i, j, k):
        return 0.0 if i == j == k == 0 else d ** 3 / np.sqrt((i * d) ** 2 + (j * d) ** 2 + (k * d) ** 2)

    def T(m):
        return np.pi ** 2 / (6 * d * d) if m == 0 else (-1.0) ** m / (d * d * m * m)

    for (i, j, k) in [(0, 0, 0), (1, 0, 0), (2, -3, 1), (-8, 8, 5), (0, 7, -2)]:
        s = sum(T(i - p) * R(p, j, k) + T(j - p) * R(i, p, k) + T(k - p) * R(i, j, p) for p in range(-nbig, nbig + 1))
        ref = -s + (2 * np.pi if i == j == k == 0 else 0.0)
        assert abs(TR[i + ns, j + ns, k + ns] - ref) < 1e-12 * max(1, abs(ref))


def test_getinverse_far_field_and_symmetry(orc):
    N, d = 12, 0.3
    t = orc.getinverse(N, 60, 8, d)
    idx = np.arange(-N, N + 1)
    I, J, K = np.meshgrid(idx, idx, idx, indexing="ij")
    far = np.maximum(np.maximum(abs(I), abs(J)), abs(K)) > 8
    r = np.sqrt((I * d) ** 2 + (J * d) ** 2 + (K * d) ** 2)
    assert np.allclose(t[far], d ** 3 / r[far], rtol=1e-15, atol=0)
    # reflections and axis permutations of the cube
    for op in (lambda a: a[::-1], lambda a: a[:, ::-1], lambda a: a.transpose(1, 0, 2), lambda a: a.transpose(2, 1, 0)):
        assert relerr(op(t), t) < 1e-13


def test_coulomb_kernel_is_inverse_laplacian(orc):
    """Near the origin the corrected kernel satisfies the sinc-DVR Poisson
    equation T3d K = 2 pi delta / D^3 (up to the nbig truncation)."""
    n, d = 7, 0.4
    k = orc.coulomb_kernel(n, d, nbig=100, nsmall=12)
    kd = k * d ** 3
    ns = 2 * n - 1
    T = _t1d(ns, d)
    lap = np.einsum("ai,ijk->ajk", T, kd) + np.einsum("bj,ijk->ibk", T, kd) + np.einsum("ck,ijk->ijc", T, kd)
    c = n - 1
    # only the centre is free of truncation-boundary effects
    assert abs(lap[c, c, c] - 2 * np.pi) < 0.05 * 2 * np.pi
    assert np.abs(lap[c - 1:c + 2, c - 1:c + 2, c - 1:c + 2]).sum() - abs(lap[c, c, c]) < 0.2


def test_coulomb_kernel_gaussian_charge(orc):
    """Physical pin of GETINVERSE + kernel placement: the grid potential of a
    normalised Gaussian charge (alpha/pi)^(3/2) exp(-alpha r^2), sampled with
    weight D^3, is erf(sqrt(alpha) r)/r (2 sqrt(alpha/pi) at r = 0).  A lost
    2 pi, a wrong sign of TR or a wrong far-field scale fails by far more."""
    from scipy.special import erf
    for n, d, al in [(21, 0.4, 1.0), (31, 0.3, 1.0)]:
        kern = orc.coulomb_kernel(n, d)
        _, _, pts = orc.sinedvr(n, d)
        X = np.meshgrid(pts, pts, pts, indexing="ij")
        c = (0.13, -0.21, 0.05)
        r2 = sum((x - cc) ** 2 for x, cc in zip(X, c))
        rho = (al / np.pi) ** 1.5 * np.exp(-al * r2) * d ** 3
        V = orc.op_tinv(n, kern, rho.ravel(order="F"))[:, 0].reshape((n, n, n), order="F")
        r = np.sqrt(r2)
        exact = erf(np.sqrt(al) * r) / r
        assert relerr(V, exact) < 1e-6                   # observed 3e-7 / 4e-7


# ------------------------------------------------------- a5/a6/a7 op_tinv
@pytest.mark.parametrize("n", [3, 5, 6])
def test_op_tinv_brute_force(orc, n):
    rng = np.random.default_rng(n)
    kern = orc.coulomb_kernel(n, 0.4, nbig=40, nsmall=8)
    den = rng.standard_normal((n ** 3, 3)) + 1j * rng.standard_normal((n ** 3, 3))
    direct = orc.toeplitz_direct(n, kern, den)
    assert relerr(orc.op_tinv(n, kern, den), direct) < 1e-14            # reference layout M = 2n
    for M in (2 * n - 1, 2 * n + 3, 3 * n):                               # any M >= 2n-1
        assert relerr(orc.op_tinv(n, kern, den, M=M), direct) < 1e-14
    assert relerr(orc.op_tinv(n, kern, den, circsize=1), direct) < 1e-14


def test_op_tinv_delta_linearity_conjugation(orc):
    n = 5
    kern = orc.coulomb_kernel(n, 0.4, nbig=40, nsmall=8)
    L = 2 * n - 1
    for (a, b, c) in [(0, 0, 0), (4, 1, 2), (2, 2, 2)]:
        den = np.zeros(n ** 3, dtype=complex)
        den[a + n * (b + n * c)] = 1.0
        v = orc.op_tinv(n, kern, den)[:, 0].reshape((n, n, n), order="F")
        ref = kern[(n - 1 - a):(2 * n - 1 - a), (n - 1 - b):(2 * n - 1 - b), (n - 1 - c):(2 * n - 1 - c)]
        assert ref.shape == (n, n, n) and L == kern.shape[0]
        assert relerr(v, ref) < 1e-14
    rng = np.random.default_rng(1)
    x = rng.standard_normal(n ** 3) + 1j * rng.standard_normal(n ** 3)
    y = rng.standard_normal(n ** 3) + 1j * rng.standard_normal(n ** 3)
    vx, vy = orc.op_tinv(n, kern, x), orc.op_tinv(n, kern, y)
    assert relerr(orc.op_tinv(n, kern, 2 * x - 3j * y), 2 * vx - 3j * vy) < 1e-14
    assert relerr(orc.op_tinv(n, kern, np.conj(x)), np.conj(vx)) < 1e-14


@pytest.mark.parametrize("n", [5, 7, 10])
def test_op_tinv_reference_golden(orc, n):
    """The oracle's op_tinv against what the reference's own circ3d_sub /
    myzfft3d / DFFTPACK (compiled into oracle/_ref) returned for the same
    seeded kernel and densities (tests/golden/ref_op_tinv.npz), at the
    reference's M = 2n and at other circulant edges."""
    g = np.load(os.path.join(GOLD, "ref_op_tinv.npz"))
    kern, den, out = g["kern_%d" % n], g["den_%d" % n], g["out_%d" % n]
    assert relerr(orc.op_tinv(n, kern, den), out) < 1e-14
    assert relerr(orc.op_tinv(n, kern, den, M=2 * n + 3), out) < 1e-14
    assert relerr(orc.op_tinv(n, kern, den, circsize=1), out) < 1e-14


def test_op_tinv_vs_reference_build(orc):
    """The oracle against the reference's convolution built from its sources
    (oracle/_ref/libref_conv.so) on pair densities of seeded orbitals, with
    circsize chunks as op_tinv_notscaled makes them."""
    if orc.conv_lib() is None:
        pytest.skip("oracle/_ref not built (reference sources not mounted)")
    from sincdvr import inputs
    for n, dx, N in [(12, 0.4, 3), (21, 0.4, 4)]:
        phi = inputs.gaussian_orbitals(n, dx, N, 3, seed=n)
        den = np.asfortranarray(phi * np.conj(phi[:, [1]]))
        kern = orc.coulomb_kernel(n, dx)
        ref = orc.ref_op_tinv(n, kern, den, circsize=1 if n == 12 else N)
        assert relerr(orc.op_tinv(n, kern, den), ref) < 1e-14


def test_op_tinv_size_error(orc):
    with pytest.raises(ValueError):
        orc.op_tinv(3, np.zeros((5, 5, 5)), np.zeros((27, 3), dtype=complex), circsize=2)


# ------------------------------------------------------ a10 twoe matel
def _brute_twoe(orc, n, kern, phi1, phi2, cnorm, low, high):
    N = phi1.shape[1]
    cj = (lambda z: z) if cnorm else np.conj
    tr = np.zeros((n ** 3, N, high - low + 1), dtype=complex, order="F")
    for b2 in range(low, high + 1):
        for a2 in range(1, N + 1):
            den = phi1[:, a2 - 1] * cj(phi2[:, b2 - 1])
            tr[:, a2 - 1, b2 - low] = orc.toeplitz_direct(n, kern, den)[:, 0]
    tm = np.zeros((N, N, N, high - low + 1), dtype=complex, order="F")
    for b1 in range(N):
        for a1 in range(N):
            myden = phi1[:, a1] * cj(phi2[:, b1])
            for b2 in range(high - low + 1):
                for a2 in range(N):
                    tm[a1, b1, a2, b2] = np.sum(tr[:, a2, b2] * myden)
    return tm, tr


@pytest.mark.parametrize("cnorm", [False, True])
@pytest.mark.parametrize("lowhigh", [(1, 3), (2, 3), (2, 2)])
def test_twoe_matel_bookkeeping(orc, cnorm, lowhigh):
    n, N = 4, 3
    rng = np.random.default_rng(7)
    kern = orc.coulomb_kernel(n, 0.5, nbig=30, nsmall=6)
    phi1 = np.asfortranarray(rng.standard_normal((n ** 3, N)) + 1j * rng.standard_normal((n ** 3, N)))
    phi2 = np.asfortranarray(rng.standard_normal((n ** 3, N)) + 1j * rng.standard_normal((n ** 3, N)))
    low, high = lowhigh
    for p2 in (phi1, phi2):
        tm, tr = orc.twoe_matel(n, kern, phi1, p2, cnorm=cnorm, low=low, high=high)
        btm, btr = _brute_twoe(orc, n, kern, phi1, p2, cnorm, low, high)
        assert relerr(tr, btr) < 1e-13
        assert relerr(tm, btm) < 1e-13


def test_twoe_matel_symmetries(orc):
    n, N = 4, 3
    rng = np.random.default_rng(3)
    kern = orc.coulomb_kernel(n, 0.5, nbig=30, nsmall=6)
    phi = np.asfortranarray(rng.standard_normal((n ** 3, N)) + 1j * rng.standard_normal((n ** 3, N)))
    tm, tr = orc.twoe_matel(n, kern, phi, phi, cnorm=False)
    assert relerr(tr.transpose(0, 2, 1), np.conj(tr)) < 1e-14           # hermnorm V_ba = conj V_ab
    assert relerr(tm, tm.transpose(2, 3, 0, 1)) < 1e-13                  # Coulomb symmetry
    tmc, trc = orc.twoe_matel(n, kern, phi, phi, cnorm=True)
    assert np.array_equal(trc, trc.transpose(0, 2, 1))                   # cnorm V_ab == V_ba bitwise
    z, _ = orc.twoe_matel(n, kern, phi, phi, notwo=True)
    assert not z.any()


# --------------------------------------------------------------- a11 KE
@pytest.mark.parametrize("ecs", [False, True])
def test_mult_ke_dense(orc, ecs):
    n, d = 5, 0.4
    ke, _, _ = orc.sinedvr(n, d)
    ij = sd = None
    if ecs:
        ij, sd = orc.ecs_arrays(n, d, 0.3, 1.0, 0.3, 0.45)
    rng = np.random.default_rng(11)
    x = rng.standard_normal(n ** 3) + 1j * rng.standard_normal(n ** 3)
    a = orc.mult_ke(n, ke, x, ij, sd)[:, 0]
    b = orc.mult_ke_dense(n, ke, x, ij, sd)
    assert relerr(a, b) < 1e-14


def test_mult_ke_wave_packet(orc):
    """The sinc-DVR KE is spectrally exact for a well-resolved wave packet
    f = exp(i k.r - a r^2):  -1/2 lap f = -1/2 sum_d [(i k_d - 2 a x_d)^2 - 2a] f."""
    n, d, al = 41, 0.3, 0.5
    ke, _, pts = orc.sinedvr(n, d)
    k = np.array([0.7, -1.1, 0.4])
    X, Y, Z = np.meshgrid(pts, pts, pts, indexing="ij")
    f = np.exp(1j * (k[0] * X + k[1] * Y + k[2] * Z) - al * (X * X + Y * Y + Z * Z))
    lap = sum((1j * kd - 2 * al * xd) ** 2 - 2 * al for kd, xd in zip(k, (X, Y, Z))) * f
    out = orc.mult_ke(n, ke, f.ravel(order="F"))[:, 0].reshape((n, n, n), order="F")
    assert relerr(out, -0.5 * lap) < 1e-7


def test_ecs_functions(orc):
    th, st, R0, S = 0.3, 1.0, 3.75, 2.25
    fac = np.exp(1j * th) * st
    assert orc.ecs_point(1.0, th, st, R0, S)[1] == 1.0
    assert abs(orc.ecs_point(6.5, th, st, R0, S)[1] - fac) < 1e-15
    assert abs(orc.ecs_point(-6.0, th, st, R0, S)[1] - fac) < 1e-15      # |x| = R0 + S reaches fac
    # J = F', J' = dJ/dx, J'' = d^2J/dx^2 inside the ramps (finite differences)
    h = 1e-5
    for x in (4.1, 5.0, -4.6, -5.7):
        F0, J0, D0, DD0 = orc.ecs_point(x, th, st, R0, S)
        Fp, Jp, Dp, _ = orc.ecs_point(x + h, th, st, R0, S)
        Fm, Jm, Dm, _ = orc.ecs_point(x - h, th, st, R0, S)
        assert abs((Fp - Fm) / (2 * h) - J0) < 1e-8
        assert abs((Jp - Jm) / (2 * h) - D0) < 1e-8
        assert abs((Dp - Dm) / (2 * h) - DD0) < 1e-7
    ij, sd = orc.ecs_arrays(5, 0.4, th, st, 10.0, 1.0)                    # ramp outside the grid
    assert np.all(ij == 1.0) and not sd.any()


def _ecs_axis(orc, pts, ecs):
    v = np.array([orc.ecs_point(p, *ecs) for p in pts])
    return v[:, 1], v[:, 2], v[:, 3]                                       # J, J', J''


@pytest.mark.parametrize("ecs", [(0.3, 1.0, 3.75, 2.25), (0.5, 1.3, 2.0, 3.0)])
def test_ecs_scalediag_pointwise(orc, ecs):
    """scalediag (SD/init_sinc_new.f90:517-521) is the diagonal remainder that
    turns J^-1 (-1/2 d^2) J^-1 into the symmetric complex-scaled kinetic
    operator -1/2 A d(A^2 d(A f)), A = J^(-1/2):
        s(x) = -1/2 [A (A^2 A')' - A^2 (A^2)''].
    Evaluated here by finite differences of J alone (no J', J'' from the
    oracle), at the grid points of one axis with the other two at the centre
    (their terms vanish: |x| <= R0)."""
    n, d = 41, 0.3
    _, _, pts = orc.sinedvr(n, d)
    ij, sd = orc.ecs_arrays(n, d, *ecs)
    c = n // 2
    assert abs(pts[c]) < 1e-12
    h = 1e-3
    w1 = np.array([1, -8, 0, 8, -1]) / (12 * h)                          # 4th-order first derivative
    w2 = np.array([-1, 16, -30, 16, -1]) / (12 * h * h)                  # 4th-order second derivative
    off = np.arange(-2, 3) * h

    def A(x):
        return 1.0 / np.sqrt(orc.ecs_point(x, *ecs)[1])

    def dA(x):
        return w1 @ np.array([A(x + o) for o in off])

    R0, S = ecs[2], ecs[3]
    checked = 0
    for i, x in enumerate(pts):
        got = sd[i + n * c + n * n * c]
        if abs(abs(x) - R0) < 0.05 or abs(abs(x) - R0 - S) < 0.05:
            continue                                                       # stencil would straddle a ramp end
        u = w1 @ np.array([A(x + o) ** 2 * dA(x + o) for o in off])       # (A^2 A')'
        v = w2 @ np.array([A(x + o) ** 2 for o in off])                    # (A^2)''
        ref = -0.5 * (A(x) * u - A(x) ** 2 * v)
        assert abs(got - ref) < 1e-7, (x, got, ref)
        checked += R0 < abs(x) < R0 + S
    assert checked >= 6                                                    # points inside the ramps


@pytest.mark.parametrize("ecs", [(0.3, 1.0, 2.0, 4.0), (0.5, 1.3, 2.0, 4.0)])
def test_ecs_mult_ke_closed_form(orc, ecs):
    """ECS mult_ke (SD/coreproject.f90:1341-1388) on a smooth, well-resolved
    complex Gaussian whose support lies inside the ramps of two axes, against
    the closed form of the complex-scaled kinetic operator,
        sum_d -1/2 [J^-2 f'' - 2 J^-3 J' f' + (5/4 J^-4 J'^2 - 1/2 J^-3 J'') f],
    with f', f'' analytic and J, J', J'' from the ECS functions (pinned by
    finite differences in test_ecs_functions).  The same comparison rejects
    scalediag tables with the coefficients swapped, the sign flipped or a
    power of J changed (errors 2e-4 .. 7e-3 against < 1e-7 here)."""
    n, d = 81, 0.15
    ke, _, pts = orc.sinedvr(n, d)
    ij, sd = orc.ecs_arrays(n, d, *ecs)
    X = np.meshgrid(pts, pts, pts, indexing="ij")
    c, al, kv = (4.0, -3.6, 0.3), 4.0, (0.8, -0.5, 0.3)
    f = np.exp(1j * sum(k * x for k, x in zip(kv, X)) - al * sum((x - cc) ** 2 for x, cc in zip(X, c)))
    J, D, DD = _ecs_axis(orc, pts, ecs)
    exact = np.zeros_like(f)
    for ax in range(3):
        g = 1j * kv[ax] - 2 * al * (X[ax] - c[ax])
        sh = [1, 1, 1]
        sh[ax] = n
        j, dj, ddj = (a.reshape(sh) for a in (J, D, DD))
        exact += -0.5 * ((g * g - 2 * al) * f / j ** 2 - 2 * dj / j ** 3 * g * f
                         + (1.25 * dj ** 2 / j ** 4 - 0.5 * ddj / j ** 3) * f)
    out = orc.mult_ke(n, ke, f.ravel(order="F"), ij, sd)[:, 0].reshape((n, n, n), order="F")
    assert relerr(out, exact) < 1e-6                                       # observed 3e-8

    I3 = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")

    def full(s1):
        return (s1[I3[0]] + s1[I3[1]] + s1[I3[2]]).ravel(order="F")

    good = 3 / 8 * D ** 2 / J ** 4 - 1 / 4 * DD / J ** 3
    assert np.abs(full(good) - sd).max() < 1e-15
    for bad in (1 / 4 * D ** 2 / J ** 4 - 3 / 8 * DD / J ** 3, -good,
                3 / 8 * D ** 2 / J ** 4 - 1 / 4 * DD / J ** 2, 3 / 8 * D ** 2 / J ** 3 - 1 / 4 * DD / J ** 3):
        o = orc.mult_ke(n, ke, f.ravel(order="F"), ij, full(bad))[:, 0].reshape((n, n, n), order="F")
        assert relerr(o, exact) > 1e-5


# ------------------------------------------------------ a12/a13 mean field
def test_mean_field(orc):
    rng = np.random.default_rng(5)
    n3, N = 64, 3
    tr = rng.standard_normal((n3, N, N)) + 1j * rng.standard_normal((n3, N, N))
    g = rng.standard_normal((N, N, N, N)) + 1j * rng.standard_normal((N, N, N, N))
    rp = orc.reducedpot(np.asfortranarray(tr), np.asfortranarray(g))
    ref = tr.reshape(n3, N * N, order="F") @ g.reshape(N * N, N * N, order="F")
    assert relerr(rp.reshape(n3, N * N, order="F"), ref) < 1e-14
    phi = rng.standard_normal((n3, N)) + 1j * rng.standard_normal((n3, N))
    out = orc.mult_reducedpot(2, 3, phi, rp[:, :, 1:3])
    ref2 = np.einsum("xki,xk->xi", rp[:, :, 1:3], phi)
    assert relerr(out, ref2) < 1e-14


# ------------------------------------------------- a4 nuclear potential
def test_nuclear_potential_oracle(orc):
    """SD/sincproject.f90:192-250.  Grid point i of a nucleus with half-spacing
    shift cs sees offset i + (cs - n + 1)/2, i.e. the nucleus sits at
    -cs*spacing/2.  Outside the kernel cube pot = -Z/r exactly (8-byte sqrt),
    inside it is -Z times the corrected kernel; the 4-byte sqrt of builds
    without -r8 differs only outside; the centershift parity rule holds."""
    n, d = 9, 0.4
    kern = orc.coulomb_kernel(n, d, nbig=40, nsmall=8)
    pts = (np.arange(1, n + 1) * d) - (d + n * d) / 2
    X, Y, Z = np.meshgrid(pts, pts, pts, indexing="ij")
    cs, zc = [[0, 0, 16], [0, 0, -4]], [1.0, 2.0]
    pot = orc.nuclear_pot(n, d, kern, cs, zc).reshape((n, n, n), order="F")
    ref = np.zeros_like(pot)
    inside_all = np.ones((n, n, n), bool)
    for (cx, cy, cz), q in zip(cs, zc):
        o = [np.arange(n) + (c - n + 1) // 2 for c in (cx, cy, cz)]
        O = np.meshgrid(*o, indexing="ij")
        inside = (abs(O[0]) <= n - 1) & (abs(O[1]) <= n - 1) & (abs(O[2]) <= n - 1)
        inside_all &= inside
        far = -q / (d * np.sqrt(np.where(inside, 1, O[0] ** 2 + O[1] ** 2 + O[2] ** 2)))
        kv = np.zeros_like(far)
        kv[inside] = -q * kern[(O[0] + n - 1)[inside], (O[1] + n - 1)[inside], (O[2] + n - 1)[inside]]
        ref += np.where(inside, kv, far)
        # the nucleus is at -cs*spacing/2: far values are -Z/|r - r_c|
        r = np.sqrt((X + cx * d / 2) ** 2 + (Y + cy * d / 2) ** 2 + (Z + cz * d / 2) ** 2)
        assert np.allclose(far[~inside], -q / r[~inside], rtol=1e-14)
    assert np.allclose(pot, ref, rtol=1e-14, atol=0)
    assert (~inside_all).any() and inside_all.any()
    p4 = orc.nuclear_pot(n, d, kern, cs, zc, realkind4=True).reshape((n, n, n), order="F")
    assert np.array_equal(p4[inside_all], pot[inside_all])
    assert 0 < np.abs(p4 - pot).max() < 1e-6 * np.abs(pot).max()
    with pytest.raises(ValueError):
        orc.nuclear_pot(n, d, kern, [[1, 0, 0]], [1.0])                # odd shift with odd n


@pytest.mark.parametrize("cnorm", [False, True])
def test_one_matel_brute_force(orc, cnorm):
    """sparseops_matel / pot_matel restatement against explicit sums with the
    dense n^3 x n^3 kinetic matrix (ECS on), sub-range and bra != ket."""
    n, d, N = 5, 0.4, 3
    ke, _, _ = orc.sinedvr(n, d)
    ij, sd = orc.ecs_arrays(n, d, 0.3, 1.0, 0.3, 0.45)
    rng = np.random.default_rng(5)
    a = np.asfortranarray(rng.standard_normal((n ** 3, N)) + 1j * rng.standard_normal((n ** 3, N)))
    b = np.asfortranarray(rng.standard_normal((n ** 3, N)) + 1j * rng.standard_normal((n ** 3, N)))
    pot = rng.standard_normal(n ** 3)
    got = orc.one_matel(n, ke, a, b, 2, 3, cnorm=cnorm, invjac=ij, scalediag=sd)
    gotp = orc.one_matel(n, ke, a, b, 2, 3, pot=pot, cnorm=cnorm)
    for jj, col in enumerate((1, 2)):
        tb = orc.mult_ke_dense(n, ke, b[:, col], ij, sd)
        for i in range(N):
            bra = a[:, i] if cnorm else np.conj(a[:, i])
            assert abs(got[i, jj] - np.sum(bra * tb)) < 1e-12 * np.abs(tb).max() * np.abs(a).sum()
            assert abs(gotp[i, jj] - np.sum(bra * pot * b[:, col])) < 1e-12 * np.abs(b).sum() * 10


def test_mask_function_closed_form(orc):
    """maskfunction: 1 inside, (1 + cos(pi ii / (k + 1))) / 2 at the ii-th point
    in from either end; overlapping ends multiply."""
    n, k = 11, 3
    m = orc.mask_function(n, k)
    r = [(1 + np.cos(np.pi * ii / (k + 1))) / 2 for ii in range(1, k + 1)]
    want = np.ones(n)
    for ii in range(1, k + 1):
        want[k - ii] *= r[ii - 1]
        want[n - k + ii - 1] *= r[ii - 1]
    assert np.array_equal(m, want)
    assert m[0] < m[1] < m[2] < 1 and m[5] == 1
    assert np.array_equal(orc.mask_function(n, 0), np.ones(n))
    mo = orc.mask_function(4, 3)          # overlapping ends
    assert mo[1] == ((1 + np.cos(np.pi * 2 / 4)) / 2) * ((1 + np.cos(np.pi * 1 / 4)) / 2)


def test_mult_ke_mask_is_similarity(orc):
    """Masked kinetic apply = M . T . M^-1 with M the diagonal product mask
    (the per-axis mask factors of the other axes cancel)."""
    n, d = 7, 0.4
    ke, _, _ = orc.sinedvr(n, d)
    m = orc.mask_function(n, 2)
    M = np.einsum("i,j,k->kji", m, m, m).ravel()          # x fastest
    rng = np.random.default_rng(3)
    x = rng.standard_normal(n ** 3) + 1j * rng.standard_normal(n ** 3)
    got = orc.mult_ke_mask(n, ke, m, x)[:, 0]
    want = M * orc.mult_ke_dense(n, ke, x / M)
    assert relerr(got, want) < 1e-13


def test_actreduced00_restatement(orc):
    """Orbital operator restatement: with denmat = 1, no potentials and no
    mean field it is the kinetic apply; the projected result is orthogonal to
    the projection orbitals; dentime multiplies by timefac . invdenmat."""
    n, d, N = 5, 0.4, 3
    ke, _, _ = orc.sinedvr(n, d)
    rng = np.random.default_rng(8)
    phi = np.asfortranarray(rng.standard_normal((n ** 3, N)) + 1j * rng.standard_normal((n ** 3, N)))
    q, _ = np.linalg.qr(phi)
    zero_rp = np.zeros((n ** 3, N, N), dtype=complex, order="F")
    out = orc.actreduced00(n, ke, phi, np.eye(N), zero_rp)
    assert relerr(out, orc.mult_ke(n, ke, phi)) < 1e-14
    dm = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    rp = np.asfortranarray(rng.standard_normal((n ** 3, N, N)) + 1j * rng.standard_normal((n ** 3, N, N)))
    pot = rng.standard_normal(n ** 3)
    full = orc.actreduced00(n, ke, phi, dm, rp, pot=pot, projspfs=q)
    assert np.abs(np.conj(q).T @ full).max() < 1e-12 * np.abs(full).max()
    inv = rng.standard_normal((N, N)) + 0j
    a = orc.actreduced00(n, ke, phi, dm, rp, pot=pot)
    b = orc.actreduced00(n, ke, phi, dm, rp, pot=pot, invdenmat=inv, timefac=-1j)
    assert relerr(b, -1j * (a @ inv)) < 1e-14


def test_dipoles_and_lenmultiply(orc):
    """Length-gauge pulse restatement.  Dipoles: the centred sine-DVR grid
    (spacing apart, summing to zero); under ECS, F(x) = x inside the unscaled
    region, the closed-form straight line beyond the ramp, and dF/dx = J of the
    scaling (central differences).  lenmultiply: the x, y and z terms of a
    separable orbital f(x) g(y) h(z) are x f g h, f y g h, f g z h (an axis
    swap fails), the sum is linear in the field, and <phi|x|phi> of a
    Gaussian centred at x0 well inside the grid is x0."""
    n, d = 9, 0.5
    dip = orc.dipoles_1d(n, d)
    assert np.allclose(np.diff(dip.real), d, atol=1e-14) and abs(dip.sum()) < 1e-13
    theta, stretch, R0, S = 0.4, 1.2, 1.0, 0.75
    dipe = orc.dipoles_1d(n, d, (theta, stretch, R0, S))
    fac = stretch * np.exp(1j * theta)
    for x, f in zip(dip.real, dipe):
        if abs(x) <= R0:
            assert f == x
        elif abs(x) >= R0 + S:
            assert abs(f - (x + (fac - 1) * (x - np.sign(x) * (R0 + S / 2)))) < 1e-14
    for x in (-1.6, -1.3, 1.2, 1.5, 1.7):
        h = 1e-5
        dF = (orc.ecs_point(x + h, theta, stretch, R0, S)[0] - orc.ecs_point(x - h, theta, stretch, R0, S)[0]) / (2 * h)
        assert abs(dF - orc.ecs_point(x, theta, stretch, R0, S)[1]) < 1e-8
    rng = np.random.default_rng(5)
    f, g, h = [rng.standard_normal(n) + 1j * rng.standard_normal(n) for _ in range(3)]
    phi = np.einsum("k,j,i->kji", h, g, f).reshape(-1, 1)           # x fastest in the point index
    for axis, (a, b, c) in enumerate([(dip * f, g, h), (f, dip * g, h), (f, g, dip * h)]):
        t = [0, 0, 0]
        t[axis] = 1.0
        want = np.einsum("k,j,i->kji", c, b, a).reshape(-1, 1)
        assert relerr(orc.lenmultiply(n, phi, dip, t), want) < 1e-14
    E1, E2 = (0.3, -0.2j, 0.1 + 0.4j), (-0.5j, 0.7, 0.2)
    lhs = orc.lenmultiply(n, phi, dipe, [a + b for a, b in zip(E1, E2)])
    assert relerr(lhs, orc.lenmultiply(n, phi, dipe, E1) + orc.lenmultiply(n, phi, dipe, E2)) < 1e-14
    n2, x0 = 25, 0.6
    dg = orc.dipoles_1d(n2, 0.3).real
    X, Y, Z = np.meshgrid(dg, dg, dg, indexing="ij")
    gau = np.exp(-2 * ((X - x0) ** 2 + Y ** 2 + Z ** 2)).transpose(2, 1, 0).reshape(-1, 1) + 0j
    xphi = orc.lenmultiply(n2, gau, dg + 0j, [1.0, 0, 0])
    assert abs(np.vdot(gau, xphi) / np.vdot(gau, gau) - x0) < 1e-10
    ke, _, _ = orc.sinedvr(n, d)
    zero_rp = np.zeros((n ** 3, 1, 1), dtype=complex, order="F")
    a = orc.actreduced00(n, ke, phi, np.eye(1), zero_rp)
    b = orc.actreduced00(n, ke, phi, np.eye(1), zero_rp, tdpot=E1, dip=dipe)
    assert relerr(b - a, orc.lenmultiply(n, phi, dipe, E1)) < 1e-13


def test_velmultiply_restatement(orc):
    """Velocity-gauge pulse restatement (velmultiply, first-derivative sinc
    matrix).  Closed form: for a Gaussian well inside the grid the axis-z term
    is (-i A_z)(-d/dz) of it (spectral accuracy of the sinc derivative), and
    each axis differentiates along its own coordinate only (an axis swap
    fails).  Structure: -F is real antisymmetric, so with ECS weights W on both
    sides the operator is antisymmetric under the bilinear (c-norm) product,
    and without weights and with real A it is Hermitian; it is linear in A; a
    function vanishing wherever J != 1 only receives the post-weight sqrt(1/J);
    under the edge mask it is the similarity M . op . M^-1."""
    n, d = 31, 0.3
    _, fd, pts = orc.sinedvr(n, d)
    X, Y, Z = np.meshgrid(pts, pts, pts, indexing="ij")
    z0, al = 0.4, 1.2
    gau = np.exp(-al * (X ** 2 + Y ** 2 + (Z - z0) ** 2))
    f = gau.transpose(2, 1, 0).reshape(-1, 1) + 0j                       # x fastest
    dz = (-2 * al * (Z - z0) * gau).transpose(2, 1, 0).reshape(-1, 1)
    dx = (-2 * al * X * gau).transpose(2, 1, 0).reshape(-1, 1)
    Az = 0.7 - 0.2j
    got = orc.velmultiply(n, f, (0, 0, Az), fd)
    assert relerr(got, (-1j * Az) * (-dz)) < 1e-8
    assert relerr(orc.velmultiply(n, f, (0.5, 0, 0), fd), (-0.5j) * (-dx)) < 1e-8
    assert relerr(got, (-1j * Az) * (-dx)) > 0.5
    n, d = 7, 0.4
    _, fd, pts = orc.sinedvr(n, d)
    rng = np.random.default_rng(12)
    u, v = [rng.standard_normal((n ** 3, 1)) + 1j * rng.standard_normal((n ** 3, 1)) for _ in range(2)]
    A1, A2 = (0.3 - 0.1j, -0.2, 0.5j), (0.1, 0.4 + 0.2j, -0.3)
    assert relerr(orc.velmultiply(n, u, [a + b for a, b in zip(A1, A2)], fd),
                  orc.velmultiply(n, u, A1, fd) + orc.velmultiply(n, u, A2, fd)) < 1e-14
    Ar = (0.3, -0.2, 0.5)
    lhs = np.vdot(v, orc.velmultiply(n, u, Ar, fd))
    assert abs(lhs - np.conj(np.vdot(u, orc.velmultiply(n, v, Ar, fd)))) < 1e-13 * abs(lhs)
    ij, _ = orc.ecs_arrays(n, d, 0.4, 1.2, 0.5, 0.4)
    sj = np.sqrt(ij)
    a = np.sum(v * orc.velmultiply(n, u, A1, fd, invsqrtjac=sj))
    b = np.sum(orc.velmultiply(n, v, A1, fd, invsqrtjac=sj) * u)
    assert abs(a + b) < 1e-13 * abs(a)
    Xs, Ys, Zs = np.meshgrid(pts, pts, pts, indexing="ij")
    mask_in = ((np.abs(Xs) <= 0.5) & (np.abs(Ys) <= 0.5) & (np.abs(Zs) <= 0.5)).transpose(2, 1, 0).reshape(-1, 1)
    w = np.where(mask_in, u, 0)
    assert np.allclose(sj[mask_in[:, 0]], 1.0)
    plain = orc.velmultiply(n, w, A1, fd)
    post = sum(orc.velmultiply(n, w, tuple(A1[k] if k == axis else 0 for k in range(3)), fd) * sj[:, [axis]]
               for axis in range(3))
    assert relerr(orc.velmultiply(n, w, A1, fd, invsqrtjac=sj), post) < 1e-14
    assert relerr(post, plain) > 1e-3                                     # the post-weight matters
    m = orc.mask_function(n, 2)
    Mv = np.einsum("i,j,k->kji", m, m, m).reshape(-1, 1)
    assert relerr(orc.velmultiply(n, u, A1, fd, mask=m), Mv * orc.velmultiply(n, u / Mv, A1, fd)) < 1e-13
    ke, _, _ = orc.sinedvr(n, d)
    zero_rp = np.zeros((n ** 3, 1, 1), dtype=complex, order="F")
    base = orc.actreduced00(n, ke, u, np.eye(1), zero_rp)
    withA = orc.actreduced00(n, ke, u, np.eye(1), zero_rp, tdpot=A1, velflag=1, fd=fd)
    assert relerr(withA - base, orc.velmultiply(n, u, A1, fd)) < 1e-13
