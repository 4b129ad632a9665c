"""Krylov propagation of the orbitals with the operator on the device.

The reference propagates the orbitals with expokit (S/expo_driver.f90,
S/expokit.f90): an Arnoldi basis of the orbital operator (jacoperate ->
actreduced00, S/derivs.f90:629-852) and the exponential of the small
Hessenberg matrix.  Here the Krylov vectors stay on the device and every
matvec is the library's actreduced00; the orthogonalisation coefficients are
formed and kept in FP64 on the host (the inner products come back as FP64
numbers, the Hessenberg matrix and its exponential live on the host), and the
vector updates run on the device.  Hermitian inner product (hermnorm).
"""
import math

import numpy as np
import torch


def _dot(a, b):
    """<a|b> of two device vectors, returned to the host as a Python complex (FP64)."""
    return complex(torch.vdot(a.reshape(-1), b.reshape(-1)).item())


def arnoldi(apply, v0, order):
    """Arnoldi basis of `apply` from v0 with two-pass classical Gram-Schmidt
    (coefficients on the host in FP64).  Returns (V list of device tensors,
    H (order+1, order) host complex128, beta = |v0|)."""
    beta = float(np.sqrt(_dot(v0, v0).real))
    V = [v0 / beta]
    H = np.zeros((order + 1, order), dtype=np.complex128)
    for k in range(order):
        w = apply(V[k])
        for _ in range(2):
            c = np.array([_dot(V[j], w) for j in range(k + 1)], dtype=np.complex128)
            H[: k + 1, k] += c
            for j in range(k + 1):
                w = w - complex(c[j]) * V[j]
        h = float(np.sqrt(_dot(w, w).real))
        H[k + 1, k] = h
        if h < 1e-300:
            return V, H[: k + 2, : k + 1], beta
        V.append(w / h)
    return V, H, beta


def expmv(apply, v0, dt, order=12):
    """exp(dt * A) v0 from an order-`order` Arnoldi basis (A = apply)."""
    from scipy.linalg import expm
    V, H, beta = arnoldi(apply, v0, order)
    m = H.shape[1]
    E = expm(dt * H[:m, :m])
    coef = beta * E[:, 0]
    out = torch.zeros_like(v0)
    for j in range(m):
        out = out + complex(coef[j]) * V[j]
    return out, H


def _nint(x):
    """Fortran NINT: nearest integer, halves away from zero."""
    return int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))


def phiv(apply, t, u, v, tol, m, anorm=1.0, guess_step=None, break_tol=1e-7, mxstep=1000):
    """w = exp(t A) v + t phi(t A) u, phi(z) = (exp(z) - 1)/z: the solution at t
    of w' = A w + u, w(0) = v, by the adaptive Krylov method of the reference's
    DGPHIV (S/expokit.f90:3904-4337; Sidje, ACM TOMS 24 (1998) 130-156) with
    the reference's changes: the first step is the caller's guess (the
    driver's running step size, S/expo_driver.f90:595-602), a rejected step
    shrinks by gamma = 0.8 without limit on rejections, the suggested next
    step never falls below the one just taken, happy breakdown at
    h(j+1, j) <= 1e-7.

    The complex vectors are treated as real vectors of twice the length, as
    the reference does (REALGO off, S/expo_driver.f90:574-576): inner
    products are Re<a|b> (DDOT of the real views), the Hessenberg matrix is
    real, built on the host in FP64; the Krylov vectors stay where `apply`
    puts them (the device) and only their inner products come back.
    Returns (w, stats) with the next guess step in stats["guess_step"]."""
    from scipy.linalg import expm

    def rdot(a, b):
        return float(torch.vdot(a.reshape(-1), b.reshape(-1)).real.item())

    def rnorm(a):
        return float(torch.linalg.vector_norm(a.reshape(-1)).item())

    gamma, delta = 0.8, 1.2
    eps = float(np.finfo(np.float64).eps)
    if tol <= eps:
        tol = math.sqrt(eps)
    rndoff = eps * anorm
    sgn = 1.0 if t >= 0 else -1.0
    t_out = abs(t)
    w = v.clone()
    t_now = t_new = t_step = t_last = 0.0
    nstep = nmult = nreject = 0
    xm = 1.0 / m
    errs = []
    avnorm = 0.0
    while t_now < t_out:
        p = apply(w) + u                                   # Krylov start A w + u
        nmult += 1
        beta = rnorm(p)
        if beta == 0.0:
            break
        V = [p / beta]
        mh = m + 3
        H = np.zeros((mh, mh))
        if nstep == 0:
            t_new = t_out if guess_step is None else float(guess_step)
        nstep += 1
        t_last = t_step
        t_step = min(t_out - t_now, t_new)
        k1, mbrk, hj1j = 3, m, 0.0
        for j in range(m):                                 # Arnoldi, modified Gram-Schmidt
            wv = apply(V[j])
            nmult += 1
            for i in range(j + 1):
                hij = rdot(V[i], wv)
                wv = wv - hij * V[i]
                H[i, j] = hij
            hj1j = rnorm(wv)
            if hj1j <= break_tol:                          # happy breakdown: finish in one step
                k1, mbrk = 0, j + 1
                t_step = t_out - t_now
                break
            H[j + 1, j] = hj1j
            V.append(wv / hj1j)
        if k1 != 0:
            avnorm = rnorm(apply(V[m]))
            nmult += 1
        H[0, mbrk] = 1.0                                   # the 3-extended scheme
        H[m, m - 1] = 0.0
        for i in range(1, k1):
            H[m + i - 1, m + i] = 1.0
        while True:
            mx = mbrk + max(1, k1)
            E = expm(sgn * t_step * H[:mx, :mx])
            phih = E[:, mbrk].copy()
            if k1 == 0:
                err = tol
            else:
                phih[m] = hj1j * E[m - 1, m + 1]
                phih[m + 1] = hj1j * E[m - 1, m + 2]
                p1 = abs(phih[m]) * beta
                p2 = abs(phih[m + 1]) * beta * avnorm
                if p1 > 10.0 * p2:
                    err, xm = p2, 1.0 / (m + 1)
                else:
                    err, xm = p1, 1.0 / m
            if k1 != 0 and err > delta * t_step * tol:      # reject: shrink and recompute
                t_step *= gamma
                nreject += 1
                continue
            break
        for i in range(mbrk + max(0, k1 - 2)):
            w = w + (beta * float(phih[i])) * V[i]
        t_new = gamma * t_step * (t_step * tol / err) ** xm
        p1 = 10.0 ** (_nint(math.log10(t_new) - math.sqrt(0.1)) - 1)
        t_new = math.trunc(t_new / p1 + 0.55) * p1
        if t_new < t_step:
            t_new = t_step
        errs.append(max(err, rndoff))
        t_now += t_step
        if mxstep and nstep >= mxstep:
            break
    stats = {"nmult": nmult, "nstep": nstep, "nreject": nreject, "t_now": sgn * t_now,
             "x_error": max(errs) if errs else 0.0, "s_error": sum(errs),
             "guess_step": t_step if nstep == 1 else t_last}
    return w, stats
