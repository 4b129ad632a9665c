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
