"""Adaptive Krylov propagator (the reference's DGPHIV) on the CPU: against the
closed form w = exp(tA)v + t phi(tA)u, read off the exponential of the
augmented matrix [[A, u], [0, 0]] (an independent route), including a happy
breakdown, step rejection and negative time."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "v1-1_amd"))


def _closed(A, u, v, t):
    from scipy.linalg import expm
    n = A.shape[0]
    aug = np.zeros((n + 1, n + 1), dtype=complex)
    aug[:n, :n] = A
    aug[:n, n] = u
    E = expm(t * aug)
    return E[:n, :n] @ v + E[:n, n]


@pytest.mark.parametrize("t,guess", [(0.7, 0.1), (-0.4, 0.05), (2.0, 2.0)])
def test_phiv_matches_closed_form(t, guess):
    import torch

    from sincdvr import krylov
    rng = np.random.default_rng(3)
    n = 40
    A = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n) - 0.3j * np.eye(n)
    u = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    At = torch.from_numpy(A)
    w, st = krylov.phiv(lambda x: At @ x, t, torch.from_numpy(u), torch.from_numpy(v), tol=1e-10, m=10,
                        guess_step=guess)
    ref = _closed(A, u, v, t)
    assert np.abs(w.numpy() - ref).max() / np.abs(ref).max() < 1e-8
    assert abs(st["t_now"] - t) < 1e-14
    assert st["nstep"] >= 1 and st["nmult"] >= st["nstep"] * 11


def test_phiv_rejects_a_too_large_guess():
    import torch

    from sincdvr import krylov
    rng = np.random.default_rng(5)
    n = 30
    A = 3.0 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    u = np.zeros(n, dtype=complex)
    v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    At = torch.from_numpy(A)
    w, st = krylov.phiv(lambda x: At @ x, 1.0, torch.from_numpy(u), torch.from_numpy(v), tol=1e-9, m=6,
                        guess_step=1.0)
    assert st["nreject"] > 0
    ref = _closed(A, u, v, 1.0)
    assert np.abs(w.numpy() - ref).max() / np.abs(ref).max() < 1e-6


def test_phiv_happy_breakdown():
    """A with two distinct eigenvalues: the real Krylov space of A w + u closes
    after a few vectors, the step finishes at once and is exact."""
    import torch

    from sincdvr import krylov
    rng = np.random.default_rng(9)
    n = 20
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    d = np.where(np.arange(n) < n // 2, -0.5, 0.25)
    A = (Q * d) @ Q.conj().T
    u = np.zeros(n, dtype=complex)
    v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    At = torch.from_numpy(A)
    w, st = krylov.phiv(lambda x: At @ x, 1.5, torch.from_numpy(u), torch.from_numpy(v), tol=1e-10, m=12,
                        guess_step=0.01)
    assert st["nstep"] == 1
    ref = _closed(A, u, v, 1.5)
    assert np.abs(w.numpy() - ref).max() / np.abs(ref).max() < 1e-9
