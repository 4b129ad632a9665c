"""Synthetic inputs for the sinc-DVR operator path: seeded Gaussian orbitals on
the sinc-DVR grid, orthonormalised with the build flavour's dot.

Grid coordinates follow SD/SineDVR.f90:16-20 (points(i) = i*spacing, shifted
by (points(1)+points(n))/2).  Orthonormalisation is the reference's two-pass
classical Gram-Schmidt (S/utils.f90:275-320) with the flavour dot: hermnorm
conjugates the bra (S/Definitions.ALL:122), cnorm conjugates nothing and takes
a complex square-root norm (S/Definitions.ALL:114, CNORMTYPE complex*16).

Random numbers come from SplitMix64 (Steele, Lea & Flood 2014) fed to
Box-Muller, so the same seed gives the same orbitals on every host.
"""
import math

import numpy as np

MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & MASK

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / 9007199254740992.0))

    def normal(self):
        u1 = self.uniform()
        u2 = self.uniform()
        u1 = max(u1, 1e-300)
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)

    def randint(self, lo, hi):
        """Uniform integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1)) % (hi - lo + 1)


def grid_points(n, spacing):
    pts = np.arange(1, n + 1, dtype=np.float64) * spacing
    return pts - (pts[0] + pts[-1]) / 2.0


def flavour_dot(a, b, cnorm):
    return np.sum(a * b) if cnorm else np.sum(np.conj(a) * b)


def gram_schmidt(phis, cnorm=False):
    """Two-pass classical Gram-Schmidt of the columns, S/utils.f90:275-320."""
    out = np.array(phis, dtype=np.complex128, order="F", copy=True)
    for k in range(out.shape[1]):
        v = out[:, k].copy()
        for _ in range(2):
            for i in range(k):
                v = v - out[:, i] * flavour_dot(out[:, i], v, cnorm)
            nrm = np.sqrt(flavour_dot(v, v, cnorm))
            v = v / nrm
        out[:, k] = v
    return out


def gaussian_orbitals(n, spacing, norb, ngauss, seed, cnorm=False):
    """(n^3, norb) complex128 Fortran-ordered orbitals: each a sum of `ngauss`
    Gaussians, exponents U[0.5,2], centres U[-1,1]^3 bohr, coefficients
    N(0,1)+iN(0,1), then orthonormalised on the grid (no Delta^3 weight)."""
    rng = SplitMix64(seed)
    x = grid_points(n, spacing)
    X = x[:, None, None]
    Y = x[None, :, None]
    Z = x[None, None, :]
    phis = np.zeros((n ** 3, norb), dtype=np.complex128, order="F")
    for o in range(norb):
        f = np.zeros((n, n, n), dtype=np.complex128)
        for _ in range(ngauss):
            alpha = rng.uniform(0.5, 2.0)
            cx, cy, cz = rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-1, 1)
            c = complex(rng.normal(), rng.normal())
            f += c * np.exp(-alpha * ((X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2))
        # (x, y, z) with x fastest in memory
        phis[:, o] = f.transpose(2, 1, 0).ravel(order="C")
    return gram_schmidt(phis, cnorm)


def perturb_orbitals(phis, seed, norm=1e-3, cnorm=False):
    """Add seeded complex noise of the given 2-norm, then re-orthonormalise."""
    rng = SplitMix64(seed)
    noise = np.array([complex(rng.normal(), rng.normal()) for _ in range(phis.size)], dtype=np.complex128)
    noise = noise.reshape(phis.shape, order="F")
    noise *= norm / np.linalg.norm(noise)
    return gram_schmidt(phis + noise, cnorm)


# Workload configurations (grid edge n, spacing, orbitals, Gaussians each).
CONFIGS = {
    "grid21_4orb": dict(n=21, spacing=0.4, norb=4, ngauss=3, seed=0),
    "grid41_10orb": dict(n=41, spacing=0.3, norb=10, ngauss=4, seed=1000),
    "grid81_16orb": dict(n=81, spacing=0.25, norb=16, ngauss=6, seed=3000),
}

# Smooth exterior scaling of the 41^3 complex-scaled case: the ramp covers the
# outer 8 points of each axis (|x| = 3.9 ... 6.0) at angle 0.3 rad.
ECS_41 = dict(theta=0.3, stretch=1.0, dist=3.75, smooth=2.25)


# Batched workload: 64 independent orbital sets, grid edges from this list at
# spacing 0.3 bohr, 4..14 orbitals of 3 Gaussians each, 1..3 nuclei of charge
# 1..3 within +-1.5 bohr.
BATCH_EDGES = (25, 31, 41, 51, 61)


def nucleus_shift(pos, n, spacing):
    """Half-spacing offset of a nucleus nearest `pos` bohr on one axis, with the
    parity SD/sincproject.f90:199-204 requires (centershift = n-1 mod 2); the
    nucleus then sits at -centershift*spacing/2."""
    cs = int(round(-2.0 * pos / spacing))
    if (cs - (n - 1)) % 2:
        cs += 1 if -2.0 * pos / spacing > cs else -1
    return cs


def batch_sets(nsets=64, seed=2000, spacing=0.3, edges=BATCH_EDGES, norb_range=(4, 14), ngauss=3):
    """Descriptions of the orbital sets of the batched workload (orbitals are
    made by batch_orbitals, so a caller can size things first)."""
    rng = SplitMix64(seed)
    out = []
    for i in range(nsets):
        n = edges[rng.randint(0, len(edges) - 1)]
        norb = rng.randint(*norb_range)
        ncen = rng.randint(1, 3)
        cs, z = [], []
        for _ in range(ncen):
            cs.append([nucleus_shift(rng.uniform(-1.5, 1.5), n, spacing) for _ in range(3)])
            z.append(float(rng.randint(1, 3)))
        out.append(dict(n=n, spacing=spacing, norb=norb, ngauss=ngauss, seed=seed * 1000 + i,
                        centershift=cs, charges=z))
    return out


def batch_orbitals(desc, cnorm=False):
    return gaussian_orbitals(desc["n"], desc["spacing"], desc["norb"], desc["ngauss"], desc["seed"], cnorm)


def gaussian_orbitals_device(n, spacing, norb, ngauss, seed, device="cuda"):
    """gaussian_orbitals' recipe (same SplitMix64 draws) evaluated with torch on
    the device, orthonormalised there (two-pass classical Gram-Schmidt, hermnorm
    dot); returns (norb, n^3) complex128.  Agrees with gaussian_orbitals to
    rounding; for load generation where the host recipe is too slow."""
    import torch
    rng = SplitMix64(seed)
    x = torch.as_tensor(grid_points(n, spacing), dtype=torch.float64, device=device)
    Z, Y, X = torch.meshgrid(x, x, x, indexing="ij")   # row-major (z, y, x): x fastest
    out = torch.zeros((norb, n ** 3), dtype=torch.complex128, device=device)
    for o in range(norb):
        f = torch.zeros((n, n, n), dtype=torch.complex128, device=device)
        for _ in range(ngauss):
            alpha = rng.uniform(0.5, 2.0)
            cx, cy, cz = rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-1, 1)
            c = complex(rng.normal(), rng.normal())
            f += c * torch.exp(-alpha * ((X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2))
        out[o] = f.reshape(-1)
    for k in range(norb):
        v = out[k].clone()
        for _ in range(2):
            if k:
                v = v - torch.mv(out[:k].transpose(0, 1), torch.mv(out[:k].conj(), v))
            v = v / torch.sqrt(torch.vdot(v, v).real)
        out[k] = v
    return out
