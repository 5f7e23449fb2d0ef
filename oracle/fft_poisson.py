"""ORACLE -- test infrastructure, not product code.

Closed-form pin for the uniform periodic operator: the 7-point Laplacian of mg_poisson.F90:20-24
is diagonal in the discrete Fourier basis with symbol

    lambda(k) = sum_d (2 cos(2 pi k_d / n) - 2) / h^2,

so the exact discrete periodic solution of L phi = f (f with zero mean) is phi_k = f_k / lambda(k),
with the k = 0 mode (the free additive constant) set to zero.  This is independent of any
multigrid detail: a converged mg_cycle must reproduce it, up to that constant.
"""
import numpy as np


def symbol(n, h):
    k = np.arange(n)
    s = (2.0 * np.cos(2.0 * np.pi * k / n) - 2.0) / (h * h)
    return s[:, None, None] + s[None, :, None] + s[None, None, :]


def solve(f, h):
    """Exact discrete periodic solve of L phi = f; returns mean-zero phi (axes k, j, i)."""
    n = f.shape[0]
    fk = np.fft.fftn(f)
    lam = symbol(n, h)
    lam[0, 0, 0] = 1.0
    phik = fk / lam
    phik[0, 0, 0] = 0.0
    return np.real(np.fft.ifftn(phik))


def apply_laplacian(u, h):
    """7-point periodic Laplacian (sum of 6 neighbours - 6u) / h^2, written with np.roll."""
    s = -6.0 * u
    for ax in range(3):
        s = s + np.roll(u, 1, axis=ax) + np.roll(u, -1, axis=ax)
    return s / (h * h)


def gradient(u, h):
    """Centred periodic gradient times -1 (fmg_poisson.F90:29-31), components (x, y, z)."""
    # arrays are indexed [k, j, i]: x is axis 2
    gx = -(np.roll(u, -1, axis=2) - np.roll(u, 1, axis=2)) / (2 * h)
    gy = -(np.roll(u, -1, axis=1) - np.roll(u, 1, axis=1)) / (2 * h)
    gz = -(np.roll(u, -1, axis=0) - np.roll(u, 1, axis=0)) / (2 * h)
    return np.stack([gx, gy, gz])


def eigenfunction(n, kx, ky, kz):
    """f = sin(2 pi kx x) sin(2 pi ky y) sin(2 pi kz z) at cell centres; L f = lambda f exactly."""
    x = (np.arange(n) + 0.5) / n
    sx = np.sin(2 * np.pi * kx * x)
    sy = np.sin(2 * np.pi * ky * x)
    sz = np.sin(2 * np.pi * kz * x)
    return sz[:, None, None] * sy[None, :, None] * sx[None, None, :]


def eigenvalue(n, h, kx, ky, kz):
    return sum((2.0 * np.cos(2.0 * np.pi * k / n) - 2.0) / (h * h) for k in (kx, ky, kz))
