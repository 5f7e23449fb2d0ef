"""Seeded synthetic inputs for the self-gravity solve.

The reference reads rho from its AMR block store (fmg_data_if.F90:27-84: rho scaled by 4 pi G and
the mean subtracted when ZERO_AVERAGE_DENSITY is set).  These generators produce that right-hand
side directly for the benchmark workloads, with G = 1/(4 pi) so that rhs = rho - <rho>.

Arrays are C-ordered [k][j][i] (x fastest), the memory order of the reference's u(i,j,k).
"""
import numpy as np


def clumpy_density(n, seed=0, nclump=8, sigma_cells=4.0, background=1.0):
    """Periodic unit box, n^3 cells: rho = 1 + sum_c A_c exp(-|x - x_c|^2 / (2 sigma^2)).

    Centres x_c ~ U[0,1)^3, amplitudes A_c ~ LogNormal(0, 1), minimum-image distances,
    sigma = ``sigma_cells`` cells.  Returns rho (float64, [k][j][i])."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(size=(nclump, 3))
    amps = rng.lognormal(mean=0.0, sigma=1.0, size=nclump)
    h = 1.0 / n
    sig = sigma_cells * h
    x = (np.arange(n) + 0.5) * h
    rho = np.full((n, n, n), background, dtype=np.float64)
    for c in range(nclump):
        g = []
        for d in range(3):
            dd = x - centres[c, d]
            dd -= np.round(dd)
            g.append(np.exp(-dd * dd / (2.0 * sig * sig)))
        gx, gy, gz = g
        rho += amps[c] * (gz[:, None, None] * gy[None, :, None] * gx[None, None, :])
    return rho


def clumpy_rhs(n, seed=0, nclump=8, sigma_cells=4.0):
    """Right-hand side 4 pi G (rho - <rho>) with G = 1/(4 pi): zero mean, solvable periodic."""
    rho = clumpy_density(n, seed=seed, nclump=nclump, sigma_cells=sigma_cells)
    return rho - rho.mean()
