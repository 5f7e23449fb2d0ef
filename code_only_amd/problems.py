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


def amr_source(hier, seed=0, nclump=4, sigma=0.06):
    """Clumpy source on every block of an AMR hierarchy ([nblk, 8, 8, 8]): seeded Gaussian clumps
    (centres uniform in the middle half of the domain, LogNormal(0, 1) amplitudes, width ``sigma``
    in domain units) on a unit background.  For periodic hierarchies the solver removes the
    leaf-volume mean (ZERO_AVERAGE_DENSITY)."""
    from .amr_hierarchy import to_blocks
    rng = np.random.default_rng(seed)
    ext = np.asarray(hier.base_dims) * 8 * hier.h0
    lo = np.asarray(hier.x0)
    centres = lo + ext * rng.uniform(0.25, 0.75, size=(nclump, 3))
    amps = rng.lognormal(0.0, 1.0, size=nclump)

    def f(x, y, z):
        out = np.ones(np.broadcast(x, y, z).shape)
        for c in range(nclump):
            r2 = (x - centres[c, 0]) ** 2 + (y - centres[c, 1]) ** 2 + (z - centres[c, 2]) ** 2
            out = out + amps[c] * np.exp(-r2 / (2 * sigma * sigma))
        return out
    return to_blocks(hier, f)


def sis_source(hier, seed=0, perturb=0.05):
    """Singular isothermal sphere for a domain centred on the origin: rho = 1 / (r^2 + r_c^2) with
    r_c one cell of the finest level (a one-cell core), times seeded lognormal perturbations of
    ``perturb`` standard deviation, one independent factor per cell of every block.  With
    G = 1/(4 pi) the source 4 pi G rho equals rho."""
    from .amr_hierarchy import to_blocks
    rc = hier.h(hier.nlev - 1)
    rng = np.random.default_rng(seed)
    rho = to_blocks(hier, lambda x, y, z: 1.0 / (x * x + y * y + z * z + rc * rc))
    s = np.sqrt(np.log1p(perturb * perturb))
    return rho * rng.lognormal(-0.5 * s * s, s, size=rho.shape)


BATCH_SIZES = (32, 64, 96, 128, 192)


def clumpy_rhs_torch(n, seed=0, nclump=8, sigma_cells=4.0, device="cuda"):
    """clumpy_rhs evaluated with torch on ``device`` (same centres and amplitudes; exp and sums
    rounded by the device, so values agree with clumpy_rhs to a few ulp, not bitwise)."""
    import torch
    rng = np.random.default_rng(seed)
    centres = rng.uniform(size=(nclump, 3))
    amps = rng.lognormal(mean=0.0, sigma=1.0, size=nclump)
    h = 1.0 / n
    sig = sigma_cells * h
    x = (torch.arange(n, dtype=torch.float64, device=device) + 0.5) * h
    rho = torch.full((n, n, n), 1.0, dtype=torch.float64, device=device)
    for c in range(nclump):
        g = []
        for d in range(3):
            dd = x - float(centres[c, d])
            dd = dd - torch.round(dd)
            g.append(torch.exp(-dd * dd / (2.0 * sig * sig)))
        gx, gy, gz = g
        rho += float(amps[c]) * (gz[:, None, None] * gy[None, :, None] * gx[None, None, :])
    return rho - rho.mean()


BATCH_SIZES = (32, 64, 96, 128, 192)


def batch_spec(nprob=48, seed=0, sizes=BATCH_SIZES):
    """The uneven batch: problem q has n_q drawn uniformly from ``sizes``, 1..16 clumps (uniform)
    and its own seed.  Returns (list of n, list of clump counts, list of seeds)."""
    rng = np.random.default_rng(seed)
    ns = [int(x) for x in rng.choice(np.asarray(sizes), size=nprob)]
    ncl = [int(x) for x in rng.integers(1, 17, size=nprob)]
    seeds = [int(x) for x in rng.integers(0, 2 ** 31 - 1, size=nprob)]
    return ns, ncl, seeds


def batch_problems(nprob=48, seed=0, sizes=BATCH_SIZES, sigma_cells=4.0, device=None):
    """Right-hand sides of the uneven batch (numpy, or torch tensors on ``device``).
    Returns (list of n, list of rhs)."""
    ns, ncl, seeds = batch_spec(nprob, seed, sizes)
    if device is None:
        rhs = [clumpy_rhs(n, seed=s, nclump=k, sigma_cells=sigma_cells) for n, s, k in zip(ns, seeds, ncl)]
    else:
        rhs = [clumpy_rhs_torch(n, seed=s, nclump=k, sigma_cells=sigma_cells, device=device)
               for n, s, k in zip(ns, seeds, ncl)]
    return ns, rhs


# Hydro block memory layouts (config.h:61-110): the self-gravitating adiabatic HD model without
# chemistry (MODEL_ART == 0) and the one with the chemistry network (MODEL_ART == 1, METAL), both
# with N_GHOST_CELL = 2 (config.h:135).  The second pads each slot to show the stride is honoured.
BLOCKMEM_LAYOUTS = {
    "hd": dict(ng=2, ncomp=9, stride=9 * 12 ** 3, mrho=0, mvx=1, mvy=2, mvz=3, mp=4, mpsi=5, mgx=6, mgy=7, mgz=8),
    "art": dict(ng=2, ncomp=21, stride=21 * 12 ** 3 + 96, mrho=0, mvx=1, mvy=2, mvz=3, mp=4, mpsi=16, mgx=17,
                mgy=18, mgz=19),
}


def blockmem_pool(hier, layout, seed=0, spare=5, src=None):
    """A pool of hydro block arrays for every block of ``hier``: three slots per block (its own
    state, the predictor stage, the past state) in a seeded random order among 3 nblk + spare
    slots.  Returns (pool [nslots, stride] float64, slot_u, slot_u1, slot_u2).  Density is
    positive everywhere (``src`` on the interiors of the own state when given, else a clumpy
    field; seeded perturbations of it in the other two), velocities in [-1, 1], pressure in
    [0.5, 1.5], the potential, g and any other component random normal."""
    rng = np.random.default_rng(seed)
    nb = hier.nblk
    W = 8 + 2 * layout["ng"]
    nslots = 3 * nb + spare
    pool = rng.standard_normal((nslots, layout["stride"]))
    perm = rng.permutation(nslots).astype(np.int32)
    slot_u, slot_u1, slot_u2 = perm[:nb], perm[nb:2 * nb], perm[2 * nb:3 * nb]
    comp = pool[:, :layout["ncomp"] * W ** 3].reshape(nslots, layout["ncomp"], W, W, W)
    comp[:, layout["mrho"]] = rng.uniform(0.5, 2.0, (nslots, W, W, W))
    for m in ("mvx", "mvy", "mvz"):
        comp[:, layout[m]] = rng.uniform(-1.0, 1.0, (nslots, W, W, W))
    comp[:, layout["mp"]] = rng.uniform(0.5, 1.5, (nslots, W, W, W))
    rho = amr_source(hier, seed=seed + 1) if src is None else np.asarray(src, np.float64)
    g0 = layout["ng"]
    inner = (slice(g0, g0 + 8),) * 3
    comp[(slot_u, layout["mrho"]) + inner] = rho
    comp[(slot_u1, layout["mrho"]) + inner] = rho * rng.uniform(0.9, 1.1, rho.shape)
    comp[(slot_u2, layout["mrho"]) + inner] = rho * rng.uniform(0.9, 1.1, rho.shape)
    return pool, slot_u, slot_u1, slot_u2
