"""GPU parity of the device-resident block-memory path against the oracle (oracle/blockmem.c with
oracle/amr.c): staging into the solver with each of fmg_prepare_interp's time slices, the solve,
the write-back of psi and g (sg_amr_solve_blockmem_dev), and the gravity source step
(sg_amr_source_g_dev), on pools whose slots are shuffled and padded.

FP64: the whole pool after the call equals the oracle's, bit for bit (staging, FP64 solve and
write-back in the same operand order, source step likewise).  Mixed precision: the written psi and
g agree with the oracle's FP64 ones to the solve's tolerance and every other value of the pool is
untouched."""
import numpy as np
import pytest

import oracle
from code_only_amd import amr_hierarchy as ah
from code_only_amd import problems
from code_only_amd._lib import SgError
from code_only_amd.amr import SG_PREC_MIXED, AmrSolver

pytestmark = pytest.mark.gpu

LAY = problems.BLOCKMEM_LAYOUTS
PI4G = 1.0   # G = 1/(4 pi)


def _hier(which):
    return {"nested-sphere": lambda: ah.nested(base=4, levels=3, bc=ah.SPHERE),
            "random-periodic": lambda: ah.random_hierarchy(4, 3, 0.35, seed=7, bc=ah.PERIODIC)}[which]()


def _slice(h, source):
    """Time-slice parameters: levels below the finest one coarse step of 4 from Step 8, the finest
    at Step 6 (half way)."""
    nl = h.nlev
    return dict(source=source, step_max=6.0, step=[8.0] * (nl - 1) + [6.0], dstep=[4.0] * (nl - 1) + [1.0])


def _oracle_chain(h, pool, lay, su, s1, s2, ts, tol, maxit):
    o = oracle.Amr(h)
    try:
        L = oracle.blockmem(**lay)
        t = oracle.timeslice(ts["source"], ts["step_max"], ts["step"], ts["dstep"])
        psi, src = oracle.bm_stage(pool, L, su, h.level, t, PI4G, s1, s2)
        _, g, st = o.solve(src, psi=psi, tol=tol, maxit=maxit)
        out = pool.copy()
        oracle.bm_restore(out, L, su, o.get(0, oracle.U), g)
        return out, st
    finally:
        o.close()


@pytest.mark.parametrize("which", ["nested-sphere", "random-periodic"])
@pytest.mark.parametrize("name", sorted(LAY))
@pytest.mark.parametrize("source", [0, 1, 2, 3])
def test_solve_blockmem_fp64_bitwise(which, name, source, gpu):
    import torch
    h = _hier(which)
    lay = LAY[name]
    pool, su, s1, s2 = problems.blockmem_pool(h, lay, seed=31 + source)
    ts = _slice(h, source)
    exp, so = _oracle_chain(h, pool, lay, su, s1, s2, ts, 1e-10, 20)
    g = AmrSolver(h)
    try:
        g.set_blockmem(lay, pool.shape[0], su, s1, s2)
        dpool = torch.from_numpy(pool).to(gpu)
        sg = g.solve_blockmem(dpool, PI4G, tol=1e-10, maxit=20, **ts)
        got = dpool.cpu().numpy()
    finally:
        g.close()
    print(which, name, source, so["iterations"], sg["iterations"], sg["rel_linf"], flush=True)
    assert so["iterations"] == sg["iterations"] and so["rel_linf"] == sg["rel_linf"]
    bad = np.argwhere(got != exp)
    assert len(bad) == 0, f"{len(bad)} pool values differ, first {bad[:5].tolist()}"


@pytest.mark.parametrize("which", ["nested-sphere", "random-periodic"])
def test_solve_blockmem_mixed(which, gpu):
    import torch
    h = _hier(which)
    lay = LAY["art"]
    pool, su, s1, s2 = problems.blockmem_pool(h, lay, seed=41)
    ts = _slice(h, 1)
    exp, _ = _oracle_chain(h, pool, lay, su, s1, s2, ts, 1e-12, 30)
    g = AmrSolver(h)
    try:
        g.set_blockmem(lay, pool.shape[0], su, s1, s2)
        dpool = torch.from_numpy(pool).to(gpu)
        sg = g.solve_blockmem(dpool, PI4G, tol=1e-10, maxit=30, precision=SG_PREC_MIXED, **ts)
        got = dpool.cpu().numpy()
    finally:
        g.close()
    assert sg["status"] == 0 and sg["rel_linf"] < 1e-10
    W = 12
    W3 = W ** 3
    written = np.zeros_like(pool, bool)
    c = written[:, :lay["ncomp"] * W3].reshape(pool.shape[0], lay["ncomp"], W, W, W)
    c[(su, lay["mpsi"]) + (slice(1, 11),) * 3] = True
    for m in ("mgx", "mgy", "mgz"):
        c[(su, lay[m]) + (slice(2, 10),) * 3] = True
    assert np.array_equal(got[~written], pool[~written])
    gc = got[:, :lay["ncomp"] * W3].reshape(pool.shape[0], lay["ncomp"], W, W, W)
    ec = exp[:, :lay["ncomp"] * W3].reshape(pool.shape[0], lay["ncomp"], W, W, W)
    ins = (slice(2, 10),) * 3
    for m in ("mpsi", "mgx", "mgy", "mgz"):
        a, b = gc[(su, lay[m]) + ins], ec[(su, lay[m]) + ins]
        if m == "mpsi" and h.bc == ah.PERIODIC:
            a, b = a - a.mean(), b - b.mean()
        err = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(which, m, err, flush=True)
        assert err < 1e-9


@pytest.mark.parametrize("name", sorted(LAY))
def test_source_g_bitwise(name, gpu):
    import torch
    h = _hier("nested-sphere")
    lay = LAY[name]
    pool, su, s1, _ = problems.blockmem_pool(h, lay, seed=53)
    dv = [h.h(l) ** 3 for l in range(h.nlev)]
    dt = [3e-3 / 2 ** l for l in range(h.nlev)]
    exp = pool.copy()
    oracle.bm_source_g(exp, oracle.blockmem(**lay), su, s1, h.level, dv, dt, 1.4)
    g = AmrSolver(h)
    try:
        g.set_blockmem(lay, pool.shape[0], su, s1)
        dpool = torch.from_numpy(pool).to(gpu)
        g.source_g(dpool, dt, 1.4)
        got = dpool.cpu().numpy()
    finally:
        g.close()
    bad = np.argwhere(got != exp)
    assert len(bad) == 0, f"{len(bad)} pool values differ, first {bad[:5].tolist()}"


def test_blockmem_argument_checks(gpu):
    import torch
    h = _hier("nested-sphere")
    lay = dict(LAY["hd"])
    pool, su, s1, s2 = problems.blockmem_pool(h, lay, seed=3)
    g = AmrSolver(h)
    try:
        dpool = torch.from_numpy(pool).to(gpu)
        with pytest.raises(SgError):
            g.solve_blockmem(dpool, PI4G)          # no layout yet
        with pytest.raises(SgError):
            g.set_blockmem(dict(lay, mgx=lay["mrho"]), pool.shape[0], su)
        with pytest.raises(SgError):
            g.set_blockmem(dict(lay, stride=100), pool.shape[0], su)
        bad = su.copy()
        bad[3] = pool.shape[0]
        with pytest.raises(SgError):
            g.set_blockmem(lay, pool.shape[0], bad)
        dup = su.copy()
        dup[3] = dup[4]                                  # two grids in one slot
        with pytest.raises(SgError):
            g.set_blockmem(lay, pool.shape[0], dup)
        with pytest.raises(SgError):
            g.set_blockmem(lay, pool.shape[0], su, su)   # predictor arrays on the grids' own
        g.set_blockmem(lay, pool.shape[0], su)
        with pytest.raises(SgError):
            g.solve_blockmem(dpool, PI4G, source=1)   # no predictor slots
        with pytest.raises(SgError):
            g.source_g(dpool, [1e-3] * h.nlev)         # likewise
        assert np.array_equal(dpool.cpu().numpy(), pool)
    finally:
        g.close()


def test_blockmem_full_sphere_hierarchy(gpu):
    """The C2 hierarchy (20 480 grids): the FP64 solve staged from the pool writes back exactly what
    the dense entry point returns for the same staged input."""
    import torch
    h = ah.nested(base=16, levels=5, bc=ah.SPHERE)
    lay = LAY["hd"]
    # one slot per grid (the current state only: time slice 0 reads nothing else), shuffled
    rng = np.random.default_rng(71)
    pool = rng.standard_normal((h.nblk, lay["stride"]))
    su = rng.permutation(h.nblk).astype(np.int32)
    c0 = pool.reshape(h.nblk, 9, 12, 12, 12)
    c0[(su, lay["mrho"]) + (slice(2, 10),) * 3] = problems.sis_source(h, seed=1)
    L = oracle.blockmem(**lay)
    psi0, src0 = oracle.bm_stage(pool, L, su, h.level, oracle.timeslice(0), PI4G)
    g = AmrSolver(h)
    try:
        pd, gd, sd = g.solve_host(src0, psi=psi0, tol=1e-10, maxit=30)
        g.set_blockmem(lay, pool.shape[0], su)
        dpool = torch.from_numpy(pool).to(gpu)
        sb = g.solve_blockmem(dpool, PI4G, tol=1e-10, maxit=30)
        got = dpool.cpu().numpy()
    finally:
        g.close()
    assert sb["iterations"] == sd["iterations"] and sb["rel_linf"] == sd["rel_linf"]
    c = got[:, :9 * 12 ** 3].reshape(-1, 9, 12, 12, 12)
    ins = (slice(2, 10),) * 3
    assert np.array_equal(c[(su, lay["mpsi"]) + ins], pd)
    for a, m in enumerate(("mgx", "mgy", "mgz")):
        assert np.array_equal(c[(su, lay[m]) + ins], gd[:, a])
