"""The dump reader (code_only_amd/dump.py) against dumps written by the oracle's restatement of the
reference's writer (oracle/dump_writer.py: dumpdata_npe, io.F90:213-248): every field read back
as written, for 1-3 ranks, 4- and 8-byte record markers and both byte orders; the block tables
rebuilt from the grids' positions equal the hierarchy the dump was made from, and the pool laid
out from it holds the grids' arrays; malformed files are rejected.  The GPU test solves from a
dump's pool and compares with the solve from the pool the dump was written from."""
import os

import numpy as np
import pytest

from code_only_amd import amr_hierarchy as ah
from code_only_amd import problems
from code_only_amd.dump import DumpFormatError, read_dump
from oracle import dump_writer

LAY = problems.BLOCKMEM_LAYOUTS
MPI_PROC_NULL = -2


def _hier(which):
    return {"nested-sphere": lambda: ah.nested(base=4, levels=3, bc=ah.SPHERE),
            "random-periodic": lambda: ah.random_hierarchy(4, 3, 0.35, seed=7, bc=ah.PERIODIC)}[which]()


def make_dump(tmp, h, lay, pool, su, s1, s2, nranks, marker=4, endian="<", seed=0, step_max=6):
    """Write `h` with its pool as a dump of `nranks` rank files; returns (paths, per-block
    (rank, gid), flux arrays per block)."""
    rng = np.random.default_rng(seed)
    nb, ng, nm = h.nblk, lay["ng"], lay["ncomp"]
    W = 8 + 2 * ng
    rank = rng.integers(0, nranks, nb)
    rank[:nranks] = np.arange(nranks)            # every rank holds a grid
    gid = np.empty(nb, np.int64)
    extra = {}
    for r in range(nranks):
        idx = np.flatnonzero(rank == r)
        ids = rng.permutation(3 * len(idx))
        gid[idx] = ids[:len(idx)]
        extra[r] = ids[len(idx):]
    u1l, u2l = np.empty(nb, np.int64), np.empty(nb, np.int64)
    for r in range(nranks):
        idx = np.flatnonzero(rank == r)
        u1l[idx] = extra[r][:len(idx)]
        u2l[idx] = extra[r][len(idx):]
    nl = h.nlev
    params = dict(Imin=0, Jmin=0, Kmin=0, Mmin=0, Gidmin=0, Lmin=0, Imax=7, Jmax=7, Kmax=7, Mmax=nm - 1,
                  Gidmax=4999, Lmax=15, Ngh=ng, Imingh=-ng, Jmingh=-ng, Kmingh=-ng, Imaxgh=7 + ng,
                  Jmaxgh=7 + ng, Kmaxgh=7 + ng)
    step = np.array([8] * (nl - 1) + [step_max], np.int64)
    dstep = np.array([4] * (nl - 1) + [1], np.int64)
    per = nm * W ** 3
    fluxes = {}
    paths = []
    for r in range(nranks):
        idx = np.flatnonzero(rank == r)
        counts = [int(np.sum(h.level[idx] == l)) for l in range(nl)]
        gl = np.full((max(counts), nl), dump_writer.UNDEF, np.int32)
        for l in range(nl):
            g_l = gid[idx[h.level[idx] == l]]
            gl[:len(g_l), l] = rng.permutation(g_l)
        header = dict(ngrid_global=nb, level_max=nl - 1, gidlistmax=np.array(counts) - 1, gidlist=gl,
                      time=np.linspace(0.1, 0.2, nl), dtime=np.array([1e-3 / 2 ** l for l in range(nl)]),
                      step=step, dstep=dstep,
                      cellwidth=np.array([[h.h(l)] * 3 for l in range(nl)]).T)
        grids = []
        for b in sorted(idx, key=lambda b: gid[b]):          # dumpdata_npe walks gids upward

            def fortran(slot):
                return pool[slot, :per].reshape(nm, W, W, W).transpose(3, 2, 1, 0)
            coords = [h.x0[a] + (h.ijk[b, a] * 8 + np.arange(-ng, 8 + ng) + 0.5) * h.h(int(h.level[b]))
                      for a in range(3)]
            nbr_gid = np.full((2, 3), dump_writer.UNDEF)
            nbr_rank = np.full((2, 3), MPI_PROC_NULL)
            for d in range(3):
                for lr in range(2):
                    m = h.nbr[b, 2 * d + lr]
                    if m >= 0 and h.level[m] == h.level[b]:
                        nbr_gid[lr, d], nbr_rank[lr, d] = gid[m], rank[m]
            child_gid = np.full((2, 2, 2), dump_writer.UNDEF)
            child_rank = np.full((2, 2, 2), MPI_PROC_NULL)
            for o in range(8):
                c = h.child[b, o]
                if c >= 0:
                    child_gid[o & 1, (o >> 1) & 1, o >> 2] = gid[c]
                    child_rank[o & 1, (o >> 1) & 1, o >> 2] = rank[c]
            p = h.parent[b]
            flux = [rng.standard_normal(64 * nm) for _ in range(12)]
            fluxes[b] = flux
            grids.append(dict(gid=gid[b], level=int(h.level[b]), u1list=u1l[b], u2list=u2l[b],
                              alloc=(True, True, True), u=fortran(su[b]), u1=fortran(s1[b]), u2=fortran(s2[b]),
                              x=coords[0], y=coords[1], z=coords[2], ijk=tuple(int(v) for v in h.ijk[b]),
                              nbr_gid=nbr_gid, nbr_rank=nbr_rank, child_gid=child_gid, child_rank=child_rank,
                              parent_gid=gid[p] if p >= 0 else dump_writer.UNDEF,
                              parent_rank=rank[p] if p >= 0 else MPI_PROC_NULL, flux=flux))
        path = os.path.join(tmp, f"st{step_max}.{r}.d")
        dump_writer.write_rank(path, params, header, grids, marker, endian)
        paths.append(path)
    return paths, rank, gid, fluxes


CASES = [("nested-sphere", "hd", 1, 4, "<"), ("nested-sphere", "art", 2, 8, ">"),
         ("random-periodic", "hd", 3, 4, ">"), ("random-periodic", "art", 1, 8, "<")]


@pytest.mark.parametrize("which,name,nranks,marker,endian", CASES)
def test_read_back(tmp_path, which, name, nranks, marker, endian):
    h = _hier(which)
    lay = LAY[name]
    pool, su, s1, s2 = problems.blockmem_pool(h, lay, seed=5)
    paths, rank, gid, fluxes = make_dump(str(tmp_path), h, lay, pool, su, s1, s2, nranks, marker, endian, seed=9)
    d = read_dump(list(reversed(paths)))          # ranks come from the file names
    assert (d.marker, d.endian) == (marker, endian)
    assert d.ncomp == lay["ncomp"] and d.ng == lay["ng"] and d.nlev == h.nlev
    assert d.ngrid_global == h.nblk and len(d.grids) == h.nblk
    assert np.array_equal(d.step, [8] * (h.nlev - 1) + [6]) and d.step.dtype == np.int64
    hb, order = d.hierarchy(h.bc)
    for k in ("level", "ijk", "parent", "child", "nbr"):
        assert np.array_equal(getattr(hb, k), getattr(h, k)), k
    assert hb.h0 == h.h0 and np.allclose(hb.x0, h.x0, rtol=0, atol=1e-15) and hb.bc == h.bc
    per = lay["ncomp"] * (8 + 2 * lay["ng"]) ** 3
    p2, t0, t1, t2 = d.pool(order)
    for b in range(h.nblk):
        g = d.grids[order[b]]
        assert (g.rank, g.gid) == (rank[b], gid[b])
        assert np.array_equal(p2[t0[b]], pool[su[b], :per])
        assert np.array_equal(p2[t1[b]], pool[s1[b], :per])
        assert np.array_equal(p2[t2[b]], pool[s2[b], :per])
        faces = [g.flux[f + "_" + k].reshape(-1) for k in ("bnd", "sum") for f in ("xl", "xr", "yl", "yr", "zl", "zr")]
        assert all(np.array_equal(a, e) for a, e in zip(faces, fluxes[b]))
        assert g.alloc == (True, True, True)
    ts = d.timeslice(3)
    assert ts["step_max"] == 6.0 and ts["dstep"][0] == 4.0


def test_rejects_malformed(tmp_path):
    h = _hier("nested-sphere")
    lay = LAY["hd"]
    pool, su, s1, s2 = problems.blockmem_pool(h, lay, seed=5)
    paths, *_ = make_dump(str(tmp_path), h, lay, pool, su, s1, s2, 1)
    raw = open(paths[0], "rb").read()
    bad = tmp_path / "bad.0.d"
    bad.write_bytes(raw[:len(raw) - 100])                       # truncated
    with pytest.raises(DumpFormatError):
        read_dump(str(bad))
    v4 = bytearray(raw)
    v4[4:8] = np.array([4], "<i4").tobytes()                    # IO_VERSION 4
    bad.write_bytes(bytes(v4))
    with pytest.raises(DumpFormatError):
        read_dump(str(bad))
    # one grid's same-level neighbour link pointing elsewhere: the rebuilt tables disagree
    d = read_dump(paths)
    g = next(g for g in d.grids if (g.nbr_gid >= 0).any())
    d1, lr = np.argwhere(g.nbr_gid >= 0)[0]
    g.nbr_gid[d1, lr] = g.gid
    with pytest.raises(DumpFormatError):
        d.hierarchy(h.bc)


@pytest.mark.gpu
def test_solve_from_dump(tmp_path, gpu):
    import torch
    from code_only_amd.amr import AmrSolver
    h = _hier("nested-sphere")
    lay = LAY["hd"]
    pool, su, s1, s2 = problems.blockmem_pool(h, lay, seed=5)
    paths, *_ = make_dump(str(tmp_path), h, lay, pool, su, s1, s2, 2, seed=3)
    d = read_dump(paths)
    hb, order = d.hierarchy(h.bc)
    p2, t0, t1, t2 = d.pool(order)
    out = []
    for hh, pl, a, b, c in ((h, pool, su, s1, s2), (hb, p2, t0, t1, t2)):
        s = AmrSolver(hh)
        try:
            s.set_blockmem(dict(lay, stride=pl.shape[1]), pl.shape[0], a, b, c)
            dp = torch.from_numpy(pl).to(gpu)
            st = s.solve_blockmem(dp, 1.0, tol=1e-10, maxit=20, **d.timeslice(1))
            assert st["status"] == 0
            out.append(dp.cpu().numpy()[a])
        finally:
            s.close()
    per = lay["ncomp"] * (8 + 2 * lay["ng"]) ** 3
    assert np.array_equal(out[0][:, :per], out[1][:, :per])


@pytest.mark.gpu
def test_dump_command_line(tmp_path, gpu):
    """python -m code_only_amd.dump FILES: reads, solves on the GPU, writes psi and g per block."""
    from code_only_amd import dump
    h = _hier("random-periodic")
    lay = LAY["hd"]
    pool, su, s1, s2 = problems.blockmem_pool(h, lay, seed=8)
    paths, *_ = make_dump(str(tmp_path), h, lay, pool, su, s1, s2, 2, seed=4)
    out = str(tmp_path / "res.npz")
    assert dump.main(paths + ["--bc", "periodic", "--precision", "mixed", "--out", out]) == 0
    r = np.load(out)
    assert r["psi"].shape == (h.nblk, 8, 8, 8) and r["g"].shape == (h.nblk, 3, 8, 8, 8)
    assert np.isfinite(r["psi"]).all() and np.abs(r["g"]).max() > 0
