"""Block tables of a block-structured AMR hierarchy, in the form the AMR entry points take.

This is the host-side counterpart of the reference's link-list construction
(``fmg_make_linklist``, fmg_data_if.F90:298-413): for every block it records its level, its
position at that level (block units), its parent and eight children, and for each of its six
faces the neighbouring block -- at the same level when one exists, otherwise the parent's
neighbour at the parent level (``NeighborParentLevel``), otherwise none (a physical boundary
face, ``TouchBoundary``).

Blocks are numbered level by level, coarsest first, and within a level in (k, j, i) order of
their positions.  Faces are numbered ``2*d + lr`` (d = x, y, z; lr = 0 left, 1 right); octants
``lri + 2*lrj + 4*lrk`` (grid.F90:9-12 Left = 0, Right = 1).

A hierarchy must be properly nested: each refined block's children may meet coarser blocks only
at the children's own parent level, so every face of every block has a neighbour at its level,
at its parent level, or on the domain boundary.  ``build`` rejects anything else.
"""
from dataclasses import dataclass, field

import numpy as np

PERIODIC = 0
SPHERE = 1
BLOCK = 8   # cells per block side (NI = NJ = NK = 8 in the reference's configurations)


@dataclass
class Hierarchy:
    level: np.ndarray          # int32 [nblk]
    ijk: np.ndarray            # int32 [nblk, 3] position at its level, block units
    parent: np.ndarray         # int32 [nblk], -1 at level 0
    child: np.ndarray          # int32 [nblk, 8], -1 for leaves
    nbr: np.ndarray            # int32 [nblk, 6], -1 on physical faces
    base_dims: tuple           # base blocks per axis
    h0: float                  # cell width at level 0
    x0: tuple                  # domain lower corner
    bc: int                    # PERIODIC or SPHERE
    nlev: int = 0
    lev_start: np.ndarray = field(default=None)

    @property
    def nblk(self):
        return int(self.level.shape[0])

    @property
    def leaf(self):
        return self.child[:, 0] < 0

    def cells(self):
        """Number of leaf cells."""
        return int(self.leaf.sum()) * BLOCK ** 3

    def h(self, lev):
        return self.h0 / 2 ** lev

    def cell_centres(self, b):
        """Cell-centre coordinates (x, y, z), each of shape (8,), of block b."""
        lev = int(self.level[b])
        h = self.h(lev)
        out = []
        for a in range(3):
            g = int(self.ijk[b, a]) * BLOCK + np.arange(BLOCK)
            out.append(self.x0[a] + (g + 0.5) * h)
        return out


def build(base_dims, refine, bc=PERIODIC, length=1.0, centred=None):
    """Build a hierarchy.

    base_dims -- base blocks per axis (nx, ny, nz).
    refine    -- list over levels 0..L-2 of sets of block positions (i, j, k) at that level to
                 refine into 8 children each; every position must exist at its level.
    bc        -- PERIODIC (neighbours wrap at every level) or SPHERE (domain faces are physical).
    length    -- domain side along x (cells are cubes: h0 = length / (nx * 8)).
    centred   -- domain centred on the origin (default: for SPHERE).
    """
    nx, ny, nz = (int(v) for v in base_dims)
    if centred is None:
        centred = bc == SPHERE
    h0 = float(length) / (nx * BLOCK)
    x0 = tuple((-0.5 * d * BLOCK * h0) if centred else 0.0 for d in (nx, ny, nz))
    levels = [sorted({(i, j, k) for k in range(nz) for j in range(ny) for i in range(nx)},
                     key=lambda p: (p[2], p[1], p[0]))]
    for lev, marks in enumerate(refine):
        have = set(levels[lev])
        kids = set()
        for p in marks:
            p = tuple(int(v) for v in p)
            if p not in have:
                raise ValueError(f"refined block {p} does not exist at level {lev}")
            for o in range(8):
                kids.add((2 * p[0] + (o & 1), 2 * p[1] + ((o >> 1) & 1), 2 * p[2] + ((o >> 2) & 1)))
        if not kids:
            break
        levels.append(sorted(kids, key=lambda p: (p[2], p[1], p[0])))
    nlev = len(levels)
    index = {}
    level, ijk = [], []
    for lev, blocks in enumerate(levels):
        for p in blocks:
            index[(lev,) + p] = len(level)
            level.append(lev)
            ijk.append(p)
    nblk = len(level)
    level = np.asarray(level, dtype=np.int32)
    ijk = np.asarray(ijk, dtype=np.int32).reshape(nblk, 3)
    parent = np.full(nblk, -1, dtype=np.int32)
    child = np.full((nblk, 8), -1, dtype=np.int32)
    for b in range(nblk):
        lev = int(level[b])
        if lev > 0:
            p = index[(lev - 1, int(ijk[b, 0]) // 2, int(ijk[b, 1]) // 2, int(ijk[b, 2]) // 2)]
            parent[b] = p
            o = (int(ijk[b, 0]) & 1) + 2 * (int(ijk[b, 1]) & 1) + 4 * (int(ijk[b, 2]) & 1)
            child[p, o] = b
    dims = np.array([nx, ny, nz])
    nbr = np.full((nblk, 6), -1, dtype=np.int32)
    for b in range(nblk):
        lev = int(level[b])
        ext = dims * 2 ** lev
        for d in range(3):
            for lr in range(2):
                q = ijk[b].copy()
                q[d] += 1 if lr else -1
                outside = q[d] < 0 or q[d] >= ext[d]
                if outside and bc != PERIODIC:
                    continue          # physical boundary face
                q[d] %= ext[d]
                m = index.get((lev, int(q[0]), int(q[1]), int(q[2])))
                if m is None:
                    if lev == 0:
                        raise ValueError("base level has a hole")
                    pm = nbr[parent[b], 2 * d + lr]
                    if pm < 0 or level[pm] != lev - 1:
                        raise ValueError(f"block {b} (level {lev}, {tuple(ijk[b])}) is not properly nested")
                    m = pm
                nbr[b, 2 * d + lr] = m
    lev_start = np.searchsorted(level, np.arange(nlev + 1)).astype(np.int32)
    return Hierarchy(level=level, ijk=ijk, parent=parent, child=child, nbr=nbr,
                     base_dims=(nx, ny, nz), h0=h0, x0=x0, bc=bc, nlev=nlev, lev_start=lev_start)


def nested(base=16, levels=5, bc=SPHERE, length=1.0):
    """Base of base^3 blocks, then levels-1 refined levels, each refining the central half-width
    (per axis) of the previous level's region: every level holds base^3 blocks."""
    if base % 4:
        raise ValueError("base must be a multiple of 4")
    refine = []
    lo, hi = 0, base                # region of the current level, in its block units
    for _ in range(levels - 1):
        w = hi - lo
        a, b = lo + w // 4, hi - w // 4
        refine.append({(i, j, k) for k in range(a, b) for j in range(a, b) for i in range(a, b)})
        lo, hi = 2 * a, 2 * b
    return build((base, base, base), refine, bc=bc, length=length)


def random_hierarchy(base, levels, frac, seed, bc=PERIODIC, length=1.0):
    """A properly nested hierarchy with a seeded random choice of refined blocks per level."""
    rng = np.random.default_rng(seed)
    refine = []
    cur = {(i, j, k) for k in range(base) for j in range(base) for i in range(base)}
    ext = base
    for _ in range(levels - 1):
        cand = sorted(cur, key=lambda p: (p[2], p[1], p[0]))
        pick = [p for p in cand if rng.random() < frac]
        # keep only picks whose 26 neighbours (with wrap when periodic) exist at this level, so
        # that the children are properly nested
        ok = set()
        for p in pick:
            good = True
            for dz in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        q = [p[0] + dx, p[1] + dy, p[2] + dz]
                        if bc == PERIODIC:
                            q = [c % ext for c in q]
                        elif min(q) < 0 or max(q) >= ext:
                            continue
                        if tuple(q) not in cur:
                            good = False
            if good:
                ok.add(p)
        if not ok:
            break
        refine.append(ok)
        cur = {(2 * p[0] + (o & 1), 2 * p[1] + ((o >> 1) & 1), 2 * p[2] + ((o >> 2) & 1))
               for p in ok for o in range(8)}
        ext *= 2
    return build((base, base, base), refine, bc=bc, length=length)


def to_blocks(h, func):
    """Evaluate func(x, y, z) (broadcasting) at every cell centre: array [nblk, 8, 8, 8] in
    (k, j, i) order, i fastest in memory."""
    out = np.empty((h.nblk, BLOCK, BLOCK, BLOCK))
    for b in range(h.nblk):
        x, y, z = h.cell_centres(b)
        out[b] = func(x[None, None, :], y[None, :, None], z[:, None, None])
    return out
