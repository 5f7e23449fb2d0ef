"""ORACLE -- test infrastructure, not product code.

Writer of the reference's dump files, restating ``dumpdata_npe`` (io.F90:213-248) with the
record layout of ``DUMP_PARAM`` / ``DUMP_HEADER`` / ``DUMP_DATA`` (io.F90:194-211) and
``reflux_write`` (reflux.F90:392-413), so that tests can produce dumps of known content for the
product's reader (``code_only_amd.dump``).  It is written from the reference's text independently
of the reader: each record is assembled here from the Fortran declarations (4-byte integers and
logicals, 8-byte reals, 64-bit Step/Dstep, column-major arrays) and framed with the sequential
unformatted record markers.  The reference ships no dump files, so this is the only source of
dumps the tests have.
"""
import numpy as np

IO_VERSION = 5
UNDEF = -1


class RecordWriter:
    def __init__(self, path, marker=4, endian="<"):
        self.f = open(path, "wb")
        self.mt = endian + ("i4" if marker == 4 else "i8")
        self.endian = endian

    def record(self, *arrays):
        """One record: the arrays' values, each in Fortran (column-major) element order."""
        body = b"".join(np.asarray(a).astype(np.asarray(a).dtype.newbyteorder(self.endian)).tobytes(order="F")
                        for a in arrays)
        n = np.array([len(body)], self.mt).tobytes()
        self.f.write(n + body + n)

    def close(self):
        self.f.close()


def i4(*v):
    return np.array(v, np.int32)


def write_rank(path, params, header, grids, marker=4, endian="<"):
    """params: dict of the DUMP_PARAM names; header: dict with level_max, gidlistmax (per level),
    gidlist (2-D [gid slot, level], Fortran order), time, dtime, step, dstep (per level),
    cellwidth ([3, levels]), ngrid_global; grids: list of dicts (gid, level, u1list, u2list,
    alloc (3 bools), u, u1, u2 as Fortran arrays (i, j, k, m), x, y, z, ijk, nbr_gid / nbr_rank
    (2, 3), child_gid / child_rank (2, 2, 2), parent_gid, parent_rank, flux: 12 arrays)."""
    w = RecordWriter(path, marker, endian)
    P = params
    w.record(i4(IO_VERSION))
    w.record(i4(header["ngrid_global"]))
    w.record(i4(len(grids)))
    w.record(i4(P["Imin"], P["Jmin"], P["Kmin"], P["Mmin"], P["Gidmin"], P["Lmin"]))
    w.record(i4(P["Imax"], P["Jmax"], P["Kmax"], P["Mmax"], P["Gidmax"], P["Lmax"]))
    w.record(i4(P["Ngh"]))
    w.record(i4(P["Imingh"], P["Jmingh"], P["Kmingh"]))
    w.record(i4(P["Imaxgh"], P["Jmaxgh"], P["Kmaxgh"]))
    w.record(i4(header["level_max"]))
    w.record(np.asarray(header["gidlistmax"], np.int32))
    w.record(np.asarray(header["gidlist"], np.int32))
    w.record(np.asarray(header["time"], np.float64))
    w.record(np.asarray(header["dtime"], np.float64))
    w.record(np.asarray(header["step"], np.int64))
    w.record(np.asarray(header["dstep"], np.int64))
    w.record(np.asarray(header["cellwidth"], np.float64))
    for g in grids:
        w.record(i4(g["gid"]))
        w.record(i4(g["level"]))
        w.record(i4(g["u1list"]))
        w.record(i4(g["u2list"]))
        w.record(np.asarray(g["alloc"], np.int32))          # default LOGICAL: 4 bytes, .TRUE. = 1
        for key in ("u", "u1", "u2"):
            w.record(np.asarray(g[key], np.float64))
        for key in ("x", "y", "z"):
            w.record(np.asarray(g[key], np.float64))
        w.record(i4(*g["ijk"]))
        w.record(np.asarray(g["nbr_gid"], np.int32), np.asarray(g["nbr_rank"], np.int32))
        w.record(np.asarray(g["child_gid"], np.int32), np.asarray(g["child_rank"], np.int32))
        w.record(i4(g["parent_gid"], g["parent_rank"]))
        for a in g["flux"]:
            w.record(np.asarray(a, np.float64))
    w.close()
