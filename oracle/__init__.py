"""ORACLE -- test infrastructure, not product code.

CPU FP64 restatement of the reference's self-gravity multigrid (``liborc.so``, built from the C
sources in this directory by its Makefile) plus closed-form checks (``fft_poisson``).  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import this package,
and only as the checker; the product path in ``code_only_amd`` never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

STOP_REFERENCE = 0
STOP_TRUE_RESIDUAL = 1
HIST = 64


class OrcParams(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("maxit", ctypes.c_int), ("mode", ctypes.c_int),
                ("ncycle", ctypes.c_int), ("npre", ctypes.c_int), ("npost", ctypes.c_int)]


class OrcStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("status", ctypes.c_int),
                ("rel_linf", ctypes.c_double), ("rel_l2", ctypes.c_double),
                ("resmaxg", ctypes.c_double), ("hist", ctypes.c_double * HIST)]


def build():
    """Compile liborc.so with the committed Makefile."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liborc.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        ci = ctypes.c_int
        cd = ctypes.c_double
        L.orc_relax_dense.argtypes = [dp, dp, ci, cd]
        L.orc_relax_dense.restype = cd
        L.orc_resid_dense.argtypes = [dp, dp, dp, ci, cd]
        L.orc_rstrct_dense.argtypes = [dp, dp, ci, ci]
        L.orc_interp_dense.argtypes = [dp, dp, ci, ci]
        L.orc_mg_cycle_dense.argtypes = [dp, dp, ci, cd, ctypes.POINTER(OrcParams), ctypes.POINTER(OrcStats)]
        L.orc_mg_cycle_dense.restype = ci
        L.orc_gradient_dense.argtypes = [dp, dp, ci, cd]
        L.orc_cubic_stencil.argtypes = [ci, ci, ci, ctypes.POINTER(ci), dp]
        L.orc_rstrct_cubic_weights.argtypes = [dp]
        _LIB = L
    return _LIB


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def relax(u, rhs, h):
    """One mg_poisson_relax call (both colours) in place on dense n^3 ``u``; returns Resmaxg."""
    n = u.shape[0]
    return lib().orc_relax_dense(_p(u), _p(np.ascontiguousarray(rhs, np.float64)), n, h)


def resid(u, rhs, h):
    n = u.shape[0]
    res = np.zeros_like(u)
    lib().orc_resid_dense(_p(res), _p(np.ascontiguousarray(u)), _p(np.ascontiguousarray(rhs)), n, h)
    return res


def restrict(uf, cubic=False):
    nc = uf.shape[0] // 2
    uc = np.zeros((nc, nc, nc))
    lib().orc_rstrct_dense(_p(uc), _p(np.ascontiguousarray(uf)), nc, int(cubic))
    return uc


def interp(uc, cubic=False):
    nf = uc.shape[0] * 2
    uf = np.zeros((nf, nf, nf))
    lib().orc_interp_dense(_p(uf), _p(np.ascontiguousarray(uc)), nf, int(cubic))
    return uf


def cubic_stencil(f, lo, hi):
    idx = (ctypes.c_int * 4)()
    a = (ctypes.c_double * 4)()
    lib().orc_cubic_stencil(f, lo, hi, idx, a)
    return list(idx), list(a)


def rstrct_cubic_weights():
    a = (ctypes.c_double * 4)()
    lib().orc_rstrct_cubic_weights(a)
    return list(a)


def mg_cycle(src, h, psi=None, tol=1e-10, maxit=50, mode=STOP_TRUE_RESIDUAL, ncycle=2, npre=2, npost=2):
    """Uniform periodic mg_cycle (mg.F90:42-94) on dense n^3 arrays.  Returns (psi, stats dict)."""
    n = src.shape[0]
    src = np.ascontiguousarray(src, np.float64)
    psi = np.zeros_like(src) if psi is None else np.ascontiguousarray(psi, np.float64).copy()
    p = OrcParams(tol, maxit, mode, ncycle, npre, npost)
    st = OrcStats()
    lib().orc_mg_cycle_dense(_p(psi), _p(src), n, h, ctypes.byref(p), ctypes.byref(st))
    stats = dict(iterations=st.iterations, status=st.status, rel_linf=st.rel_linf, rel_l2=st.rel_l2,
                 resmaxg=st.resmaxg, hist=[st.hist[i] for i in range(min(st.iterations + 1, HIST))])
    return psi, stats


def gradient(u, h):
    n = u.shape[0]
    g = np.zeros((3, n, n, n))
    lib().orc_gradient_dense(_p(g), _p(np.ascontiguousarray(u)), n, h)
    return g
