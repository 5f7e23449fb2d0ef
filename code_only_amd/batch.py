"""Batched uniform solves over libsg's C ABI: many independent periodic mg_cycle problems
(mg.F90:42-94), each its own n^3 grid (n = 2^a or 3 * 2^a) and spacing, solved together with one
launch per multigrid step (sg_batch_*).  Arrays are dense float64 GPU tensors [k][j][i]."""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import (SG_COARSE_EXACT, SG_COARSE_VCYCLE, SG_ERR_NONFINITE, SG_ERR_NOT_CONVERGED, SG_PREC_FP64,
                   SG_PREC_MIXED, SG_STOP_TRUE_RESIDUAL)

__all__ = ["BatchSolver", "SG_PREC_FP64", "SG_PREC_MIXED", "SG_COARSE_EXACT", "SG_COARSE_VCYCLE"]


class BatchSolver:
    """Level hierarchies of every problem of the batch on the current GPU (one mg_init per
    problem).  ``h`` defaults to 1/n (unit box) per problem."""

    def __init__(self, sizes, h=None, precisions=SG_PREC_FP64 | SG_PREC_MIXED):
        self.sizes = [int(n) for n in sizes]
        self.h = [1.0 / n for n in self.sizes] if h is None else [float(x) for x in h]
        if len(self.h) != len(self.sizes):
            raise ValueError("one spacing per problem")
        self._lib = _lib.load()
        ctx = ctypes.c_void_p()
        n_arr = (ctypes.c_int * len(self.sizes))(*self.sizes)
        h_arr = (ctypes.c_double * len(self.h))(*self.h)
        _lib.check(self._lib.sg_batch_create(ctypes.byref(ctx), len(self.sizes), n_arr, h_arr, int(precisions)))
        self._ctx = ctx

    def close(self):
        if getattr(self, "_ctx", None):
            self._lib.sg_batch_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, psis, srcs, tol=1e-10, maxit=30, precision=SG_PREC_MIXED, ncycle=2, npre=2, npost=2,
              use_graph=True, raise_on_fail=True, coarse=SG_COARSE_EXACT):
        """Solve every problem in place (psis[q] in = initial guess).  Returns per-problem stats.
        coarse: as UniformSolver.cycle's (mixed precision only)."""
        npb = len(self.sizes)
        if len(psis) != npb or len(srcs) != npb:
            raise ValueError(f"expected {npb} psi and src tensors")
        for q, (a, b) in enumerate(zip(psis, srcs)):
            n = self.sizes[q]
            for t in (a, b):
                if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous() and tuple(t.shape) == (n,) * 3):
                    raise ValueError(f"problem {q}: expected a contiguous float64 GPU tensor of shape {(n,) * 3}")
        p = _lib.default_params(tol=tol, maxit=maxit, stop_mode=SG_STOP_TRUE_RESIDUAL, precision=precision,
                                ncycle=ncycle, npre=npre, npost=npost, use_graph=int(bool(use_graph)), coarse=coarse)
        psi_arr = (ctypes.c_void_p * npb)(*[t.data_ptr() for t in psis])
        src_arr = (ctypes.c_void_p * npb)(*[t.data_ptr() for t in srcs])
        st = (_lib.SgStats * npb)()
        torch.cuda.synchronize()
        rc = self._lib.sg_batch_solve_dev(self._ctx, psi_arr, src_arr, ctypes.byref(p), st)
        allow = () if raise_on_fail else (SG_ERR_NOT_CONVERGED, SG_ERR_NONFINITE)
        _lib.check(rc, allow)
        return [st[q].as_dict() for q in range(npb)]

    def solve_host(self, srcs, **kw):
        """numpy in/out convenience: returns (list of psi arrays, stats)."""
        dev = torch.device("cuda")
        s_t = [torch.tensor(np.ascontiguousarray(s, np.float64), device=dev) for s in srcs]
        p_t = [torch.zeros_like(s) for s in s_t]
        st = self.solve(p_t, s_t, **kw)
        return [p.cpu().numpy() for p in p_t], st

    def tile_info(self, level):
        cfg, ns, fs, fx = (ctypes.c_int() for _ in range(4))
        ms = (ctypes.c_float * 32)()
        _lib.check(self._lib.sg_batch_tile_info(self._ctx, level, ctypes.byref(cfg), ctypes.byref(ns),
                                                ctypes.byref(fs), ctypes.byref(fx), ms, 32))
        return dict(cfg=cfg.value, sweeps_per_launch=ns.value, source_in_lds=bool(fs.value),
                    cubic_fx=fx.value, ms=[ms[i] for i in range(32)])

    def levels(self, q):
        nl, lc = ctypes.c_int(), ctypes.c_int()
        _lib.check(self._lib.sg_batch_levels(self._ctx, q, ctypes.byref(nl), ctypes.byref(lc)))
        return nl.value, lc.value
