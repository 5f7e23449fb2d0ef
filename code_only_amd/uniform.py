"""Host-side mirror of the reference's uniform-grid multigrid interface (module ``mg``,
mg.F90:14: mg_init, mg_cycle, mg_finalize) over libsg's C ABI, for torch device tensors.

Arrays are dense n^3 float64 CUDA(HIP) tensors, C order [k][j][i], x fastest.
"""
import ctypes

import torch

from . import _lib
from ._lib import (SG_COARSE_EXACT, SG_COARSE_VCYCLE, SG_ERR_NONFINITE, SG_ERR_NOT_CONVERGED, SG_PREC_FP64,
                   SG_PREC_MIXED, SG_STOP_REFERENCE, SG_STOP_TRUE_RESIDUAL)

__all__ = ["UniformSolver", "SG_PREC_FP64", "SG_PREC_MIXED", "SG_STOP_REFERENCE", "SG_STOP_TRUE_RESIDUAL",
           "SG_COARSE_EXACT", "SG_COARSE_VCYCLE"]


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class UniformSolver:
    """One periodic n^3 level hierarchy on the current GPU (mg_init / mg_data_init,
    mg.F90:525-528, mg_data.F90:45-90).  ``precisions`` selects which level arrays to allocate."""

    def __init__(self, n, h=None, precisions=SG_PREC_FP64 | SG_PREC_MIXED):
        self.n = int(n)
        self.h = float(h) if h is not None else 1.0 / n
        self._lib = _lib.load()
        ctx = ctypes.c_void_p()
        _lib.check(self._lib.sg_uniform_create(ctypes.byref(ctx), self.n, self.h, int(precisions)))
        self._ctx = ctx

    def close(self):
        if getattr(self, "_ctx", None):
            self._lib.sg_uniform_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, t, dtype=torch.float64, shape=None):
        if not (t.is_cuda and t.dtype == dtype and t.is_contiguous()):
            raise ValueError(f"expected a contiguous {dtype} GPU tensor, got {t.dtype} on {t.device}")
        if shape is not None and tuple(t.shape) != tuple(shape):
            raise ValueError(f"expected shape {shape}, got {tuple(t.shape)}")

    def cycle(self, psi, src, tol=1e-10, maxit=50, stop_mode=SG_STOP_TRUE_RESIDUAL, precision=SG_PREC_MIXED,
              ncycle=2, npre=2, npost=2, use_graph=True, fused=True, raise_on_fail=True, coarse=SG_COARSE_EXACT):
        """mg_cycle(ju, jrho) (mg.F90:42-94): psi updated in place; returns stats dict.  coarse
        (mixed precision only): SG_COARSE_EXACT solves the first level of <= 16^3 cells exactly,
        SG_COARSE_VCYCLE keeps the reference's levels down to 1^3."""
        n = self.n
        self._check(psi, shape=(n, n, n))
        self._check(src, shape=(n, n, n))
        p = _lib.default_params(tol=tol, maxit=maxit, stop_mode=stop_mode, precision=precision, ncycle=ncycle,
                                npre=npre, npost=npost, use_graph=int(bool(use_graph)), fused=int(bool(fused)),
                                coarse=coarse)
        st = _lib.SgStats()
        torch.cuda.synchronize()
        rc = self._lib.sg_uniform_cycle_dev(self._ctx, _ptr(psi), _ptr(src), ctypes.byref(p), ctypes.byref(st))
        allow = () if raise_on_fail else (SG_ERR_NOT_CONVERGED, SG_ERR_NONFINITE)
        _lib.check(rc, allow=allow)
        return st.as_dict()

    def gradient(self, psi):
        """g = -grad psi, centred (fmg_poisson.F90:29-31); returns (3, n, n, n)."""
        n = self.n
        self._check(psi, shape=(n, n, n))
        g = torch.empty((3, n, n, n), dtype=torch.float64, device=psi.device)
        torch.cuda.synchronize()
        _lib.check(self._lib.sg_uniform_gradient_dev(self._ctx, _ptr(psi), _ptr(g)))
        return g

    def adjoint(self, adj, tol=1e-10, maxit=50, stop_mode=SG_STOP_TRUE_RESIDUAL, precision=SG_PREC_MIXED,
                use_graph=True, fused=True, coarse=SG_COARSE_EXACT):
        """d loss / d src = P L^+ P adj for the periodic solve (sg_uniform_adjoint_dev)."""
        n = self.n
        self._check(adj, shape=(n, n, n))
        grad = torch.empty_like(adj)
        p = _lib.default_params(tol=tol, maxit=maxit, stop_mode=stop_mode, precision=precision,
                                use_graph=int(bool(use_graph)), fused=int(bool(fused)), coarse=coarse)
        st = _lib.SgStats()
        torch.cuda.synchronize()
        _lib.check(self._lib.sg_uniform_adjoint_dev(self._ctx, _ptr(adj), _ptr(grad), ctypes.byref(p), ctypes.byref(st)))
        return grad, st.as_dict()

    def gradient_sample(self, psi, cells):
        """(npts, 3): centred grad psi at linear cell indices ``cells`` (int32 GPU tensor)."""
        self._check(psi, shape=(self.n,) * 3)
        self._check(cells, dtype=torch.int32)
        out = torch.empty((cells.numel(), 3), dtype=torch.float64, device=psi.device)
        torch.cuda.synchronize()
        _lib.check(self._lib.sg_uniform_gradient_sample_dev(self._ctx, _ptr(psi), _ptr(cells), cells.numel(), _ptr(out)))
        return out

    def gradient_sample_adjoint(self, cells, w):
        """Transpose of gradient_sample: the n^3 field sum_p w_p . d(grad psi(c_p))/d psi."""
        self._check(cells, dtype=torch.int32)
        self._check(w, shape=(cells.numel(), 3))
        out = torch.empty((self.n,) * 3, dtype=torch.float64, device=w.device)
        torch.cuda.synchronize()
        _lib.check(self._lib.sg_uniform_gradient_sample_adjoint_dev(self._ctx, _ptr(cells), _ptr(w), cells.numel(),
                                                                     _ptr(out)))
        return out

    def tile_info(self, level):
        """Autotuned sweep configuration of fused level ``level`` and all candidate timings."""
        cfg, ns, fs = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ms = (ctypes.c_float * 32)()
        _lib.check(self._lib.sg_uniform_tile_info(self._ctx, level, ctypes.byref(cfg), ctypes.byref(ns),
                                                  ctypes.byref(fs), ms, 32))
        return dict(cfg=cfg.value, ns=ns.value, fs=fs.value, ms=list(ms))

    def set_tile(self, level, cfg, ns=1, fs=0):
        """Force the sweep configuration of fused level ``level`` (values do not change)."""
        _lib.check(self._lib.sg_uniform_set_tile(self._ctx, level, cfg, ns, fs))

    # single operators (parity tests)
    def relax(self, u, rhs, h, prec=SG_PREC_FP64):
        r = ctypes.c_double()
        torch.cuda.synchronize()
        _lib.check(self._lib.sg_uniform_relax_dev(self._ctx, prec, _ptr(u), _ptr(rhs), u.shape[0], h, ctypes.byref(r)))
        return r.value

    def resid(self, u, rhs, h, prec=SG_PREC_FP64):
        res = torch.empty_like(u)
        torch.cuda.synchronize()
        _lib.check(self._lib.sg_uniform_resid_dev(self._ctx, prec, _ptr(res), _ptr(u), _ptr(rhs), u.shape[0], h))
        return res

    def restrict(self, uf, prec=SG_PREC_FP64):
        nc = uf.shape[0] // 2
        uc = torch.empty((nc, nc, nc), dtype=uf.dtype, device=uf.device)
        torch.cuda.synchronize()
        _lib.check(self._lib.sg_uniform_restrict_dev(self._ctx, prec, _ptr(uc), _ptr(uf), nc))
        return uc

    def interp(self, uc, cubic, prec=SG_PREC_FP64):
        nf = uc.shape[0] * 2
        uf = torch.empty((nf, nf, nf), dtype=uc.dtype, device=uc.device)
        torch.cuda.synchronize()
        _lib.check(self._lib.sg_uniform_interp_dev(self._ctx, prec, _ptr(uf), _ptr(uc), nf, int(bool(cubic))))
        return uf
