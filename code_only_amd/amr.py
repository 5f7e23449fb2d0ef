"""Host-side mirror of the reference's AMR self-gravity interface (module ``fmg``: fmg_poisson,
fmg.F90:63-74, with fmg_data's per-block arrays) over libsg's C ABI.

Block arrays are float64, dense [nblk, 8, 8, 8] in (k, j, i) order (x fastest); the hierarchy
comes from ``code_only_amd.amr_hierarchy``.  ``solve`` takes torch GPU tensors; ``solve_host``
numpy arrays.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import (SG_BC_PERIODIC, SG_BC_SPHERE, SG_ERR_NONFINITE, SG_ERR_NOT_CONVERGED, SG_PREC_FP64,
                   SG_PREC_MIXED, SG_STOP_REFERENCE, SG_STOP_TRUE_RESIDUAL)

__all__ = ["AmrSolver", "SG_BC_PERIODIC", "SG_BC_SPHERE", "SG_PREC_FP64", "SG_PREC_MIXED", "SG_STOP_REFERENCE",
           "SG_STOP_TRUE_RESIDUAL"]

U, RES, RHS, RHO, SRC, PSI = range(6)


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class AmrSolver:
    """One hierarchy with all of its FMG-level block arrays on the current GPU (fmg_data_init,
    fmg_alloc_LoL and fmg_make_linklist, done once)."""

    def __init__(self, hier, zero_average=None):
        self.h = hier
        self._lib = _lib.load()
        self._keep = [np.ascontiguousarray(a, np.int32) for a in (hier.level, hier.ijk, hier.parent, hier.child,
                                                                  hier.nbr, np.asarray(hier.base_dims))]
        self._x0 = np.ascontiguousarray(hier.x0, np.float64)
        ctx = ctypes.c_void_p()
        _lib.check(self._lib.sg_amr_create(ctypes.byref(ctx), hier.nblk, hier.nlev, *[_ip(a) for a in self._keep],
                                           float(hier.h0), _dp(self._x0), int(hier.bc)))
        self._ctx = ctx
        if zero_average is not None:
            _lib.check(self._lib.sg_amr_set_zero_average(self._ctx, int(bool(zero_average))))

    def set_zero_average(self, on):
        """ZERO_AVERAGE_DENSITY on or off for later solves (fmg_data_if.F90:53-78, vmg.F90:213-216)."""
        _lib.check(self._lib.sg_amr_set_zero_average(self._ctx, int(bool(on))))

    def close(self):
        if getattr(self, "_ctx", None):
            self._lib.sg_amr_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _params(tol, maxit, stop_mode, precision=SG_PREC_FP64, use_graph=True, fused=True):
        return _lib.default_params(tol=tol, maxit=maxit, stop_mode=stop_mode, precision=precision,
                                   use_graph=int(bool(use_graph)), fused=int(bool(fused)))

    def info(self):
        """Context facts: FMG levels whose mixed-solve refreshes run fused (bit mask), kernel nodes of the
        captured mixed outer iteration, blocks, coarse-fine faces, AMR levels whose vmg_ghostcell
        runs as one gather in the fused mode (bit mask), FMG levels whose base-level
        fmg_ghostcell_fix runs as one gather in the fused mode (bit mask)."""
        v = (ctypes.c_int * 6)()
        _lib.check(self._lib.sg_amr_info(self._ctx, v, 6))
        return dict(fused_levels=v[0], graph_nodes=v[1], nblk=v[2], cf_faces=v[3], vmg_fill_levels=v[4],
                    base_gather_levels=v[5])

    def tile_info(self):
        """The autotuner's threads per 8^3 block of the FMG-level-0 relaxation / residual kernels and
        the candidates' ms per launch."""
        out = {}
        for which, name in enumerate(("relax_fp32", "resid_fp32", "relax_fp64", "resid_fp64")):
            c = ctypes.c_int()
            ms = (ctypes.c_float * 3)()
            _lib.check(self._lib.sg_amr_tile_info(self._ctx, which, ctypes.byref(c), ms, 3))
            out[name] = {"threads_per_block": c.value,
                         "ms_per_launch": {str(nt): round(ms[q], 5) for q, nt in enumerate((64, 128, 256)) if ms[q] > 0}}
        return out

    def set_base_solver(self, chip):
        """Base-grid mg_cycle of the plain FP64 solve, and the mixed solve's exact base solve, spread
        over the chip (True, default) or in one workgroup (False); the values are the same."""
        _lib.check(self._lib.sg_amr_set_base_solver(self._ctx, 1 if chip else 0))

    def set_fp64_refresh(self, fused):
        """Plain FP64 solve with the staged refresh of the reference (False, default) or the mixed
        solve's fused launch structure (True); the values are the same."""
        _lib.check(self._lib.sg_amr_set_fp64_refresh(self._ctx, 1 if fused else 0))

    def time_smoother(self, precision=SG_PREC_MIXED, reps=20):
        """ms per full red-black sweep (both colours) of the FMG-level-0 relaxation, HIP events on
        the library's stream."""
        ms = ctypes.c_double()
        _lib.check(self._lib.sg_amr_time_smoother(self._ctx, precision, reps, ctypes.byref(ms)))
        return ms.value

    def solve(self, psi, src, g=None, tol=1e-10, maxit=30, stop_mode=SG_STOP_TRUE_RESIDUAL, precision=SG_PREC_FP64,
              use_graph=True, raise_on_fail=True, fused=True):
        """fmg_poisson on torch float64 GPU tensors [nblk, 8, 8, 8]; psi in place, g [nblk, 3, 8, 8, 8].
        With raise_on_fail False, SG_ERR_NOT_CONVERGED / SG_ERR_NONFINITE come back in the stats'
        status instead of an exception."""
        import torch
        for t in (psi, src) + ((g,) if g is not None else ()):
            if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
                raise ValueError("expected contiguous float64 GPU tensors")
        if tuple(psi.shape) != (self.h.nblk, 8, 8, 8) or psi.shape != src.shape:
            raise ValueError("psi and src must be [nblk, 8, 8, 8]")
        p = self._params(tol, maxit, stop_mode, precision, use_graph, fused)
        st = _lib.SgStats()
        torch.cuda.synchronize()
        _lib.check(self._lib.sg_amr_solve_dev(self._ctx, ctypes.c_void_p(psi.data_ptr()),
                                              ctypes.c_void_p(src.data_ptr()),
                                              ctypes.c_void_p(g.data_ptr()) if g is not None else None,
                                              ctypes.byref(p), ctypes.byref(st)),
                   allow=() if raise_on_fail else (SG_ERR_NOT_CONVERGED, SG_ERR_NONFINITE))
        return st.as_dict()

    def adjoint(self, psi_bar, tol=1e-10, maxit=200, precision=SG_PREC_MIXED, use_graph=True, fused=True,
                raise_on_fail=True):
        """src_bar = (d psi / d src)^T psi_bar for the sphere-boundary solve (sg_amr_adjoint_dev), torch
        float64 GPU tensors [nblk, 8, 8, 8]; precision is that of the forward solves preconditioning the
        adjoint's FGMRES, tol its relative residual.  Returns (src_bar, stats)."""
        import torch
        if not (psi_bar.is_cuda and psi_bar.dtype == torch.float64 and psi_bar.is_contiguous()):
            raise ValueError("expected a contiguous float64 GPU tensor")
        if tuple(psi_bar.shape) != (self.h.nblk, 8, 8, 8):
            raise ValueError("psi_bar must be [nblk, 8, 8, 8]")
        out = torch.empty_like(psi_bar)
        p = self._params(tol, maxit, SG_STOP_TRUE_RESIDUAL, precision, use_graph, fused)
        st = _lib.SgStats()
        torch.cuda.synchronize()
        _lib.check(self._lib.sg_amr_adjoint_dev(self._ctx, ctypes.c_void_p(out.data_ptr()),
                                                ctypes.c_void_p(psi_bar.data_ptr()), ctypes.byref(p), ctypes.byref(st)),
                   allow=() if raise_on_fail else (SG_ERR_NOT_CONVERGED,))
        return out, st.as_dict()

    def adjoint_apply(self, x):
        """y = A^T x for the homogeneous composite operator A at FMG level 0 (sg_amr_adjoint_apply)."""
        import torch
        if not (x.is_cuda and x.dtype == torch.float64 and x.is_contiguous()) or tuple(x.shape) != (self.h.nblk, 8, 8, 8):
            raise ValueError("expected a contiguous float64 GPU tensor [nblk, 8, 8, 8]")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        _lib.check(self._lib.sg_amr_adjoint_apply(self._ctx, ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(x.data_ptr())))
        return y

    def solve_host(self, src, psi=None, want_g=True, tol=1e-10, maxit=30, stop_mode=SG_STOP_TRUE_RESIDUAL,
                   precision=SG_PREC_FP64, use_graph=True, raise_on_fail=True, fused=True):
        src = np.ascontiguousarray(src, np.float64)
        psi = np.zeros_like(src) if psi is None else np.ascontiguousarray(psi, np.float64).copy()
        g = np.zeros((self.h.nblk, 3, 8, 8, 8)) if want_g else None
        p = self._params(tol, maxit, stop_mode, precision, use_graph, fused)
        st = _lib.SgStats()
        _lib.check(self._lib.sg_amr_solve_host(self._ctx, _dp(psi), _dp(src), _dp(g) if want_g else None,
                                               ctypes.byref(p), ctypes.byref(st)),
                   allow=() if raise_on_fail else (SG_ERR_NOT_CONVERGED, SG_ERR_NONFINITE))
        return psi, g, st.as_dict()

    # ---- hydro block memory on the device (BlockMem(gid)%u, grid.F90:146-194) ----
    def set_blockmem(self, layout, pool_slots, slot_u, slot_u1=None, slot_u2=None):
        """Layout (dict of sg_blockmem's fields: ng, ncomp, stride, mrho, mvx, mvy, mvz, mp, mpsi, mgx,
        mgy, mgz) and per block the pool slots of its own, predictor-stage and past arrays."""
        lay = _lib.SgBlockmem(**{k: int(v) for k, v in layout.items()})
        tabs = [None if a is None else np.ascontiguousarray(a, np.int32) for a in (slot_u, slot_u1, slot_u2)]
        for a in tabs:
            if a is not None and a.shape != (self.h.nblk,):
                raise ValueError("slot tables hold one entry per block")
        _lib.check(self._lib.sg_amr_set_blockmem(self._ctx, ctypes.byref(lay), int(pool_slots),
                                                 *[None if a is None else _ip(a) for a in tabs]))

    def solve_blockmem(self, pool, pi4g, source=0, step_max=0.0, step=(), dstep=(), tol=1e-10, maxit=30,
                       stop_mode=SG_STOP_TRUE_RESIDUAL, precision=SG_PREC_FP64, use_graph=True, raise_on_fail=True,
                       fused=True):
        """fmg_poisson on a float64 GPU tensor holding the pool: stage psi and 4 pi G rho (time slice
        `source` of fmg_prepare_interp), solve, write psi back over the FMG array's extent and g
        over the interior."""
        import torch
        if not (pool.is_cuda and pool.dtype == torch.float64 and pool.is_contiguous()):
            raise ValueError("expected a contiguous float64 GPU tensor")
        ts = _lib.SgTimeslice(source=int(source), step_max=float(step_max))
        for l, v in enumerate(step):
            ts.step[l] = float(v)
        for l, v in enumerate(dstep):
            ts.dstep[l] = float(v)
        p = self._params(tol, maxit, stop_mode, precision, use_graph, fused)
        st = _lib.SgStats()
        torch.cuda.synchronize()
        _lib.check(self._lib.sg_amr_solve_blockmem_dev(self._ctx, ctypes.c_void_p(pool.data_ptr()), ctypes.byref(ts),
                                                       float(pi4g), ctypes.byref(p), ctypes.byref(st)),
                   allow=() if raise_on_fail else (SG_ERR_NOT_CONVERGED, SG_ERR_NONFINITE))
        return st.as_dict()

    def source_g(self, pool, dtime, gamma=1.4):
        """source_g_all_level on the pool (float64 GPU tensor), dtime[l] = Dtime(l) per AMR level."""
        import torch
        if not (pool.is_cuda and pool.dtype == torch.float64 and pool.is_contiguous()):
            raise ValueError("expected a contiguous float64 GPU tensor")
        dt = np.ascontiguousarray(dtime, np.float64)
        if dt.shape != (self.h.nlev,):
            raise ValueError("dtime holds one value per AMR level")
        torch.cuda.synchronize()
        _lib.check(self._lib.sg_amr_source_g_dev(self._ctx, ctypes.c_void_p(pool.data_ptr()), _dp(dt), float(gamma)))

    # ---- hydro steps on the pool (grid_boundary.F90, fg2cg.F90, reflux.F90) ----
    @staticmethod
    def _pool_ok(t):
        import torch
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
            raise ValueError("expected a contiguous float64 GPU tensor")

    def hydro_boundary(self, pool, level, mode=0):
        """boundary_grid(level, mode): mode 0 complete, 1 predictor, 2 corrector."""
        import torch
        self._pool_ok(pool)
        torch.cuda.synchronize()
        _lib.check(self._lib.sg_amr_hydro_boundary_dev(self._ctx, ctypes.c_void_p(pool.data_ptr()), int(level), int(mode)))

    def fg2cg(self, pool, level, gamma=1.4):
        """fg2cg_u(level): the level's grids restricted into their parents."""
        import torch
        self._pool_ok(pool)
        torch.cuda.synchronize()
        _lib.check(self._lib.sg_amr_fg2cg_dev(self._ctx, ctypes.c_void_p(pool.data_ptr()), int(level), float(gamma)))

    def fluxcorrection(self, pool, fb, level, dtime, gamma=1.4):
        """fluxcorrection(level) with the face-flux buffers fb [nblk, 2, 6, 9, 64] (GPU float64)."""
        import torch
        self._pool_ok(pool)
        self._pool_ok(fb)
        if tuple(fb.shape) != (self.h.nblk, 2, 6, 9, 64):
            raise ValueError("fb must be [nblk, 2, 6, 9, 64]")
        dt = np.ascontiguousarray(dtime, np.float64)
        if dt.shape != (self.h.nlev,):
            raise ValueError("dtime holds one value per AMR level")
        torch.cuda.synchronize()
        _lib.check(self._lib.sg_amr_fluxcorrection_dev(self._ctx, ctypes.c_void_p(pool.data_ptr()),
                                                       ctypes.c_void_p(fb.data_ptr()), int(level), _dp(dt), float(gamma)))

    # ---- operator-level access (parity tests) ----
    def get(self, L, code):
        n = 8 >> L
        a = np.empty((self.h.nblk, n + 2, n + 2, n + 2))
        _lib.check(self._lib.sg_amr_get_array(self._ctx, L, code, _dp(a)))
        return a

    def set(self, L, code, a):
        n = 8 >> L
        a = np.ascontiguousarray(a, np.float64).reshape(self.h.nblk, n + 2, n + 2, n + 2)
        _lib.check(self._lib.sg_amr_set_array(self._ctx, L, code, _dp(a)))

    @property
    def fill0_level(self):
        v = ctypes.c_int()
        _lib.check(self._lib.sg_amr_get_fill0_level(self._ctx, ctypes.byref(v)))
        return v.value

    @fill0_level.setter
    def fill0_level(self, v):
        _lib.check(self._lib.sg_amr_set_fill0_level(self._ctx, int(v)))

    def op(self, op, L=0, a=0, b=0, c=0, flag=0):
        r = ctypes.c_double()
        _lib.check(self._lib.sg_amr_op(self._ctx, op, L, a, b, c, flag, ctypes.byref(r)))
        return r.value
