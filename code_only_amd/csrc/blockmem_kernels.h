// Kernels between the hydro block memory and the AMR solver, and the gravity source step, for a
// device pool of per-grid arrays u(i, j, k, m) (BlockMem(gid)%u, grid.F90:146-194): i, j, k in
// [-ng, 8 + ng), m in [0, ncomp), i fastest, grid slot s at pool + s * stride.  One thread per
// cell; consecutive threads take consecutive i, so every component's loads and stores coalesce.
// All of it is byte traffic with a handful of flops per cell: HBM-bound.
#pragma once

#include <hip/hip_runtime.h>

#include "amr_kernels.h"
#include "sg_common.h"

namespace sg {
namespace amr {

// fmg_prepare_interp's coefficients per AMR level (fmg_data_if.F90:186-200): fine[l] when
// Step(l) == Step(LevelMax), else stp1 = stp - (Step(l) - Dstep(l)), stp2 = Step(l) - stp and
// dstp = Dstep(l) with stp = Step(LevelMax)
struct SliceCoef {
    int source;
    int fine[16];
    double stp1[16], stp2[16], dstp[16];
};

__device__ __forceinline__ long bm_cell(const sg_blockmem &L, int i, int j, int k)
{
    const long W = 8 + 2 * L.ng;
    return ((long)(k + L.ng) * W + (j + L.ng)) * W + (i + L.ng);
}

// fmg_make_physvar's Poisson staging (fmg_data_if.F90:38-51): psi <- u(MPSI) of the current
// state, src <- 4 pi G times the time slice of u(MRHO), dense [nblk][8^3] interiors (the solve
// reads no ghost cell of either before it writes it).  Time slice 3 in the reference's operand
// order: (u0 - (2 u1 - u2)) (stp1/dstp)^2 + (u2 stp2 + (2 u1 - u2) stp1) / dstp.
__global__ void k_bm_stage(double *psi, double *src, const double *pool, const int *su, const int *s1, const int *s2,
                           int nblk, const int *level, sg_blockmem L, SliceCoef C, double pi4g)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nblk * 512) return;
    const int b = (int)(t >> 9), c = (int)(t & 511);
    const long W = 8 + 2 * L.ng, W3 = W * W * W, q = bm_cell(L, c & 7, (c >> 3) & 7, c >> 6);
    const double *u0 = pool + (long)su[b] * L.stride;
    psi[t] = u0[L.mpsi * W3 + q];
    const long rq = L.mrho * W3 + q;
    const int l = level[b];
    double v;
    if (C.source == 0 || (C.source == 3 && C.fine[l])) {
        v = u0[rq];
    } else if (C.source == 1) {
        v = pool[(long)s1[b] * L.stride + rq];
    } else if (C.source == 2) {
        v = pool[(long)s2[b] * L.stride + rq];
    } else {
        const double x0 = u0[rq], x1 = pool[(long)s1[b] * L.stride + rq], x2 = pool[(long)s2[b] * L.stride + rq];
        const double stp1 = C.stp1[l], stp2 = C.stp2[l], dstp = C.dstp[l];
        const double r = stp1 / dstp;
        const double w = 2.0 * x1 - x2;
        const double a = (x0 - w) * (r * r);
        const double bb = (x2 * stp2 + w * stp1) / dstp;
        v = a + bb;
    }
    src[t] = v * pi4g;
}

// fmg_restore_u((/MPSI/), IU) over the FMG array's extent (interior and one ghost layer: the
// solver's U, [nblk][10^3] ghosted, i fastest) and fmg_psi2g's g on the interior (g dense
// [nblk][3][8^3])
__global__ void k_bm_restore(double *pool, const int *su, int nblk, sg_blockmem L, const double *U, const double *g)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nblk * 1000) return;
    const int b = (int)(t / 1000), c = (int)(t % 1000);
    const int i = c % 10 - 1, j = (c / 10) % 10 - 1, k = c / 100 - 1;
    const long W = 8 + 2 * L.ng, W3 = W * W * W, q = bm_cell(L, i, j, k);
    double *u = pool + (long)su[b] * L.stride;
    u[L.mpsi * W3 + q] = U[t];
    if (i >= 0 && i < 8 && j >= 0 && j < 8 && k >= 0 && k < 8) {
        const double *g0 = g + (long)b * 1536 + (k * 8 + j) * 8 + i;
        u[L.mgx * W3 + q] = g0[0];
        u[L.mgy * W3 + q] = g0[512];
        u[L.mgz * W3 + q] = g0[1024];
    }
}

// source_g_all_level (timestep.F90:528-578) with eos_adiabatic's u2w / w2u
// (eos_adiabatic.F90:226-266), every cell of every grid; the source on the interior only.
// dv[l] = CellWidth(0,l) CellWidth(1,l) CellWidth(2,l), dt[l] = Dtime(l).
__global__ void k_source_g(double *pool, const int *su, const int *s1, const int *level, int nblk, sg_blockmem L,
                           LevelScalars dvl, LevelScalars dtl, double gamma)
{
    const long W = 8 + 2 * L.ng, W3 = W * W * W;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nblk * W3) return;
    const int b = (int)(t / W3);
    const long c = t % W3;
    const int i = (int)(c % W) - L.ng, j = (int)((c / W) % W) - L.ng, k = (int)(c / (W * W)) - L.ng;
    const int l = level[b];
    const double dv = dvl.v[l], dt = dtl.v[l];
    double *u = pool + (long)su[b] * L.stride + c;
    const double rho = u[L.mrho * W3], vx = u[L.mvx * W3], vy = u[L.mvy * W3], vz = u[L.mvz * W3];
    const double p = u[L.mp * W3], ps = u[L.mpsi * W3];
    const double gx = u[L.mgx * W3], gy = u[L.mgy * W3], gz = u[L.mgz * W3];
    // u2w
    const double wr = rho * dv;
    double wx = vx * wr, wy = vy * wr, wz = vz * wr;
    double kin = vx * vx + vy * vy;
    kin = kin + vz * vz;
    double wp = p / (gamma - 1.0) + rho * kin / 2.0;
    wp = dv * wp;
    const double wps = ps * dv;
    const double wgx = gx * rho * dv, wgy = gy * rho * dv, wgz = gz * rho * dv;
    if (i >= 0 && i < 8 && j >= 0 && j < 8 && k >= 0 && k < 8) {
        const double *u1 = pool + (long)s1[b] * L.stride + c;
        const double dtdvrho = dt * dv * u1[L.mrho * W3];
        wx = wx + gx * dtdvrho;
        wy = wy + gy * dtdvrho;
        wz = wz + gz * dtdvrho;
        double work = u1[L.mvx * W3] * gx + u1[L.mvy * W3] * gy;
        work = work + u1[L.mvz * W3] * gz;
        wp = wp + dtdvrho * work;
    }
    // w2u
    const double r2 = wr / dv;
    const double x2 = wx / wr, y2 = wy / wr, z2 = wz / wr;
    double k2 = x2 * x2 + y2 * y2;
    k2 = k2 + z2 * z2;
    u[L.mrho * W3] = r2;
    u[L.mvx * W3] = x2;
    u[L.mvy * W3] = y2;
    u[L.mvz * W3] = z2;
    u[L.mp * W3] = (wp / dv - 0.5 * r2 * k2) * (gamma - 1.0);
    u[L.mpsi * W3] = wps / dv;
    u[L.mgx * W3] = wgx / wr;
    u[L.mgy * W3] = wgy / wr;
    u[L.mgz * W3] = wgz / wr;
}

} // namespace amr
} // namespace sg
