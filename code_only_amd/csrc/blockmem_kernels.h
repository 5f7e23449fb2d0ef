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

// ---- hydro steps around the solve: ghost exchange (boundary_grid), fine-to-coarse restriction
//      (fg2cg_u), refluxing (fluxcorrection).  Faces are 2 d + lr as the reference's
//      encode_surf (grid.F90:461-465). ----

// One direction pass of boundary_grid (grid_boundary.F90:65-105) over items {block, lr, kind |
// parity << 1, source}: kind 0 fix_from_samelev (grid_boundary.F90:238-342), the ghost slab (ng
// layers along d, the whole extent of the other two directions) copied from the same-level
// neighbour's boundary slab, on the first ntab of the current, predictor and past arrays; kind 1
// fix_from_parent (:347-405) on the current array: interp_space's minmod-limited linear
// interpolation (:409-530) from the coarse grid across the face, whose array the step mode picks
// (SINGLE_STEP interp_time, :622-639: past for the predictor step, predictor for the corrector
// step, else current).  same_only skips kind 1 (the parent level's pass).  Grid (items, chunks):
// one thread per cell of a slab per array and component, consecutive threads along the slab's
// contiguous axis (i; for x slabs that is the ng-wide layer, then j); a pass writes only ghost
// slabs and reads only interior slabs along d and coarse cells, so its items are independent.
__device__ __forceinline__ double minmod(double a, double b)
{
    const double s = copysign(1.0, a);
    return s * fmax(0.0, fmin(fabs(a), s * b));
}

__global__ void k_hx_pass(double *pool, const int4 *items, int nitems, const int *su, const int *s1, const int *s2,
                          int ntab, int same_only, int mode, sg_blockmem L, int d)
{
    const int ng = L.ng, W = 8 + 2 * ng;
    const long W3 = (long)W * W * W;
    const int slab = ng * W * W, per_tab = slab * L.ncomp;
    const int it = blockIdx.x;
    int rem = blockIdx.y * blockDim.x + threadIdx.x;
    if (rem >= per_tab * ntab) return;
    const int tab = rem / per_tab;
    rem -= tab * per_tab;
    const int c = rem / slab;
    rem -= c * slab;
    int r, t1, t2;
    if (d == 0) {   // (t2, t1, r): r fastest
        r = rem % ng;
        t1 = (rem / ng) % W - ng;
        t2 = rem / (ng * W) - ng;
    } else {        // (r, t2, t1): t1 (= i) fastest
        t1 = rem % W - ng;
        t2 = (rem / W) % W - ng;
        r = rem / (W * W);
    }
    const int4 e = items[it];
    const int b = e.x, lr = e.y, kind = e.z & 1, par = e.z >> 1;
    if (kind == 1 && (same_only || tab > 0)) return;
    const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
    int p[3];
    p[d] = lr == 0 ? -ng + r : 8 + r;
    p[ta] = t1;
    p[tb] = t2;
    const int *tb_slots = tab == 0 ? su : (tab == 1 ? s1 : s2);
    double *ud = pool + (long)tb_slots[b] * L.stride + c * W3;
    if (kind == 0) {
        int q[3] = {p[0], p[1], p[2]};
        q[d] = lr == 0 ? p[d] + 8 : p[d] - 8;
        const double *us = pool + (long)tb_slots[e.w] * L.stride + c * W3;
        ud[bm_cell(L, p[0], p[1], p[2])] = us[bm_cell(L, q[0], q[1], q[2])];
        return;
    }
    // the coarse buffer Ics..Ice (grid_boundary.F90:182-234) and the fine slab's start, per axis
    int ci[3];
    for (int a = 0; a < 3; ++a) {
        int bs, fs;
        if (a == d) {
            bs = lr == 0 ? 7 - ng / 2 : -ng / 2;
            fs = lr == 0 ? -ng : 8;
        } else {
            bs = ((par >> a) & 1) ? 4 - ng : -ng;
            fs = -ng;
        }
        ci[a] = (p[a] - fs) / 2 + bs + 1;   // IJKC (config.h:249) relative to the buffer
    }
    const int *cs = mode == 1 ? s2 : (mode == 2 ? s1 : su);
    const double *bu = pool + (long)cs[e.w] * L.stride + c * W3;
    const double v = bu[bm_cell(L, ci[0], ci[1], ci[2])];
    double g3[3];
    for (int a = 0; a < 3; ++a) {
        int up[3] = {ci[0], ci[1], ci[2]}, dn[3] = {ci[0], ci[1], ci[2]};
        up[a] += 1;
        dn[a] -= 1;
        const double dr = (bu[bm_cell(L, up[0], up[1], up[2])] - v) / 2.0;
        const double dl = (v - bu[bm_cell(L, dn[0], dn[1], dn[2])]) / 2.0;
        g3[a] = minmod(dr, dl);
    }
    double res = v + g3[0] * ((double)(p[0] & 1) - 0.5);
    res = res + g3[1] * ((double)(p[1] & 1) - 0.5);
    res = res + g3[2] * ((double)(p[2] & 1) - 0.5);
    ud[bm_cell(L, p[0], p[1], p[2])] = res;
}

// the ideal-gas u2w / w2u of one cell (eos_adiabatic.F90:226-266) over the nine labelled
// components, w in the order rho, rho v, E, psi, rho g
__device__ __forceinline__ void u2w9(const double *u, long s, const sg_blockmem &L, double dv, double gamma, double w[9])
{
    const double rho = u[L.mrho * s], vx = u[L.mvx * s], vy = u[L.mvy * s], vz = u[L.mvz * s];
    w[0] = rho * dv;
    w[1] = vx * w[0];
    w[2] = vy * w[0];
    w[3] = vz * w[0];
    double kin = vx * vx + vy * vy;
    kin = kin + vz * vz;
    w[4] = dv * (u[L.mp * s] / (gamma - 1.0) + rho * kin / 2.0);
    w[5] = u[L.mpsi * s] * dv;
    w[6] = u[L.mgx * s] * rho * dv;
    w[7] = u[L.mgy * s] * rho * dv;
    w[8] = u[L.mgz * s] * rho * dv;
}

__device__ __forceinline__ void w2u9(const double w[9], double *u, long s, const sg_blockmem &L, double dv, double gamma)
{
    const double r = w[0] / dv, x = w[1] / w[0], y = w[2] / w[0], z = w[3] / w[0];
    double k2 = x * x + y * y;
    k2 = k2 + z * z;
    u[L.mrho * s] = r;
    u[L.mvx * s] = x;
    u[L.mvy * s] = y;
    u[L.mvz * s] = z;
    u[L.mp * s] = (w[4] / dv - 0.5 * r * k2) * (gamma - 1.0);
    u[L.mpsi * s] = w[5] / dv;
    u[L.mgx * s] = w[6] / w[0];
    u[L.mgy * s] = w[7] / w[0];
    u[L.mgz * s] = w[8] / w[0];
}

__device__ __forceinline__ int label9(const sg_blockmem &L, int q)
{
    int m = L.mrho;
    m = q == 1 ? L.mvx : m;
    m = q == 2 ? L.mvy : m;
    m = q == 3 ? L.mvz : m;
    m = q == 4 ? L.mp : m;
    m = q == 5 ? L.mpsi : m;
    m = q == 6 ? L.mgx : m;
    m = q == 7 ? L.mgy : m;
    m = q == 8 ? L.mgz : m;
    return m;
}

// fg2cg_u (fg2cg.F90:22-194): items {fine block, parent, octant}; one thread per coarse cell of
// the octant: u2w (dv 1) of the eight children, summed in restrict_u's order, w2u (dv 8) into the
// parent
__global__ void k_fg2cg(double *pool, const int4 *items, int nitems, const int *su, sg_blockmem L, double gamma)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nitems * 64) return;
    const int4 e = items[t >> 6];
    const int q = (int)(t & 63), ic = q & 3, jc = (q >> 2) & 3, kc = q >> 4;
    const long W = 8 + 2 * L.ng, W3 = W * W * W;
    const double *uf = pool + (long)su[e.x] * L.stride;
    double *uc = pool + (long)su[e.y] * L.stride;
    const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    double sum[9];
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        double w[9];
        u2w9(uf + bm_cell(L, 2 * ic + off[n][0], 2 * jc + off[n][1], 2 * kc + off[n][2]), W3, L, 1.0, gamma, w);
#pragma unroll
        for (int m = 0; m < 9; ++m) sum[m] = n == 0 ? w[m] : sum[m] + w[m];
    }
    const int o = e.z;
    w2u9(sum, uc + bm_cell(L, ic + (o & 1) * 4, jc + ((o >> 1) & 1) * 4, kc + (o >> 2) * 4), W3, L, 8.0, gamma);
}

// fluxcorrection's com_restrict (reflux.F90:193-295) for one round of items {fine block, face +
// 8 parity, coarse grid, -}: items of a round touch disjoint coarse cells, and rounds keep the
// reference's order (grid by grid, face by face) for cells two faces correct.  One thread per
// coarse cell of the face quarter: restrict_f's 2 x 2 sum (reflux.F90:300-388) times dsf / dsc,
// u2w, w -/+= (that - the coarse boundary flux) dt ds, w2u.  fb[b][kind][face][m][64] (kind 0
// _bnd, 1 _sum).
__global__ void k_reflux_round(double *pool, const int4 *items, int nitems, const int *su, const double *fb,
                               sg_blockmem L, double dsf, double dsc, double dv, double dt, double gamma)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nitems * 16) return;
    const int4 e = items[t >> 4];
    const int c1 = (int)(t & 3), c2 = (int)((t >> 2) & 3);
    const int f = e.y & 7, par = e.y >> 3, d = f >> 1, lr = f & 1;
    const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
    const long W = 8 + 2 * L.ng, W3 = W * W * W;
    int pc[3];
    pc[d] = lr == 0 ? 7 : 0;
    pc[ta] = c1 + ((par >> ta) & 1) * 4;
    pc[tb] = c2 + ((par >> tb) & 1) * 4;
    double *u = pool + (long)su[e.z] * L.stride + bm_cell(L, pc[0], pc[1], pc[2]);
    const double *ff = fb + (((long)e.x * 2 + 1) * 6 + f) * 9 * 64;
    const double *fcb = fb + (((long)e.z * 2 + 0) * 6 + (2 * d + 1 - lr)) * 9 * 64;
    double w[9];
    u2w9(u, W3, L, dv, gamma, w);
    const int s1 = 2 * c1, s2 = 2 * c2;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const int m = label9(L, q);
        const double *a = ff + m * 64;
        double sm = a[s1 + 8 * s2] + a[s1 + 8 * (s2 + 1)];
        sm = sm + a[s1 + 1 + 8 * s2];
        sm = sm + a[s1 + 1 + 8 * (s2 + 1)];
        const double fr = sm * dsf / dsc;
        const double corr = (fr - fcb[m * 64 + pc[ta] + 8 * pc[tb]]) * dt * dsc;
        w[q] = lr == 0 ? w[q] - corr : w[q] + corr;
    }
    w2u9(w, u, W3, L, dv, gamma);
}

} // namespace amr
} // namespace sg
