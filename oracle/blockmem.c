/*
 * ORACLE -- test infrastructure, not product code.
 *
 * CPU FP64 restatement of the steps on either side of the AMR gravity solve that move data
 * between the hydro block memory and the solver, and of the gravity source step that consumes
 * its result, used only as the checker by tests/:
 *
 *   fmg_make_physvar (Poisson part: U <- psi, Rho <- slice(rho) * 4 pi G)
 *                                  fmg_data_if.F90:13-52
 *   fmg_prepare_interp (time slices 0..3)
 *                                  fmg_data_if.F90:169-211
 *   fmg_restore_u((/MPSI/), IU) after fmg_psi2g(IU)
 *                                  fmg.F90:210-215, fmg_data_if.F90:275-294
 *   source_g_all_level             timestep.F90:528-578 (SINGLE_STEP form)
 *   u2w / w2u, ideal gas           eos_adiabatic.F90:226-266
 *
 * Block memory (BlockMem(gid)%u, grid.F90:146-194): per grid one array u(i, j, k, m), i, j, k in
 * [-ng, 8 + ng), m in [0, ncomp), i fastest; the grids in one pool, grid `slot` at
 * pool + slot * stride.  The solver side is the dense form the oracle's orc_amr_solve takes:
 * [nblk][8^3] interiors (i fastest) and g [nblk][3][8^3].  The staging writes interiors only: the
 * solve overwrites the ghost layer of U and PSI before reading it, and reads RHO and SRC in the
 * interior only.
 *
 * Arithmetic keeps the reference's operand order (Fortran evaluates left to right and honours
 * parentheses; x**2 is x*x); this file is built with -ffp-contract=off.
 */
#include <stddef.h>

#include "oracle.h"

#define NB 8

static long cell(const orc_blockmem *L, int i, int j, int k)
{
    const long W = NB + 2 * L->ng;
    return ((long)(k + L->ng) * W + (j + L->ng)) * W + (i + L->ng);
}

static double *comp(double *pool, const orc_blockmem *L, int slot, int m)
{
    const long W = NB + 2 * L->ng;
    return pool + (long)slot * L->stride + (long)m * W * W * W;
}

/* fmg_prepare_interp for one value: u0 current, u1 first-order future (predictor), u2 past */
static double slice_value(int ts, int fine, double u0, double u1, double u2, double stp1, double stp2, double dstp)
{
    if (ts == 0) return u0;
    if (ts == 1) return u1;
    if (ts == 2) return u2;
    if (fine) return u0;   /* Step(amrlev) == Step(LevelMax) */
    const double r = stp1 / dstp;
    const double w = 2 * u1 - u2;
    const double a = (u0 - w) * (r * r);
    const double b = (u2 * stp2 + w * stp1) / dstp;
    return a + b;
}

void orc_bm_stage(const double *pool, const int *slot_u, const int *slot_u1, const int *slot_u2, int nblk,
                  const int *level, const orc_blockmem *L, const orc_timeslice *ts, double pi4g, double *psi,
                  double *src)
{
    for (int b = 0; b < nblk; ++b) {
        const int l = level[b];
        /* fmg_data_if.F90:186-200: stp = Step(LevelMax), dstp = Dstep(l),
         * stp1 = stp - (Step(l) - Dstep(l)), stp2 = Step(l) - stp */
        const double stp = ts->step_max, dstp = ts->dstep[l];
        const double stp1 = stp - (ts->step[l] - ts->dstep[l]);
        const double stp2 = ts->step[l] - stp;
        const int fine = ts->step[l] == ts->step_max;
        const double *up = comp((double *)pool, L, slot_u[b], L->mpsi);
        const double *r0 = comp((double *)pool, L, slot_u[b], L->mrho);
        const double *r1 = (ts->source == 1 || ts->source == 3) ? comp((double *)pool, L, slot_u1[b], L->mrho) : NULL;
        const double *r2 = (ts->source == 2 || ts->source == 3) ? comp((double *)pool, L, slot_u2[b], L->mrho) : NULL;
        for (int k = 0; k < NB; ++k)
            for (int j = 0; j < NB; ++j)
                for (int i = 0; i < NB; ++i) {
                    const long c = cell(L, i, j, k);
                    const size_t q = (size_t)b * 512 + ((size_t)k * NB + j) * NB + i;
                    psi[q] = up[c];   /* timesliceUnknwon = 0 */
                    const double v = slice_value(ts->source, fine, r0[c], r1 ? r1[c] : 0.0, r2 ? r2[c] : 0.0,
                                                 stp1, stp2, dstp);
                    src[q] = v * pi4g;   /* Rho := 4 pi G rho */
                }
    }
}

void orc_bm_restore(double *pool, const int *slot_u, int nblk, const orc_blockmem *L, const double *u_blocks,
                    const double *g)
{
    for (int b = 0; b < nblk; ++b) {
        const double *u = u_blocks + (size_t)b * 1000;
        double *psi = comp(pool, L, slot_u[b], L->mpsi);
        double *gx = comp(pool, L, slot_u[b], L->mgx), *gy = comp(pool, L, slot_u[b], L->mgy);
        double *gz = comp(pool, L, slot_u[b], L->mgz);
        /* fmg_restore_u: the FMG array's whole extent, its ghost layer included */
        for (int k = -1; k <= NB; ++k)
            for (int j = -1; j <= NB; ++j)
                for (int i = -1; i <= NB; ++i) psi[cell(L, i, j, k)] = u[((k + 1) * 10 + (j + 1)) * 10 + (i + 1)];
        /* fmg_psi2g writes the interior of MGX, MGY, MGZ */
        const double *g0 = g + (size_t)b * 3 * 512;
        for (int k = 0; k < NB; ++k)
            for (int j = 0; j < NB; ++j)
                for (int i = 0; i < NB; ++i) {
                    const size_t q = ((size_t)k * NB + j) * NB + i;
                    const long c = cell(L, i, j, k);
                    gx[c] = g0[q];
                    gy[c] = g0[512 + q];
                    gz[c] = g0[1024 + q];
                }
    }
}

void orc_bm_source_g(double *pool, const int *slot_u, const int *slot_u1, int nblk, const int *level,
                     const double *dv_lev, const double *dt_lev, const orc_blockmem *L, double gamma)
{
    const long W = NB + 2 * L->ng, W3 = W * W * W;
    for (int b = 0; b < nblk; ++b) {
        const double dv = dv_lev[level[b]], dt = dt_lev[level[b]];
        double *u = pool + (long)slot_u[b] * L->stride;
        const double *u1 = pool + (long)slot_u1[b] * L->stride;
        for (int k = -L->ng; k < NB + L->ng; ++k)
            for (int j = -L->ng; j < NB + L->ng; ++j)
                for (int i = -L->ng; i < NB + L->ng; ++i) {
                    const long c = cell(L, i, j, k);
                    const double rho = u[L->mrho * W3 + c], vx = u[L->mvx * W3 + c], vy = u[L->mvy * W3 + c];
                    const double vz = u[L->mvz * W3 + c], p = u[L->mp * W3 + c], ps = u[L->mpsi * W3 + c];
                    const double gx = u[L->mgx * W3 + c], gy = u[L->mgy * W3 + c], gz = u[L->mgz * W3 + c];
                    /* u2w (eos_adiabatic.F90:232-242) */
                    const double wr = rho * dv;
                    double wx = vx * wr, wy = vy * wr, wz = vz * wr;
                    double kin = vx * vx + vy * vy;
                    kin = kin + vz * vz;
                    double wp = p / (gamma - 1.0) + rho * kin / 2;
                    wp = dv * wp;
                    const double wps = ps * dv;
                    const double wgx = gx * rho * dv, wgy = gy * rho * dv, wgz = gz * rho * dv;
                    /* the source on the interior (timestep.F90:550-565) */
                    if (i >= 0 && i < NB && j >= 0 && j < NB && k >= 0 && k < NB) {
                        const double dtdvrho = dt * dv * u1[L->mrho * W3 + c];
                        wx = wx + gx * dtdvrho;
                        wy = wy + gy * dtdvrho;
                        wz = wz + gz * dtdvrho;
                        double work = u1[L->mvx * W3 + c] * gx + u1[L->mvy * W3 + c] * gy;
                        work = work + u1[L->mvz * W3 + c] * gz;
                        wp = wp + dtdvrho * work;
                    }
                    /* w2u (eos_adiabatic.F90:254-265) */
                    const double r2 = wr / dv;
                    const double x2 = wx / wr, y2 = wy / wr, z2 = wz / wr;
                    double k2 = x2 * x2 + y2 * y2;
                    k2 = k2 + z2 * z2;
                    const double p2 = (wp / dv - 0.5 * r2 * k2) * (gamma - 1.0);
                    u[L->mrho * W3 + c] = r2;
                    u[L->mvx * W3 + c] = x2;
                    u[L->mvy * W3 + c] = y2;
                    u[L->mvz * W3 + c] = z2;
                    u[L->mp * W3 + c] = p2;
                    u[L->mpsi * W3 + c] = wps / dv;
                    u[L->mgx * W3 + c] = wgx / wr;
                    u[L->mgy * W3 + c] = wgy / wr;
                    u[L->mgz * W3 + c] = wgz / wr;
                }
    }
}
