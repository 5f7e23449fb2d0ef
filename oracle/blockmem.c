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
 *   boundary_grid (SINGLE_STEP)    grid_boundary.F90:65-405, 409-530, 622-639: the hydro ghost
 *                                  exchange, same level and from the parent level
 *   fg2cg_u                        fg2cg.F90:22-194: fine-to-coarse restriction of u
 *   fluxcorrection                 reflux.F90:114-295, 300-388: refluxing at coarse-fine faces
 *
 * Hierarchy tables for the last three are this package's (level, parent, child[8], nbr[6] with
 * faces 2 d + lr -- the same code as the reference's encode_surf, grid.F90:461-483 -- holding the
 * same-level neighbour, else the parent-level one, else -1).  The reference's NeighborGid is the
 * same-level entry; the parent's NeighborGid across a face is nbr[parent] when at the parent's
 * level.  Face-flux buffers (Fb, reflux.F90:24-36) are fb[b][kind][face][m][64], kind 0 the
 * boundary flux (_bnd) and 1 the accumulated one (_sum), the 8 x 8 face values in the Fortran order
 * of the face array (x faces (j, k), y faces (i, k), z faces (i, j), first index fastest).
 * These steps use the ideal-gas u2w / w2u on every component, so they take the nine-component
 * layout only (ncomp 9: rho, v, p, psi, g).
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
#include <math.h>
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


/* ---- boundary_grid ---- */

static const double *arr_c(const double *pool, const orc_blockmem *L, int slot) { return pool + (long)slot * L->stride; }

/* fix_from_samelev (grid_boundary.F90:238-342) of level lev, direction d, on the arrays of the
 * given slot tables: the ghost slab of each face with a same-level neighbour from that
 * neighbour's boundary slab, the full extent of the other two directions (ghosts included) */
static void samelev(double *pool, const int *const *tabs, int ntab, int nblk, const int *level, const int *nbr,
                    const orc_blockmem *L, int lev, int d)
{
    const int ng = L->ng, W = NB + 2 * ng;
    const long W3 = (long)W * W * W;
    for (int b = 0; b < nblk; ++b) {
        if (level[b] != lev) continue;
        for (int lr = 0; lr < 2; ++lr) {
            const int m = nbr[6 * b + 2 * d + lr];
            if (m < 0 || level[m] != lev) continue;
            for (int t = 0; t < ntab; ++t) {
                double *ud = pool + (long)tabs[t][b] * L->stride;
                const double *us = pool + (long)tabs[t][m] * L->stride;
                for (int c = 0; c < L->ncomp; ++c)
                    for (int r = 0; r < ng; ++r)          /* recv index along d, send = recv +/- 8 */
                        for (int t2 = -ng; t2 < NB + ng; ++t2)
                            for (int t1 = -ng; t1 < NB + ng; ++t1) {
                                const int rv = lr == 0 ? -ng + r : NB + r, sv = lr == 0 ? rv + NB : rv - NB;
                                int pr[3], ps[3];
                                pr[d] = rv;
                                ps[d] = sv;
                                pr[d == 0 ? 1 : 0] = ps[d == 0 ? 1 : 0] = t1;
                                pr[d == 2 ? 1 : 2] = ps[d == 2 ? 1 : 2] = t2;
                                ud[c * W3 + cell(L, pr[0], pr[1], pr[2])] = us[c * W3 + cell(L, ps[0], ps[1], ps[2])];
                            }
            }
        }
    }
}

static double minmod(double a, double b)
{
    const double s = copysign(1.0, a);
    return s * fmax(0.0, fmin(fabs(a), s * b));
}

/* fix_from_parent (grid_boundary.F90:347-405) of level lev, direction d: every face without a
 * same-level neighbour whose parent has one gets its ghost slab by interp_space (minmod-limited
 * linear interpolation, grid_boundary.F90:409-530) from the buffer interp_time (SINGLE_STEP,
 * :622-639) takes from that coarse grid: its u2 (predictor), u1 (corrector) or u (complete) over
 * Ics..Ice (grid_boundary.F90:182-234) */
static void from_parent(double *pool, const int *su, const int *s1, const int *s2, int nblk, const int *level,
                        const int *parent, const int *child, const int *nbr, const orc_blockmem *L, int lev, int d,
                        int mode)
{
    const int ng = L->ng, W = NB + 2 * ng;
    const long W3 = (long)W * W * W;
    for (int b = 0; b < nblk; ++b) {
        if (level[b] != lev || lev == 0) continue;
        const int p = parent[b];
        int o = 0;
        while (child[8 * p + o] != b) ++o;
        const int par[3] = {o & 1, (o >> 1) & 1, o >> 2};
        for (int lr = 0; lr < 2; ++lr) {
            const int f = 2 * d + lr, m = nbr[6 * b + f];
            if (m >= 0 && level[m] == lev) continue;
            const int g = nbr[6 * p + f];
            if (g < 0 || level[g] != lev - 1) continue;
            const double *uc = arr_c(pool, L, mode == 1 ? s2[g] : (mode == 2 ? s1[g] : su[g]));
            double *ud = pool + (long)su[b] * L->stride;
            int bs[3], fs[3], fe[3];
            for (int e = 0; e < 3; ++e) {
                if (e == d) {
                    bs[e] = lr == 0 ? NB - 1 - ng / 2 : -ng / 2;   /* buffer Ics..Ice along d */
                    fs[e] = lr == 0 ? -ng : NB;
                    fe[e] = lr == 0 ? -1 : NB - 1 + ng;
                } else {
                    bs[e] = par[e] ? NB / 2 - ng : -ng;
                    fs[e] = -ng;
                    fe[e] = NB - 1 + ng;
                }
            }
            for (int c = 0; c < L->ncomp; ++c) {
                const double *bu = uc + c * W3;
                double *fu = ud + c * W3;
                for (int k = fs[2]; k <= fe[2]; ++k)
                    for (int j = fs[1]; j <= fe[1]; ++j)
                        for (int i = fs[0]; i <= fe[0]; ++i) {
                            const int ic = (i - fs[0]) / 2 + bs[0] + 1, jc = (j - fs[1]) / 2 + bs[1] + 1;
                            const int kc = (k - fs[2]) / 2 + bs[2] + 1;
                            const double v = bu[cell(L, ic, jc, kc)];
                            const double hc = 2.0;
                            const double dxr = (bu[cell(L, ic + 1, jc, kc)] - v) / hc, dxl = (v - bu[cell(L, ic - 1, jc, kc)]) / hc;
                            const double dyr = (bu[cell(L, ic, jc + 1, kc)] - v) / hc, dyl = (v - bu[cell(L, ic, jc - 1, kc)]) / hc;
                            const double dzr = (bu[cell(L, ic, jc, kc + 1)] - v) / hc, dzl = (v - bu[cell(L, ic, jc, kc - 1)]) / hc;
                            const double di = ((i & 1) - 0.5), dj = ((j & 1) - 0.5), dk = ((k & 1) - 0.5);
                            double r = v + minmod(dxr, dxl) * di;
                            r = r + minmod(dyr, dyl) * dj;
                            r = r + minmod(dzr, dzl) * dk;
                            fu[cell(L, i, j, k)] = r;
                        }
            }
        }
    }
}

void orc_bm_grid_boundary(double *pool, const int *su, const int *s1, const int *s2, int nblk, const int *level,
                          const int *parent, const int *child, const int *nbr, const orc_blockmem *L, int lev, int mode)
{
    const int *all[3] = {su, s1, s2};
    if (lev >= 1)   /* the parent level: current, predictor and past arrays, same level only */
        for (int d = 0; d < 3; ++d) samelev(pool, all, 3, nblk, level, nbr, L, lev - 1, d);
    for (int d = 0; d < 3; ++d) {
        from_parent(pool, su, s1, s2, nblk, level, parent, child, nbr, L, lev, d, mode);
        samelev(pool, all, 1, nblk, level, nbr, L, lev, d);
    }
}

/* ---- u2w / w2u on whole cells (ideal gas), the nine components ---- */
static void u2w_cell(const double *u, long s, const orc_blockmem *L, double dv, double gamma, double w[9])
{
    const double rho = u[L->mrho * s], vx = u[L->mvx * s], vy = u[L->mvy * s], vz = u[L->mvz * s];
    w[0] = rho * dv;
    w[1] = vx * w[0];
    w[2] = vy * w[0];
    w[3] = vz * w[0];
    double kin = vx * vx + vy * vy;
    kin = kin + vz * vz;
    w[4] = dv * (u[L->mp * s] / (gamma - 1.0) + rho * kin / 2);
    w[5] = u[L->mpsi * s] * dv;
    w[6] = u[L->mgx * s] * rho * dv;
    w[7] = u[L->mgy * s] * rho * dv;
    w[8] = u[L->mgz * s] * rho * dv;
}

static void w2u_cell(const double w[9], double *u, long s, const orc_blockmem *L, double dv, double gamma)
{
    const double r = w[0] / dv, x = w[1] / w[0], y = w[2] / w[0], z = w[3] / w[0];
    double k2 = x * x + y * y;
    k2 = k2 + z * z;
    u[L->mrho * s] = r;
    u[L->mvx * s] = x;
    u[L->mvy * s] = y;
    u[L->mvz * s] = z;
    u[L->mp * s] = (w[4] / dv - 0.5 * r * k2) * (gamma - 1.0);
    u[L->mpsi * s] = w[5] / dv;
    u[L->mgx * s] = w[6] / w[0];
    u[L->mgy * s] = w[7] / w[0];
    u[L->mgz * s] = w[8] / w[0];
}

/* the nine labels in w's order */
static int lab(const orc_blockmem *L, int q)
{
    const int t[9] = {L->mrho, L->mvx, L->mvy, L->mvz, L->mp, L->mpsi, L->mgx, L->mgy, L->mgz};
    return t[q];
}

/* fg2cg_u (fg2cg.F90:22-194): every grid of level lev restricted into its parent's octant:
 * u2w with dv = 1 on the fine interior, the eight children summed, w2u with dv = 8 */
void orc_bm_fg2cg(double *pool, const int *su, int nblk, const int *level, const int *parent, const int *child,
                  const orc_blockmem *L, int lev, double gamma)
{
    const int W = NB + 2 * L->ng;
    const long W3 = (long)W * W * W;
    for (int b = 0; b < nblk; ++b) {
        if (level[b] != lev || lev == 0) continue;
        const int p = parent[b];
        int o = 0;
        while (child[8 * p + o] != b) ++o;
        const double *uf = pool + (long)su[b] * L->stride;
        double *uc = pool + (long)su[p] * L->stride;
        for (int kc = 0; kc < NB / 2; ++kc)
            for (int jc = 0; jc < NB / 2; ++jc)
                for (int ic = 0; ic < NB / 2; ++ic) {
                    double w[8][9];
                    const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                           {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
                    for (int q = 0; q < 8; ++q)
                        u2w_cell(uf + cell(L, 2 * ic + off[q][0], 2 * jc + off[q][1], 2 * kc + off[q][2]), W3, L, 1.0,
                                 gamma, w[q]);
                    double sum[9];
                    for (int c = 0; c < 9; ++c) {
                        double a = w[0][c] + w[1][c];
                        for (int q = 2; q < 8; ++q) a = a + w[q][c];
                        sum[c] = a;
                    }
                    w2u_cell(sum, uc + cell(L, ic + (o & 1) * NB / 2, jc + ((o >> 1) & 1) * NB / 2, kc + (o >> 2) * NB / 2),
                             W3, L, 8.0, gamma);
                }
    }
}

/* fluxcorrection(level) (reflux.F90:114-124, com_restrict :193-295, restrict_f :300-388): for
 * every grid of level lev + 1 and face 2 d + lr with no same-level neighbour where the parent has
 * one (the coarse grid g): the fine accumulated flux restricted to g's face quarter (the 2 x 2 sum
 * times dsf / dsc), and g's boundary layer there in conserved form corrected by -/+ (that - g's own
 * boundary flux) dt ds (lr left / right), in the order grid by grid, face by face */
void orc_bm_fluxcorrection(double *pool, const int *su, const double *fb, int nblk, const int *level,
                           const int *parent, const int *child, const int *nbr, const double *h_lev,
                           const double *dt_lev, const orc_blockmem *L, int lev, double gamma)
{
    const int W = NB + 2 * L->ng, nm = L->ncomp;
    const long W3 = (long)W * W * W;
    const double hf = h_lev[lev + 1], hc = h_lev[lev];
    const double dsf = hf * hf, dsc = hc * hc, dv = hc * hc * hc, dt = dt_lev[lev];
    for (int b = 0; b < nblk; ++b) {
        if (level[b] != lev + 1) continue;
        const int p = parent[b];
        int o = 0;
        while (child[8 * p + o] != b) ++o;
        const int par[3] = {o & 1, (o >> 1) & 1, o >> 2};
        for (int f = 0; f < 6; ++f) {
            const int d = f >> 1, lr = f & 1;
            const int m = nbr[6 * b + f];
            if (m >= 0 && level[m] == lev + 1) continue;
            const int g = nbr[6 * p + f];
            if (g < 0 || level[g] != lev) continue;
            const int fc_face = 2 * d + (1 - lr);
            const double *ff = fb + (((long)b * 2 + 1) * 6 + f) * nm * 64;
            const double *fcb = fb + (((long)g * 2 + 0) * 6 + fc_face) * nm * 64;
            double *u = pool + (long)su[g] * L->stride;
            const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;   /* tangential axes: t1, t2 */
            for (int c2 = 0; c2 < NB / 2; ++c2)
                for (int c1 = 0; c1 < NB / 2; ++c1) {
                    int pc[3];
                    pc[d] = lr == 0 ? NB - 1 : 0;      /* the coarse grid's layer facing the fine grid */
                    pc[ta] = c1 + par[ta] * NB / 2;
                    pc[tb] = c2 + par[tb] * NB / 2;
                    double w[9];
                    u2w_cell(u + cell(L, pc[0], pc[1], pc[2]), W3, L, dv, gamma, w);
                    for (int q = 0; q < 9; ++q) {
                        const int c = lab(L, q);
                        const double *a = ff + (long)c * 64;
                        const int t1 = 2 * c1, t2 = 2 * c2;
                        double sm = a[t1 + 8 * t2] + a[t1 + 8 * (t2 + 1)];
                        sm = sm + a[t1 + 1 + 8 * t2];
                        sm = sm + a[t1 + 1 + 8 * (t2 + 1)];
                        const double fr = sm * dsf / dsc;
                        const double fcv = fcb[(long)c * 64 + pc[ta] + 8 * pc[tb]];
                        const double corr = (fr - fcv) * dt * dsc;
                        w[q] = lr == 0 ? w[q] - corr : w[q] + corr;
                    }
                    w2u_cell(w, u + cell(L, pc[0], pc[1], pc[2]), W3, L, dv, gamma);
                }
        }
    }
}
