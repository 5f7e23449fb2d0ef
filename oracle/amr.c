/*
 * ORACLE -- test infrastructure, not product code.
 *
 * CPU FP64 restatement of the reference's block-structured AMR multigrid (the cross-level FMG of
 * fmg.F90 with its VMG coarse solve of vmg.F90 and the base-grid MG of mg.F90), used only as the
 * checker by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.  The product path
 * (code_only_amd/csrc) never links, loads or calls anything in this directory.
 *
 * Data model (fmg_data.F90:35-96, 380-427): every block owns, per FMG level L (cells per side
 * n = NB >> L), one array per code (U, RES, RHS, RHO, SRC, PSI) with one ghost layer, index range
 * -1..n per axis, i fastest -- the layout of the uniform oracle -- plus a face-flux array F for
 * the three directions.  Blocks are numbered level by level (coarsest first); within a level the
 * order is the caller's and only fixes the order of sums.  Arrays start at zero for each solve,
 * as fmg_alloc_arr zero-fills them (fmg_data.F90:395,398).
 *
 * Every operation runs over all blocks in block order and touches exactly the cells the
 * reference's loops touch, including ghost edges and corners, in the reference's staged order;
 * arithmetic keeps the reference's left-to-right operand order and this file is built with
 * -ffp-contract=off.
 *
 * Sums: where the reference accumulates one running sum over blocks (multipole moments,
 * fmg_boundary_phys_sphere.F90:132-174; averages, fmg_data.F90:730-769, fmg_data_if.F90:53-78) --
 * an order that depends on the MPI rank count there -- this restatement fixes the order: each
 * block's cells are summed sequentially (k, j, i), then the block partials are added by pairwise
 * halving over the block list zero-padded to a power of two.
 *
 * Reference functions restated here:
 *   fmg_cycle                      fmg.F90:134-233          -> orc_amr_solve
 *   fmg_lin                        fmg.F90:270-377          -> fmg_lin
 *   fmg_rstrct / fmg_interp / fmg_addint / fmg_fill0 / fmg_copy / fmg_add
 *                                  fmg.F90:750-946
 *   fmg_get_h2absmax               fmg.F90:1017-1039
 *   fmg_poisson_flux/relax/resid   fmg_poisson.F90:53-265
 *   fmg_psi2g                      fmg_poisson.F90:4-49
 *   fmg_ghostcell_fix, fmg_ghfix_samelev, fmg_ghfix_parentlev (interp, interpCubic)
 *                                  fmg_ghostcell.F90:43-771
 *   fmg_converge_c2p(_lev)         fmg_converge.F90:74-188
 *   fmg_fluxcorrection             fmg_reflux.F90:28-258
 *   fmg_boundary_fill0 / _extrap   fmg_boundary.F90:15-146
 *   fmg_boundary_physical_grid     fmg_boundary_phys_sphere.F90:80-354 (sphere; periodic: no-op)
 *   vmg_fas and helpers            vmg.F90:116-243, 387-916
 *   vmg_poisson_tau/flux/relax     vmg_poisson.F90:6-145
 *   vmg_interp_p2c_lev             vmg_interpol.F90:100-164
 *   mg_dataPrepare / Restore       mg_data.F90:179-340 (base gather / scatter)
 *   fmg_make_physvar (mean)        fmg_data_if.F90:53-78 (ZERO_AVERAGE_DENSITY)
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define GI(n, i, j, k) ((size_t)((k) + 1) * (size_t)((n) + 2) * (size_t)((n) + 2) + \
                        (size_t)((j) + 1) * (size_t)((n) + 2) + (size_t)((i) + 1))

enum { C_U = 0, C_RES = 1, C_RHS = 2, C_RHO = 3, C_SRC = 4, C_PSI = 5 };
#define C_RUF C_PSI /* vmg.F90:115: FAS keeps R u_f in the PSI slot */
#define NFMG 3      /* fmg_data.F90:115 with 8^3 blocks: FMG levels 0..2 (8^3, 4^3, 2^3) */
#define LEFT 0
#define RIGHT 1

struct orc_amr {
    int nblk, nlev, nb;
    int *level, *ijk, *parent, *child, *nbr, *nkind, *octant, *lev_start, *covered;
    int base_dims[3];
    double h0, x0[3];
    int bc;                 /* ORC_AMR_PERIODIC or ORC_AMR_SPHERE */
    int n[NFMG];
    size_t bl[NFMG];
    double *A[NFMG][ORC_AMR_NCODE];
    double *F[NFMG];        /* [nblk][3][bl] */
    int fill0_level;        /* FmgLevel_fill0 (fmg_data.F90:92) */
    double resh2maxg;       /* Resh2maxg of the last fmg_relax */
    double resmaxg[ORC_AMR_MAXLEV];
    int zero_average;       /* ZERO_AVERAGE_DENSITY */
    long base_calls, base_iters; /* statistics of the base-grid solves */
};

/* ------------------------------------------------------------------------------------------ */
/* small helpers                                                                              */
/* ------------------------------------------------------------------------------------------ */

static inline double *arr(orc_amr *s, int L, int code, int b) { return s->A[L][code] + (size_t)b * s->bl[L]; }
static inline double *flx(orc_amr *s, int L, int b, int d) { return s->F[L] + ((size_t)b * 3 + d) * s->bl[L]; }
static inline double hof(const orc_amr *s, int l, int L) { return ldexp(s->h0, L - l); } /* fmg_data.F90:288 */
static inline int lev_begin(const orc_amr *s, int l) { return s->lev_start[l]; }
static inline int lev_end(const orc_amr *s, int l) { return s->lev_start[l + 1]; }
/* fmg_skip_grid (fmg_data.F90:618-628): a block with children, below the finest level */
static inline int skip(const orc_amr *s, int b) { return s->level[b] != s->nlev - 1 && s->covered[b]; }

/* config.h:249 IJKC with origin 0: floor(i/2). */
static inline int ijkc(int i) { return (i >= 0) ? i / 2 : -((1 - i) / 2); }
/* NEAR(n, nc) = nc + 2*modulo(n,2) - 1 (fmg_ghostcell.F90:409) */
static inline int near_(int f, int c) { return c + 2 * (((f % 2) + 2) % 2) - 1; }

/* Element (a, b, c) of a block array, where (a, b, c) are the coordinates along (d, t1, t2) for a
 * face normal d: t1 < t2 are the two other axes in x, y, z order. */
static inline size_t gi_rot(int n, int d, int a, int t1, int t2)
{
    if (d == 0) return GI(n, a, t1, t2);
    if (d == 1) return GI(n, t1, a, t2);
    return GI(n, t1, t2, a);
}

/* Pairwise halving over a zero-padded power-of-two list (the fixed order for block sums). */
static double tree_sum(double *a, int m)
{
    int p = 1;
    while (p < m) p <<= 1;
    for (int q = m; q < p; ++q) a[q] = 0.0;
    for (int st = p / 2; st >= 1; st /= 2)
        for (int q = 0; q < st; ++q) a[q] = a[q] + a[q + st];
    return p > 0 ? a[0] : 0.0;
}

/* ------------------------------------------------------------------------------------------ */
/* fmg_converge: child -> parent octant                                                       */
/* ------------------------------------------------------------------------------------------ */

static double st_cubic[4][4][4];
static int st_ready = 0;
static void st_init(void)
{
    if (st_ready) return;
    double s1d[4];
    orc_rstrct_cubic_weights(s1d);
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b)
            for (int c = 0; c < 4; ++c) st_cubic[c][b][a] = (s1d[a] * s1d[b]) * s1d[c];
    st_ready = 1;
}

/* rstrct_linear (restriction.F90:127-151) at fine origin (i, j, k). */
static inline double rstr_lin(const double *uf, int n, int i, int j, int k)
{
    double s = uf[GI(n, i, j, k)] + uf[GI(n, i + 1, j, k)];
    s = s + uf[GI(n, i, j + 1, k)];
    s = s + uf[GI(n, i, j, k + 1)];
    s = s + uf[GI(n, i + 1, j + 1, k)];
    s = s + uf[GI(n, i + 1, j, k + 1)];
    s = s + uf[GI(n, i, j + 1, k + 1)];
    s = s + uf[GI(n, i + 1, j + 1, k + 1)];
    return s * (1.0 / 8.0);
}

/* rstrct_cubic (restriction.F90:71-119) at fine origin i1 = 2 ic (stencil i1-1 .. i1+2). */
static inline double rstr_cub(const double *uf, int n, int i1, int j1, int k1)
{
    double s = 0.0;
    int first = 1;
    for (int c = 0; c < 4; ++c)
        for (int b = 0; b < 4; ++b)
            for (int a = 0; a < 4; ++a) {
                double t = uf[GI(n, i1 - 1 + a, j1 - 1 + b, k1 - 1 + c)] * st_cubic[c][b][a];
                s = first ? t : s + t;
                first = 0;
            }
    return s;
}

/* fmg_converge_c2p_lev (fmg_converge.F90:82-187): every block of level lc+1 restricted into its
 * parent's octant (put_buf2uc). */
static void c2p_lev(orc_amr *s, int L, int code, int lc, int cubic)
{
    const int n = s->n[L], h = n / 2;
    if (cubic) st_init();
    for (int b = lev_begin(s, lc + 1); b < lev_end(s, lc + 1); ++b) {
        const double *uf = arr(s, L, code, b);
        double *uc = arr(s, L, code, s->parent[b]);
        int o = s->octant[b];
        int oi = (o & 1) * h, oj = ((o >> 1) & 1) * h, ok = ((o >> 2) & 1) * h;
        for (int kc = 0; kc < h; ++kc)
            for (int jc = 0; jc < h; ++jc)
                for (int ic = 0; ic < h; ++ic)
                    uc[GI(n, oi + ic, oj + jc, ok + kc)] =
                        cubic ? rstr_cub(uf, n, 2 * ic, 2 * jc, 2 * kc) : rstr_lin(uf, n, 2 * ic, 2 * jc, 2 * kc);
    }
}

/* fmg_converge_c2p (fmg_converge.F90:74-81): finest level first. */
static void c2p(orc_amr *s, int L, int code, int cubic)
{
    for (int lc = s->nlev - 2; lc >= 0; --lc) c2p_lev(s, L, code, lc, cubic);
}

/* ------------------------------------------------------------------------------------------ */
/* fmg_ghostcell                                                                              */
/* ------------------------------------------------------------------------------------------ */

/* fmg_ghfix_samelev (fmg_ghostcell.F90:128-197): the whole ghost slab of face (d, lr), tangential
 * ghosts included, from the neighbour's first interior layer. */
static void ghfix_samelev(orc_amr *s, int l, int L, int d, int code)
{
    const int n = s->n[L];
    for (int b = lev_begin(s, l); b < lev_end(s, l); ++b)
        for (int lr = 0; lr < 2; ++lr) {
            int f = 2 * d + lr;
            if (s->nkind[b * 6 + f] != ORC_AMR_NBR_SAME) continue;
            const double *us = arr(s, L, code, s->nbr[b * 6 + f]);
            double *ud = arr(s, L, code, b);
            int ad = lr == LEFT ? -1 : n;         /* Ins/Ine(lr, d, Recv) */
            int as = lr == LEFT ? n - 1 : 0;      /* Ins/Ine(lr, d, Send) */
            for (int t2 = -1; t2 <= n; ++t2)
                for (int t1 = -1; t1 <= n; ++t1) ud[gi_rot(n, d, ad, t1, t2)] = us[gi_rot(n, d, as, t1, t2)];
        }
}

/* fmg_ghfix_parentlev (fmg_ghostcell.F90:317-514, 724-758): a buffer from the parent-level
 * neighbour (its first layer facing this block, tangential half chosen by this block's octant,
 * with that neighbour's tangential ghosts), interpolated tangentially (bilinear (9,3,3,1)/16, or
 * separable cubic when requested and the buffer is at least 4 wide), then the ghost layer
 * (tangential ghosts included) set by the quadratic (10 u0 + 8 uH - 3 uN)/15. */
static void ghfix_parentlev(orc_amr *s, int l, int L, int d, int code, int cubic)
{
    if (l <= 0) return;
    const int n = s->n[L], h = n / 2, w = h + 2; /* buffer: normal 1, tangential -1..h */
    const int use_cubic = cubic && w >= 4;       /* interpCubic_checkSize, MINBLOCKSIZE = 4 */
    double *buf = (double *)malloc(sizeof(double) * (size_t)w * w);
    double *bt = (double *)malloc(sizeof(double) * (size_t)(n + 2) * w);
    const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2; /* tangential axes t1 < t2 */
    for (int b = lev_begin(s, l); b < lev_end(s, l); ++b)
        for (int lr = 0; lr < 2; ++lr) {
            int f = 2 * d + lr;
            if (s->nkind[b * 6 + f] != ORC_AMR_NBR_PARENT) continue;
            const double *us = arr(s, L, code, s->nbr[b * 6 + f]);
            double *ud = arr(s, L, code, b);
            int o = s->octant[b];
            int eo1 = (o >> ta) & 1, eo2 = (o >> tb) & 1; /* fmg_left_or_right parity */
            int an = lr == LEFT ? n - 1 : 0;              /* Ics(lr, d, :) */
            for (int c2 = -1; c2 <= h; ++c2)
                for (int c1 = -1; c1 <= h; ++c1)
                    buf[(c2 + 1) * w + (c1 + 1)] = us[gi_rot(n, d, an, c1 + eo1 * h, c2 + eo2 * h)];
            const int ag = lr == LEFT ? -1 : n;   /* ghost layer */
            const int inner = lr == LEFT ? 1 : -1;
            if (use_cubic) {
                /* first tangential pass into bt(t1 fine, c2 coarse) */
                for (int c2 = -1; c2 <= h; ++c2)
                    for (int t1 = -1; t1 <= n; ++t1) {
                        int idx[4]; double a[4];
                        orc_cubic_stencil(t1, -1, h, idx, a);
                        const double *row = buf + (c2 + 1) * w + 1;
                        double v = a[0] * row[idx[0]] + a[1] * row[idx[1]];
                        v = v + a[2] * row[idx[2]];
                        v = v + a[3] * row[idx[3]];
                        bt[(c2 + 1) * (n + 2) + (t1 + 1)] = v;
                    }
            }
            for (int t2 = -1; t2 <= n; ++t2)
                for (int t1 = -1; t1 <= n; ++t1) {
                    double uH;
                    if (use_cubic) {
                        int idx[4]; double a[4];
                        orc_cubic_stencil(t2, -1, h, idx, a);
                        const double *col = bt + (t1 + 1);
                        uH = a[0] * col[(idx[0] + 1) * (n + 2)] + a[1] * col[(idx[1] + 1) * (n + 2)];
                        uH = uH + a[2] * col[(idx[2] + 1) * (n + 2)];
                        uH = uH + a[3] * col[(idx[3] + 1) * (n + 2)];
                    } else {
                        int c1 = ijkc(t1), c2 = ijkc(t2);
                        int c1n = near_(t1, c1), c2n = near_(t2, c2);
                        double va = buf[(c2 + 1) * w + (c1 + 1)];
                        double vb = buf[(c2n + 1) * w + (c1n + 1)];
                        double vc = buf[(c2n + 1) * w + (c1 + 1)];
                        double vd = buf[(c2 + 1) * w + (c1n + 1)];
                        double S = 9.0 * va + vb;
                        S = S + 3.0 * vc;
                        S = S + 3.0 * vd;
                        uH = S / 16.0;
                    }
                    double u0 = ud[gi_rot(n, d, ag + inner, t1, t2)];
                    double uN = ud[gi_rot(n, d, ag + 2 * inner, t1, t2)];
                    double q = 10.0 * u0 + 8.0 * uH;
                    q = q - 3.0 * uN;
                    ud[gi_rot(n, d, ag, t1, t2)] = q / 15.0;
                }
        }
    free(buf);
    free(bt);
}

/* fmg_ghostcell_fix (fmg_ghostcell.F90:43-55): levels coarse to fine, x, y, z; parent-level faces
 * then same-level faces for each direction. */
static void ghostcell_fix(orc_amr *s, int L, int code, int cubic)
{
    for (int l = 0; l < s->nlev; ++l)
        for (int d = 0; d < 3; ++d) {
            ghfix_parentlev(s, l, L, d, code, cubic);
            ghfix_samelev(s, l, L, d, code);
        }
}

/* ------------------------------------------------------------------------------------------ */
/* fmg_boundary                                                                               */
/* ------------------------------------------------------------------------------------------ */

/* fmg_boundary_fill0 / vmg_boundary_fill0 (fmg_boundary.F90:15-53, vmg.F90:746-778) for the
 * blocks of levels [l0, l1). */
static void fill0_faces(orc_amr *s, int L, int code, int l0, int l1)
{
    const int n = s->n[L];
    for (int b = lev_begin(s, l0); b < lev_end(s, l1 - 1); ++b) {
        double *u = arr(s, L, code, b);
        for (int f = 0; f < 6; ++f) {
            if (s->nkind[b * 6 + f] != ORC_AMR_NBR_NONE) continue;
            int d = f / 2, a = (f & 1) ? n : -1;
            for (int t2 = -1; t2 <= n; ++t2)
                for (int t1 = -1; t1 <= n; ++t1) u[gi_rot(n, d, a, t1, t2)] = 0.0;
        }
    }
}

/* fmg_boundary_extrap / vmg_boundary_extrap (fmg_boundary.F90:87-146, vmg.F90:782-818). */
static void extrap_faces(orc_amr *s, int L, int code, int l0, int l1)
{
    const int n = s->n[L];
    for (int b = lev_begin(s, l0); b < lev_end(s, l1 - 1); ++b) {
        double *u = arr(s, L, code, b);
        for (int f = 0; f < 6; ++f) {
            if (s->nkind[b * 6 + f] != ORC_AMR_NBR_NONE) continue;
            int d = f / 2, a = (f & 1) ? n : -1, a1 = (f & 1) ? n - 1 : 0;
            for (int t2 = -1; t2 <= n; ++t2)
                for (int t1 = -1; t1 <= n; ++t1) u[gi_rot(n, d, a, t1, t2)] = -u[gi_rot(n, d, a1, t1, t2)];
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* fmg_poisson: flux, reflux, relax, resid                                                    */
/* ------------------------------------------------------------------------------------------ */

/* fmg_reflux_do (fmg_reflux.F90:139-258): each face of a level >= 1 block whose neighbour is at
 * the parent level hands the 2x2 average of its own face fluxes to the matching quarter of that
 * neighbour's face fluxes. */
static void fluxcorrection(orc_amr *s, int L)
{
    const int n = s->n[L], h = n / 2;
    for (int l = 1; l < s->nlev; ++l)
        for (int b = lev_begin(s, l); b < lev_end(s, l); ++b)
            for (int d = 0; d < 3; ++d)
                for (int lr = 0; lr < 2; ++lr) {
                    int f = 2 * d + lr;
                    if (s->nkind[b * 6 + f] != ORC_AMR_NBR_PARENT) continue;
                    const double *ff = flx(s, L, b, d);
                    double *fc = flx(s, L, s->nbr[b * 6 + f], d);
                    const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
                    int o = s->octant[b];
                    int off1 = ((o >> ta) & 1) * h, off2 = ((o >> tb) & 1) * h;
                    int af = lr == LEFT ? -1 : n - 1;   /* Ifs(lr, d) */
                    int ac = lr == LEFT ? n - 1 : -1;   /* Ics(lr, d, :) */
                    for (int c2 = 0; c2 < h; ++c2)
                        for (int c1 = 0; c1 < h; ++c1) {
                            int f1 = 2 * c1, f2 = 2 * c2;
                            double v = ff[gi_rot(n, d, af, f1, f2)] + ff[gi_rot(n, d, af, f1, f2 + 1)];
                            v = v + ff[gi_rot(n, d, af, f1 + 1, f2)];
                            v = v + ff[gi_rot(n, d, af, f1 + 1, f2 + 1)];
                            fc[gi_rot(n, d, ac, off1 + c1, off2 + c2)] = v * 0.25;
                        }
                }
}

static void block_flux(orc_amr *s, int L, int b, const double *u, double hi)
{
    const int n = s->n[L];
    double *fx = flx(s, L, b, 0), *fy = flx(s, L, b, 1), *fz = flx(s, L, b, 2);
    for (int k = -1; k < n; ++k)
        for (int j = -1; j < n; ++j)
            for (int i = -1; i < n; ++i) {
                size_t q = GI(n, i, j, k);
                fx[q] = (u[GI(n, i + 1, j, k)] - u[q]) * hi;
                fy[q] = (u[GI(n, i, j + 1, k)] - u[q]) * hi;
                fz[q] = (u[GI(n, i, j, k + 1)] - u[q]) * hi;
            }
}

/* fmg_poisson_flux (fmg_poisson.F90:53-124). */
static void poisson_flux(orc_amr *s, int L, int ju, int fill0flag)
{
    if (s->fill0_level != L) {
        c2p(s, L, ju, 0);
        ghostcell_fix(s, L, ju, 0);
        if (fill0flag) fill0_faces(s, L, ju, 0, s->nlev);
    }
    s->fill0_level = -1;
    for (int l = 0; l < s->nlev; ++l) {
        double hi = 1.0 / hof(s, l, L);
        for (int b = lev_begin(s, l); b < lev_end(s, l); ++b)
            if (!skip(s, b)) block_flux(s, L, b, arr(s, L, ju, b), hi);
    }
    fluxcorrection(s, L);
}

/* sum of flux differences at (i, j, k) in the reference's operand order */
static inline double divf(const double *fx, const double *fy, const double *fz, int n, int i, int j, int k)
{
    double v = fx[GI(n, i, j, k)] - fx[GI(n, i - 1, j, k)];
    v = v + fy[GI(n, i, j, k)];
    v = v - fy[GI(n, i, j - 1, k)];
    v = v + fz[GI(n, i, j, k)];
    v = v - fz[GI(n, i, j, k - 1)];
    return v;
}

/* fmg_poisson_relax (fmg_poisson.F90:130-213). */
static void poisson_relax(orc_amr *s, int L, int ju, int jrhs, int fill0flag)
{
    const int n = s->n[L];
    const double sixth = 1.0 / 6.0;
    double resh2max = 0.0;
    for (int ipass = 1; ipass <= 2; ++ipass) {
        poisson_flux(s, L, ju, fill0flag);
        for (int l = 0; l < s->nlev; ++l) {
            double h = hof(s, l, L);
            for (int b = lev_begin(s, l); b < lev_end(s, l); ++b) {
                if (skip(s, b)) continue;
                double *u = arr(s, L, ju, b);
                const double *rhs = arr(s, L, jrhs, b);
                const double *fx = flx(s, L, b, 0), *fy = flx(s, L, b, 1), *fz = flx(s, L, b, 2);
                for (int k = 0; k < n; ++k)
                    for (int j = 0; j < n; ++j)
                        for (int i = (j + k + ipass) % 2; i < n; i += 2) {
                            double r = divf(fx, fy, fz, n, i, j, k) - rhs[GI(n, i, j, k)] * h;
                            double resh2 = r * h;
                            u[GI(n, i, j, k)] = u[GI(n, i, j, k)] + sixth * resh2;
                            double a = fabs(resh2);
                            if (a > resh2max) resh2max = a;
                        }
            }
        }
    }
    s->resh2maxg = resh2max;
}

/* fmg_poisson_resid (fmg_poisson.F90:217-265). */
static void poisson_resid(orc_amr *s, int L, int jres, int ju, int jrhs, int fill0flag)
{
    const int n = s->n[L];
    poisson_flux(s, L, ju, fill0flag);
    for (int l = 0; l < s->nlev; ++l) {
        double hi = 1.0 / hof(s, l, L);
        for (int b = lev_begin(s, l); b < lev_end(s, l); ++b) {
            if (skip(s, b)) continue;
            double *res = arr(s, L, jres, b);
            const double *rhs = arr(s, L, jrhs, b);
            const double *fx = flx(s, L, b, 0), *fy = flx(s, L, b, 1), *fz = flx(s, L, b, 2);
            for (int k = 0; k < n; ++k)
                for (int j = 0; j < n; ++j)
                    for (int i = 0; i < n; ++i)
                        res[GI(n, i, j, k)] = rhs[GI(n, i, j, k)] - divf(fx, fy, fz, n, i, j, k) * hi;
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* fmg.F90 helpers                                                                            */
/* ------------------------------------------------------------------------------------------ */

/* fmg_rstrct (fmg.F90:750-778), linear: non-skipped blocks, level Lc from Lc-1. */
static void fmg_rstrct(orc_amr *s, int Lc, int iuc, int iuf)
{
    for (int b = 0; b < s->nblk; ++b) {
        if (skip(s, b)) continue;
        orc_rstrct_linear(arr(s, Lc, iuc, b), arr(s, Lc - 1, iuf, b), s->n[Lc]);
    }
}

/* fmg_interp (fmg.F90:786-828) with fmg_interp_cubic / fmg_interp_linear (fmg.F90:832-846,
 * fmg_interpol_cubic.F90:84-115). */
static void fmg_interp(orc_amr *s, int Lf, int iuf, int iuc, int cubic)
{
    const int Lc = Lf + 1;
    c2p(s, Lc, iuc, cubic);
    ghostcell_fix(s, Lc, iuc, cubic);
    extrap_faces(s, Lc, iuc, 0, s->nlev);
    for (int b = 0; b < s->nblk; ++b) {
        if (skip(s, b)) continue;
        if (cubic) orc_interp_tricubic(arr(s, Lf, iuf, b), arr(s, Lc, iuc, b), s->n[Lf]);
        else orc_interp_trilinear(arr(s, Lf, iuf, b), arr(s, Lc, iuc, b), s->n[Lf]);
    }
    fill0_faces(s, Lc, iuc, 0, s->nlev);
    fill0_faces(s, Lf, iuf, 0, s->nlev);
}

/* fmg_fill0 / fmg_copy / fmg_add (fmg.F90:850-946): whole arrays (ghosts included). */
static void fmg_fill0(orc_amr *s, int L, int code)
{
    for (int b = 0; b < s->nblk; ++b)
        if (!skip(s, b)) memset(arr(s, L, code, b), 0, sizeof(double) * s->bl[L]);
}

static void fmg_copy(orc_amr *s, int L, int out, int in, int noskip)
{
    for (int b = 0; b < s->nblk; ++b)
        if (noskip || !skip(s, b)) memcpy(arr(s, L, out, b), arr(s, L, in, b), sizeof(double) * s->bl[L]);
}

static void fmg_add(orc_amr *s, int L, int a, int da)
{
    for (int b = 0; b < s->nblk; ++b) {
        if (skip(s, b)) continue;
        double *x = arr(s, L, a, b);
        const double *y = arr(s, L, da, b);
        for (size_t q = 0; q < s->bl[L]; ++q) x[q] = x[q] + y[q];
    }
}

/* fmg_addint (fmg.F90:942-946). */
static void fmg_addint(orc_amr *s, int Lf, int juf, int juc, int jres)
{
    fmg_interp(s, Lf, jres, juc, 0);
    fmg_add(s, Lf, juf, jres);
}

/* fmg_get_h2absmax (fmg.F90:1017-1039). */
static double h2absmax(orc_amr *s, int L, int code)
{
    const int n = s->n[L];
    double m = 0.0;
    for (int l = 0; l < s->nlev; ++l) {
        double hh = hof(s, l, L);
        double h2 = hh * hh;
        for (int b = lev_begin(s, l); b < lev_end(s, l); ++b) {
            if (skip(s, b)) continue;
            const double *a = arr(s, L, code, b);
            double bm = 0.0;
            for (int k = 0; k < n; ++k)
                for (int j = 0; j < n; ++j)
                    for (int i = 0; i < n; ++i) {
                        double v = fabs(a[GI(n, i, j, k)]);
                        if (v > bm) bm = v;
                    }
            double t = bm * h2;
            if (t > m) m = t;
        }
    }
    return m;
}

/* ------------------------------------------------------------------------------------------ */
/* vmg.F90: FAS V-cycle over AMR levels at the coarsest FMG level                             */
/* ------------------------------------------------------------------------------------------ */

static void lev_copy(orc_amr *s, int L, int l, int out, int in)
{
    size_t off = (size_t)lev_begin(s, l) * s->bl[L];
    size_t len = (size_t)(lev_end(s, l) - lev_begin(s, l)) * s->bl[L];
    memcpy(s->A[L][out] + off, s->A[L][in] + off, sizeof(double) * len);
}

/* vmg_ghostcell (vmg.F90:728-742): same-level faces x, y, z, then zero on physical faces. */
static void vmg_ghostcell(orc_amr *s, int L, int l, int code)
{
    for (int d = 0; d < 3; ++d) ghfix_samelev(s, l, L, d, code);
    fill0_faces(s, L, code, l, l + 1);
}

/* vmg_gridboundary (vmg.F90:708-724): parent-level faces x, y, z (linear), then physical zero. */
static void vmg_gridboundary(orc_amr *s, int L, int l, int code)
{
    for (int d = 0; d < 3; ++d) ghfix_parentlev(s, l, L, d, code, 0);
    fill0_faces(s, L, code, l, l + 1);
}

/* vmg_poisson_flux (vmg_poisson.F90:59-84): all blocks of the level, no reflux. */
static void vmg_flux(orc_amr *s, int L, int l, int ju)
{
    vmg_ghostcell(s, L, l, ju);
    double hi = 1.0 / hof(s, l, L);
    for (int b = lev_begin(s, l); b < lev_end(s, l); ++b) block_flux(s, L, b, arr(s, L, ju, b), hi);
}

/* vmg_poisson_relax (vmg_poisson.F90:88-145): every block of the level, covered ones included. */
static void vmg_relax(orc_amr *s, int L, int l, int ju, int jrhs)
{
    const int n = s->n[L];
    const double sixth = 1.0 / 6.0;
    double resmax = 0.0, h = hof(s, l, L);
    for (int ipass = 1; ipass <= 2; ++ipass) {
        vmg_flux(s, L, l, ju);
        for (int b = lev_begin(s, l); b < lev_end(s, l); ++b) {
            double *u = arr(s, L, ju, b);
            const double *rhs = arr(s, L, jrhs, b);
            const double *fx = flx(s, L, b, 0), *fy = flx(s, L, b, 1), *fz = flx(s, L, b, 2);
            for (int k = 0; k < n; ++k)
                for (int j = 0; j < n; ++j)
                    for (int i = (j + k + ipass) % 2; i < n; i += 2) {
                        double r = divf(fx, fy, fz, n, i, j, k) - rhs[GI(n, i, j, k)] * h;
                        double resh2 = r * h;
                        u[GI(n, i, j, k)] = u[GI(n, i, j, k)] + sixth * resh2;
                        double a = fabs(resh2);
                        if (a > resmax) resmax = a;
                    }
        }
    }
    s->resmaxg[l] = resmax / (h * h);
}

/* vmg_poisson_tau (vmg_poisson.F90:6-55) then vmg_tau_crop (vmg.F90:462-506): tau at level lc. */
static void vmg_tau(orc_amr *s, int L, int lc, int jtau, int ju)
{
    const int n = s->n[L], lf = lc + 1;
    if (lf > s->nlev - 1) return;
    vmg_flux(s, L, lf, ju);
    double hi = 1.0 / hof(s, lf, L);
    for (int b = lev_begin(s, lf); b < lev_end(s, lf); ++b) {
        double *tau = arr(s, L, jtau, b);
        const double *fx = flx(s, L, b, 0), *fy = flx(s, L, b, 1), *fz = flx(s, L, b, 2);
        memset(tau, 0, sizeof(double) * s->bl[L]);
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) tau[GI(n, i, j, k)] = -(divf(fx, fy, fz, n, i, j, k) * hi);
    }
    c2p_lev(s, L, jtau, lc, 0);
    vmg_flux(s, L, lc, ju);
    hi = 1.0 / hof(s, lc, L);
    for (int b = lev_begin(s, lc); b < lev_end(s, lc); ++b) {
        if (!s->covered[b]) continue;
        double *tau = arr(s, L, jtau, b);
        const double *fx = flx(s, L, b, 0), *fy = flx(s, L, b, 1), *fz = flx(s, L, b, 2);
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    tau[GI(n, i, j, k)] = tau[GI(n, i, j, k)] + divf(fx, fy, fz, n, i, j, k) * hi;
    }
    /* crop: a child face that meets the parent level zeroes the parent's boundary plane there */
    for (int b = lev_begin(s, lf); b < lev_end(s, lf); ++b)
        for (int f = 0; f < 6; ++f) {
            if (s->nkind[b * 6 + f] != ORC_AMR_NBR_PARENT) continue;
            double *tau = arr(s, L, jtau, s->parent[b]);
            int d = f / 2, a = (f & 1) ? n - 1 : 0;
            for (int t2 = -1; t2 <= n; ++t2)
                for (int t1 = -1; t1 <= n; ++t1) tau[gi_rot(n, d, a, t1, t2)] = 0.0;
        }
}

/* vmg_interp (vmg.F90:592-600) with vmg_interp_p2c_lev (vmg_interpol.F90:100-164), trilinear:
 * level l-1 ghosts refreshed, then each child's interior from its parent's octant with ghosts. */
static void vmg_interp(orc_amr *s, int L, int l, int juf, int juc)
{
    const int n = s->n[L], h = n / 2, w = h + 2;
    vmg_ghostcell(s, L, l - 1, juc);
    extrap_faces(s, L, juc, l - 1, l);
    double *buf = (double *)malloc(sizeof(double) * (size_t)w * w * w);
    for (int b = lev_begin(s, l); b < lev_end(s, l); ++b) {
        const double *uc = arr(s, L, juc, s->parent[b]);
        int o = s->octant[b];
        int oi = (o & 1) * h, oj = ((o >> 1) & 1) * h, ok = ((o >> 2) & 1) * h;
        for (int k = -1; k <= h; ++k)
            for (int j = -1; j <= h; ++j)
                for (int i = -1; i <= h; ++i) buf[GI(h, i, j, k)] = uc[GI(n, oi + i, oj + j, ok + k)];
        orc_interp_trilinear(arr(s, L, juf, b), buf, n);
    }
    free(buf);
    fill0_faces(s, L, juc, l - 1, l);
}

/* vmg_zeroAverage (vmg.F90:890-915) with fmg_ave (fmg_data.F90:730-769) on the base level. */
static void vmg_zero_average(orc_amr *s, int L, int code)
{
    const int n = s->n[L];
    int nb0 = lev_end(s, 0) - lev_begin(s, 0);
    double *part = (double *)malloc(sizeof(double) * (size_t)(2 * nb0 + 1));
    for (int b = 0; b < nb0; ++b) {
        const double *a = arr(s, L, code, lev_begin(s, 0) + b);
        double t = 0.0;
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) t = t + a[GI(n, i, j, k)];
        part[b] = t;
    }
    double ave = tree_sum(part, nb0) / (double)(n * n * n * nb0);
    free(part);
    for (int b = lev_begin(s, 0); b < lev_end(s, 0); ++b) {
        double *a = arr(s, L, code, b);
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) a[GI(n, i, j, k)] = a[GI(n, i, j, k)] - ave;
    }
}

/* vmg_slvsml (vmg.F90:387-432): base blocks gathered into one uniform grid (mg_dataPrepare),
 * solved by mg_cycle with the reference's own stopping test, scattered back (mg_dataRestore);
 * without a power-of-two base, 10 base-level relaxations instead. */
static void vmg_slvsml(orc_amr *s, int L, int ju, int jrhs)
{
    const int n = s->n[L];
    const int nx = s->base_dims[0], ny = s->base_dims[1], nz = s->base_dims[2];
    int pow2 = (nx & (nx - 1)) == 0 && (ny & (ny - 1)) == 0 && (nz & (nz - 1)) == 0;
    if (!pow2 || nx != ny || ny != nz) {
        for (int it = 0; it < 10; ++it) vmg_relax(s, L, 0, ju, jrhs);
        return;
    }
    const int N = nx * n;
    size_t len = orc_ghosted_len(N);
    double *psi = (double *)calloc(len, sizeof(double)), *src = (double *)calloc(len, sizeof(double));
    for (int b = lev_begin(s, 0); b < lev_end(s, 0); ++b) {
        const int *p = s->ijk + 3 * b;
        const double *u = arr(s, L, ju, b), *r = arr(s, L, jrhs, b);
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    size_t q = GI(N, p[0] * n + i, p[1] * n + j, p[2] * n + k);
                    psi[q] = u[GI(n, i, j, k)];
                    src[q] = r[GI(n, i, j, k)];
                }
    }
    /* mg.F90:50-52: errormax is the single-precision literal 1.e-2 widened to double */
    orc_params prm = {(double)1.e-2f, 10, ORC_STOP_REFERENCE, 2, 2, 2};
    orc_stats st;
    orc_set_bc(s->bc == ORC_AMR_SPHERE ? ORC_BC_DIRICHLET : ORC_BC_PERIODIC);
    orc_mg_cycle(psi, src, N, hof(s, 0, L), &prm, &st);
    orc_set_bc(ORC_BC_PERIODIC);
    s->base_calls += 1;
    s->base_iters += st.iterations;
    for (int b = lev_begin(s, 0); b < lev_end(s, 0); ++b) {
        const int *p = s->ijk + 3 * b;
        double *u = arr(s, L, ju, b);
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) u[GI(n, i, j, k)] = psi[GI(N, p[0] * n + i, p[1] * n + j, p[2] * n + k)];
    }
    free(psi);
    free(src);
}

/* vmg_fas (vmg.F90:116-243), linear Poisson: Ncycle = 1, Npre = Npost = 2. */
static void vmg_fas(orc_amr *s, int jrho)
{
    const int L = NFMG - 1;
    if (jrho != C_RHO)
        for (int l = 0; l < s->nlev; ++l) lev_copy(s, L, l, C_RHO, jrho);
    c2p(s, L, C_U, 0);
    c2p(s, L, C_RHO, 0);
    fill0_faces(s, L, C_U, 0, s->nlev);
    for (int l = 0; l < s->nlev; ++l) lev_copy(s, L, l, C_RHS, C_RHO);
    for (int v = s->nlev - 1; v >= 1; --v) {
        vmg_gridboundary(s, L, v, C_U);
        for (int p = 0; p < 2; ++p) vmg_relax(s, L, v, C_U, C_RHS);
        /* vmg_rstrct(v-1, U, cubic): this level's ghosts (parent- then same-level per direction,
         * parent faces with the cubic request), then cubic restriction into the parents */
        for (int d = 0; d < 3; ++d) {
            ghfix_parentlev(s, v, L, d, C_U, 1);
            ghfix_samelev(s, v, L, d, C_U);
        }
        c2p_lev(s, L, C_U, v - 1, 1);
        lev_copy(s, L, v - 1, C_RUF, C_U);
        c2p_lev(s, L, C_RHS, v - 1, 0);
        vmg_tau(s, L, v - 1, C_RES, C_U);
        for (int b = lev_begin(s, v - 1); b < lev_end(s, v - 1); ++b) {   /* vmg_add_onov */
            if (!s->covered[b]) continue;
            double *a = arr(s, L, C_RHS, b);
            const double *da = arr(s, L, C_RES, b);
            for (size_t q = 0; q < s->bl[L]; ++q) a[q] = a[q] + da[q];
        }
    }
    if (s->zero_average) vmg_zero_average(s, L, C_RHS);
    vmg_slvsml(s, L, C_U, C_RHS);
    for (int v = 1; v < s->nlev; ++v) {
        for (int b = lev_begin(s, v - 1); b < lev_end(s, v - 1); ++b) {   /* vmg_sub */
            double *o = arr(s, L, C_RES, b);
            const double *a = arr(s, L, C_U, b), *c = arr(s, L, C_RUF, b);
            for (size_t q = 0; q < s->bl[L]; ++q) o[q] = a[q] - c[q];
        }
        vmg_interp(s, L, v, C_RES, C_RES);
        for (int b = lev_begin(s, v); b < lev_end(s, v); ++b) {           /* vmg_add */
            double *a = arr(s, L, C_U, b);
            const double *da = arr(s, L, C_RES, b);
            for (size_t q = 0; q < s->bl[L]; ++q) a[q] = a[q] + da[q];
        }
        vmg_gridboundary(s, L, v, C_U);
        for (int p = 0; p < 2; ++p) vmg_relax(s, L, v, C_U, C_RHS);
    }
}

/* ------------------------------------------------------------------------------------------ */
/* fmg_lin / fmg_cycle                                                                        */
/* ------------------------------------------------------------------------------------------ */

/* fmg_lin (fmg.F90:270-377): Ncycle = 2, Npre = Npost = 2, zero physical ghosts. */
static void fmg_lin(orc_amr *s)
{
    const int Lmax = NFMG - 1;
    for (int L = 1; L <= Lmax; ++L) fmg_rstrct(s, L, C_RHO, C_RHO);
    fmg_fill0(s, Lmax, C_U);
    vmg_fas(s, C_RHO);
    for (int Lf = Lmax - 1; Lf >= 0; --Lf) {
        fmg_interp(s, Lf, C_U, C_U, 1);
        fmg_copy(s, Lf, C_RHS, C_RHO, 0);
        for (int jc = 0; jc < 2; ++jc) {
            for (int v = Lf; v <= Lmax - 1; ++v) {
                for (int p = 0; p < 2; ++p) poisson_relax(s, v, C_U, C_RHS, 1);
                poisson_resid(s, v, C_RES, C_U, C_RHS, 1);
                fmg_rstrct(s, v + 1, C_RHS, C_RES);
                fmg_fill0(s, v + 1, C_U);
                s->fill0_level = v + 1;
            }
            vmg_fas(s, C_RHS);
            for (int v = Lmax - 1; v >= Lf; --v) {
                fmg_addint(s, v, C_U, C_U, C_RES);
                for (int p = 0; p < 2; ++p) poisson_relax(s, v, C_U, C_RHS, 1);
            }
        }
    }
    c2p(s, 0, C_U, 0);
}

/* Cell-centre coordinate along axis a of cell index i (-1..n) in block b at FMG level 0. */
static inline double coord(const orc_amr *s, int b, int a, int i)
{
    int g = s->ijk[3 * b + a] * s->nb + i;
    return s->x0[a] + ((double)g + 0.5) * hof(s, s->level[b], 0);
}

/* fmg_boundary_physical_grid (fmg_boundary_phys_sphere.F90:80-354): mask the source outside the
 * inscribed sphere, take the l <= 3 multipole moments of the base level, set the ghost layer of
 * every physical face of PSI to the multipole potential there. */
void orc_amr_sphere_bc(orc_amr *s, int ju, int jrho)
{
    const int n = s->nb;
    const int NX = s->base_dims[0] * n, NY = s->base_dims[1] * n, NZ = s->base_dims[2] * n;
    double xmin = s->x0[0] + 0.5 * s->h0, ymin = s->x0[1] + 0.5 * s->h0, zmin = s->x0[2] + 0.5 * s->h0;
    double xmax = s->x0[0] + ((double)(NX - 1) + 0.5) * s->h0;
    double ymax = s->x0[1] + ((double)(NY - 1) + 0.5) * s->h0;
    double zmax = s->x0[2] + ((double)(NZ - 1) + 0.5) * s->h0;
    double m = xmax;
    if (ymax < m) m = ymax;
    if (zmax < m) m = zmax;
    if (-xmin < m) m = -xmin;
    if (-ymin < m) m = -ymin;
    if (-zmin < m) m = -zmin;
    const double rad2 = m * m;
    for (int b = 0; b < s->nblk; ++b) {
        double *rho = arr(s, 0, jrho, b);
        for (int k = -1; k <= n; ++k)
            for (int j = -1; j <= n; ++j)
                for (int i = -1; i <= n; ++i) {
                    double x = coord(s, b, 0, i), y = coord(s, b, 1, j), z = coord(s, b, 2, k);
                    if (x * x + y * y + z * z >= rad2) rho[GI(n, i, j, k)] = 0.0;
                }
    }
    /* moments over the base level */
    const int nb0 = lev_end(s, 0) - lev_begin(s, 0);
    double *part = (double *)malloc(sizeof(double) * 16 * (size_t)(2 * nb0 + 1));
    const double dv = (s->h0 * s->h0) * s->h0;
    for (int bb = 0; bb < nb0; ++bb) {
        int b = lev_begin(s, 0) + bb;
        const double *rho = arr(s, 0, jrho, b);
        double a[16];
        for (int q = 0; q < 16; ++q) a[q] = 0.0;
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    double x = coord(s, b, 0, i), y = coord(s, b, 1, j), z = coord(s, b, 2, k);
                    double r2 = x * x + y * y + z * z;
                    double r = sqrt(r2);
                    double r3 = r2 * r;
                    double rc = sqrt(x * x + y * y);
                    double ri = 1.0 / r, rci = 1.0 / rc;
                    double cost = z * ri, sint = rc * ri, cosp = x * rci, sinp = y * rci;
                    double c2 = cost * cost, s2 = sint * sint, s3 = s2 * sint;
                    double cp2 = cosp * cosp, sp2 = sinp * sinp;
                    double rdv = rho[GI(n, i, j, k)] * dv;
                    a[0] = a[0] + rdv;
                    a[1] = a[1] + rdv * r * cost;
                    a[2] = a[2] + rdv * r * sint * cosp;
                    a[3] = a[3] + rdv * r * sint * sinp;
                    a[4] = a[4] + rdv * r2 * (3.0 * c2 - 1.0);
                    a[5] = a[5] + rdv * r2 * sint * cost * cosp;
                    a[6] = a[6] + rdv * r2 * sint * cost * sinp;
                    a[7] = a[7] + rdv * r2 * s2 * (cp2 - sp2);
                    a[8] = a[8] + rdv * r2 * s2 * 2.0 * sinp * cosp;
                    a[9] = a[9] + rdv * r3 * cost * (5.0 * c2 - 3.0);
                    a[10] = a[10] + rdv * r3 * sint * (5.0 * c2 - 1.0) * cosp;
                    a[11] = a[11] + rdv * r3 * sint * (5.0 * c2 - 1.0) * sinp;
                    a[12] = a[12] + rdv * r3 * cost * s2 * (cp2 - sp2);
                    a[13] = a[13] + rdv * r3 * cost * s2 * 2.0 * sinp * cosp;
                    a[14] = a[14] + rdv * r3 * s3 * cosp * (cp2 - 3.0 * sp2);
                    a[15] = a[15] + rdv * r3 * s3 * sinp * (3.0 * cp2 - sp2);
                }
        for (int q = 0; q < 16; ++q) part[(size_t)q * (2 * nb0 + 1) + bb] = a[q];
    }
    double A[16];
    for (int q = 0; q < 16; ++q) A[q] = tree_sum(part + (size_t)q * (2 * nb0 + 1), nb0);
    free(part);
    const double pi = 3.14159265358979323846;   /* 4 atan(1) rounded to double */
    /* fmg_boundary_phys_sphere.F90:213-228 normalisation */
    const double a00 = -A[0] / (4.0 * pi), a10 = -A[1] / (4.0 * pi);
    const double a11r = -A[2] / (8.0 * pi), a11i = -A[3] / (8.0 * pi);
    const double a20 = -A[4] / (16.0 * pi);
    const double a21r = -A[5] / (8.0 * pi) * 3.0, a21i = -A[6] / (8.0 * pi) * 3.0;
    const double a22r = -A[7] / (32.0 * pi) * 3.0, a22i = -A[8] / (32.0 * pi) * 3.0;
    const double a30 = -A[9] / (16.0 * pi);
    const double a31r = -A[10] / (64.0 * pi) * 3.0, a31i = -A[11] / (64.0 * pi) * 3.0;
    const double a32r = -A[12] / (32.0 * pi) * 15.0, a32i = -A[13] / (32.0 * pi) * 15.0;
    const double a33r = -A[14] / (64.0 * pi) * 5.0, a33i = -A[15] / (64.0 * pi) * 5.0;
    for (int b = 0; b < s->nblk; ++b) {
        double *u = arr(s, 0, ju, b);
        for (int f = 0; f < 6; ++f) {
            if (s->nkind[b * 6 + f] != ORC_AMR_NBR_NONE) continue;
            int d = f / 2, ag = (f & 1) ? n : -1;
            for (int t2 = -1; t2 <= n; ++t2)
                for (int t1 = -1; t1 <= n; ++t1) {
                    int ii, jj, kk;
                    if (d == 0) { ii = ag; jj = t1; kk = t2; }
                    else if (d == 1) { ii = t1; jj = ag; kk = t2; }
                    else { ii = t1; jj = t2; kk = ag; }
                    double x = coord(s, b, 0, ii), y = coord(s, b, 1, jj), z = coord(s, b, 2, kk);
                    double r2i = 1.0 / (x * x + y * y + z * z);
                    double ri = sqrt(r2i);
                    double r3i = r2i * ri, r4i = r2i * r2i;
                    double rc = sqrt(x * x + y * y);
                    double rci = 1.0 / rc;
                    double cost = z * ri, sint = rc * ri, cosp = x * rci, sinp = y * rci;
                    double c2 = cost * cost, s2 = sint * sint, s3 = s2 * sint;
                    double cp2 = cosp * cosp, sp2 = sinp * sinp;
                    double ub = ri * a00;
                    ub = ub + r2i * cost * a10;
                    ub = ub + r2i * sint * (cosp * a11r + sinp * a11i) * 2.0;
                    ub = ub + r3i * (3.0 * c2 - 1.0) * a20;
                    ub = ub + r3i * sint * cost * (cosp * a21r + sinp * a21i) * 2.0;
                    ub = ub + r3i * s2 * ((cp2 - sp2) * a22r + 2.0 * sinp * cosp * a22i) * 2.0;
                    ub = ub + r4i * cost * (5.0 * c2 - 3.0) * a30;
                    ub = ub + r4i * sint * (5.0 * c2 - 1.0) * (cosp * a31r + sinp * a31i) * 2.0;
                    ub = ub + r4i * cost * s2 * ((cp2 - sp2) * a32r + 2.0 * sinp * cosp * a32i) * 2.0;
                    ub = ub + r4i * s3 * (cosp * (cp2 - 3.0 * sp2) * a33r + sinp * (3.0 * cp2 - sp2) * a33i) * 2.0;
                    u[gi_rot(n, d, ag, t1, t2)] = ub;
                }
        }
    }
}

/* fmg_psi2g (fmg_poisson.F90:4-49): g = -(F(i) + F(i-1))/2 on leaf blocks (fluxes after reflux),
 * centred difference on covered blocks.  g is [nblk][3][n^3] dense (i fastest). */
static void psi2g(orc_amr *s, int ju, double *g)
{
    const int n = s->nb;
    const size_t n3 = (size_t)n * n * n;
    poisson_flux(s, 0, ju, 0);
    for (int b = 0; b < s->nblk; ++b) {
        const double hi = 1.0 / hof(s, s->level[b], 0);
        const double *u = arr(s, 0, ju, b);
        const double *fx = flx(s, 0, b, 0), *fy = flx(s, 0, b, 1), *fz = flx(s, 0, b, 2);
        double *gx = g + (size_t)b * 3 * n3, *gy = gx + n3, *gz = gy + n3;
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    size_t q = ((size_t)k * n + j) * n + i;
                    if (skip(s, b)) {
                        gx[q] = -(u[GI(n, i + 1, j, k)] - u[GI(n, i - 1, j, k)]) * hi * 0.5;
                        gy[q] = -(u[GI(n, i, j + 1, k)] - u[GI(n, i, j - 1, k)]) * hi * 0.5;
                        gz[q] = -(u[GI(n, i, j, k + 1)] - u[GI(n, i, j, k - 1)]) * hi * 0.5;
                    } else {
                        gx[q] = -(fx[GI(n, i, j, k)] + fx[GI(n, i - 1, j, k)]) * 0.5;
                        gy[q] = -(fy[GI(n, i, j, k)] + fy[GI(n, i, j - 1, k)]) * 0.5;
                        gz[q] = -(fz[GI(n, i, j, k)] + fz[GI(n, i, j, k - 1)]) * 0.5;
                    }
                }
    }
}

/* Leaf-cell norms of an FMG-level-0 array: max |a| and the volume-weighted sum of a^2. */
static void leaf_norms(orc_amr *s, int code, double *amax, double *sq)
{
    const int n = s->nb;
    double mx = 0.0, t = 0.0;
    for (int b = 0; b < s->nblk; ++b) {
        if (skip(s, b)) continue;
        double hh = hof(s, s->level[b], 0);
        double dv = (hh * hh) * hh;
        const double *a = arr(s, 0, code, b);
        double bs = 0.0;
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    double v = a[GI(n, i, j, k)];
                    if (fabs(v) > mx) mx = fabs(v);
                    bs = bs + v * v;
                }
        t = t + bs * dv;
    }
    *amax = mx;
    *sq = t;
}

/* fmg_make_physvar mean removal (fmg_data_if.F90:53-78): mass / volume over leaf cells. */
static void remove_leaf_mean(orc_amr *s, int code)
{
    const int n = s->nb;
    double *pm = (double *)malloc(sizeof(double) * (size_t)(2 * s->nblk + 1));
    double *pv = (double *)malloc(sizeof(double) * (size_t)(2 * s->nblk + 1));
    for (int b = 0; b < s->nblk; ++b) {
        pm[b] = 0.0;
        pv[b] = 0.0;
        if (s->covered[b]) continue;
        double hh = hof(s, s->level[b], 0);
        double dv = (hh * hh) * hh;
        const double *a = arr(s, 0, code, b);
        double t = 0.0;
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) t = t + a[GI(n, i, j, k)];
        pm[b] = t * dv;
        pv[b] = (double)(n * n * n) * dv;
    }
    double mass = tree_sum(pm, s->nblk), vol = tree_sum(pv, s->nblk);
    double ave = mass / vol;
    free(pm);
    free(pv);
    for (int b = 0; b < s->nblk; ++b) {
        double *a = arr(s, 0, code, b);
        for (size_t q = 0; q < s->bl[0]; ++q) a[q] = a[q] - ave;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* public interface                                                                           */
/* ------------------------------------------------------------------------------------------ */

orc_amr *orc_amr_create(int nblk, int nlev, const int *level, const int *ijk, const int *parent,
                        const int *child, const int *nbr, const int base_dims[3], double h0,
                        const double x0[3], int bc)
{
    if (nblk <= 0 || nlev <= 0 || nlev > ORC_AMR_MAXLEV) return NULL;
    orc_amr *s = (orc_amr *)calloc(1, sizeof(orc_amr));
    s->nblk = nblk;
    s->nlev = nlev;
    s->nb = 8;
    s->h0 = h0;
    s->bc = bc;
    s->zero_average = bc == ORC_AMR_PERIODIC;
    for (int a = 0; a < 3; ++a) { s->x0[a] = x0[a]; s->base_dims[a] = base_dims[a]; }
    s->level = (int *)malloc(sizeof(int) * nblk);
    s->ijk = (int *)malloc(sizeof(int) * 3 * nblk);
    s->parent = (int *)malloc(sizeof(int) * nblk);
    s->child = (int *)malloc(sizeof(int) * 8 * nblk);
    s->nbr = (int *)malloc(sizeof(int) * 6 * nblk);
    s->nkind = (int *)malloc(sizeof(int) * 6 * nblk);
    s->octant = (int *)malloc(sizeof(int) * nblk);
    s->covered = (int *)malloc(sizeof(int) * nblk);
    s->lev_start = (int *)calloc(nlev + 1, sizeof(int));
    memcpy(s->level, level, sizeof(int) * nblk);
    memcpy(s->ijk, ijk, sizeof(int) * 3 * nblk);
    memcpy(s->parent, parent, sizeof(int) * nblk);
    memcpy(s->child, child, sizeof(int) * 8 * nblk);
    memcpy(s->nbr, nbr, sizeof(int) * 6 * nblk);
    for (int b = 0; b < nblk; ++b) {
        if (b > 0 && level[b] < level[b - 1]) { orc_amr_destroy(s); return NULL; }
        s->lev_start[level[b] + 1] = b + 1;
        s->covered[b] = child[8 * b] >= 0;
        s->octant[b] = parent[b] < 0 ? -1 : (ijk[3 * b] & 1) + 2 * (ijk[3 * b + 1] & 1) + 4 * (ijk[3 * b + 2] & 1);
        for (int f = 0; f < 6; ++f) {
            int m = nbr[6 * b + f];
            if (m < 0) s->nkind[6 * b + f] = ORC_AMR_NBR_NONE;
            else if (level[m] == level[b]) s->nkind[6 * b + f] = ORC_AMR_NBR_SAME;
            else if (level[m] == level[b] - 1) s->nkind[6 * b + f] = ORC_AMR_NBR_PARENT;
            else { orc_amr_destroy(s); return NULL; }
        }
    }
    for (int l = 1; l <= nlev; ++l)
        if (s->lev_start[l] < s->lev_start[l - 1]) s->lev_start[l] = s->lev_start[l - 1];
    for (int L = 0; L < NFMG; ++L) {
        s->n[L] = s->nb >> L;
        s->bl[L] = orc_ghosted_len(s->n[L]);
        for (int c = 0; c < ORC_AMR_NCODE; ++c) s->A[L][c] = (double *)calloc((size_t)nblk * s->bl[L], sizeof(double));
        s->F[L] = (double *)calloc((size_t)nblk * 3 * s->bl[L], sizeof(double));
    }
    s->fill0_level = -1;
    return s;
}

void orc_amr_destroy(orc_amr *s)
{
    if (!s) return;
    for (int L = 0; L < NFMG; ++L) {
        for (int c = 0; c < ORC_AMR_NCODE; ++c) free(s->A[L][c]);
        free(s->F[L]);
    }
    free(s->level); free(s->ijk); free(s->parent); free(s->child); free(s->nbr);
    free(s->nkind); free(s->octant); free(s->covered); free(s->lev_start);
    free(s);
}

void orc_amr_set_zero_average(orc_amr *s, int on) { s->zero_average = on; }
size_t orc_amr_block_len(const orc_amr *s, int L) { return s->bl[L]; }
int orc_amr_get_fill0_level(const orc_amr *s) { return s->fill0_level; }
void orc_amr_set_fill0_level(orc_amr *s, int v) { s->fill0_level = v; }
double orc_amr_resh2maxg(const orc_amr *s) { return s->resh2maxg; }
void orc_amr_base_stats(const orc_amr *s, long *calls, long *iters) { *calls = s->base_calls; *iters = s->base_iters; }

/* Copy a whole state array ([nblk][(n+2)^3]) in or out. */
void orc_amr_set_array(orc_amr *s, int L, int code, const double *src)
{
    memcpy(s->A[L][code], src, sizeof(double) * (size_t)s->nblk * s->bl[L]);
}

void orc_amr_get_array(const orc_amr *s, int L, int code, double *dst)
{
    memcpy(dst, s->A[L][code], sizeof(double) * (size_t)s->nblk * s->bl[L]);
}

/* Single operations on the current state, for operator-by-operator parity tests. */
double orc_amr_op(orc_amr *s, int op, int L, int a, int b, int c, int flag)
{
    switch (op) {
    case ORC_AMR_OP_C2P: c2p(s, L, a, flag); return 0.0;
    case ORC_AMR_OP_GHOSTFIX: ghostcell_fix(s, L, a, flag & 1); return 0.0;   /* bit 2: GPU-side form only */
    case ORC_AMR_OP_FILL0: fill0_faces(s, L, a, 0, s->nlev); return 0.0;
    case ORC_AMR_OP_EXTRAP: extrap_faces(s, L, a, 0, s->nlev); return 0.0;
    case ORC_AMR_OP_RELAX: poisson_relax(s, L, a, b, flag); return s->resh2maxg;
    case ORC_AMR_OP_RESID: poisson_resid(s, L, a, b, c, flag); return 0.0;
    case ORC_AMR_OP_RSTRCT: fmg_rstrct(s, L, a, b); return 0.0;
    case ORC_AMR_OP_INTERP: fmg_interp(s, L, a, b, flag); return 0.0;
    case ORC_AMR_OP_ADDINT: fmg_addint(s, L, a, b, c); return 0.0;
    case ORC_AMR_OP_VMG_FAS: vmg_fas(s, a); return 0.0;
    case ORC_AMR_OP_FMG_LIN: fmg_lin(s); return s->resh2maxg;
    case ORC_AMR_OP_SPHERE_BC: orc_amr_sphere_bc(s, a, b); return 0.0;
    case ORC_AMR_OP_REMOVE_MEAN: remove_leaf_mean(s, a); return 0.0;
    case ORC_AMR_OP_VMG_TAU:   /* vmg_tau(AMR level flag, jtau a, ju b) at FMG level L (vmg.F90:440-458) */
        if (flag < 0 || flag >= s->nlev) return -1.0;
        vmg_tau(s, L, flag, a, b);
        return 0.0;
    default: return -1.0;
    }
}

/* fmg_cycle (fmg.F90:134-233) for the Poisson equation.  psi: [nblk][n^3] initial guess in,
 * potential out; src: [nblk][n^3] source 4 pi G rho (mean removed here when zero_average is on,
 * as fmg_make_physvar does); g: [nblk][3][n^3] out (may be NULL).  Returns ORC_OK when converged,
 * ORC_NOT_CONVERGED when the loop ended on maxit or on the stall test. */
int orc_amr_solve(orc_amr *s, double *psi, const double *src, double *g, const orc_params *p, orc_amr_stats *st)
{
    const int n = s->nb;
    const size_t n3 = (size_t)n * n * n;
    for (int L = 0; L < NFMG; ++L) {
        for (int c = 0; c < ORC_AMR_NCODE; ++c) memset(s->A[L][c], 0, sizeof(double) * (size_t)s->nblk * s->bl[L]);
        memset(s->F[L], 0, sizeof(double) * (size_t)s->nblk * 3 * s->bl[L]);
    }
    s->fill0_level = -1;
    for (int b = 0; b < s->nblk; ++b) {
        double *u = arr(s, 0, C_U, b), *r = arr(s, 0, C_RHO, b);
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    size_t q = ((size_t)b * n3) + ((size_t)k * n + j) * n + i;
                    u[GI(n, i, j, k)] = psi[q];
                    r[GI(n, i, j, k)] = src[q];
                }
    }
    if (s->zero_average) remove_leaf_mean(s, C_RHO);
    memset(st, 0, sizeof(*st));
    fmg_copy(s, 0, C_PSI, C_U, 0);
    fmg_copy(s, 0, C_SRC, C_RHO, 0);
    c2p(s, 0, C_SRC, 0);
    c2p(s, 0, C_PSI, 0);
    if (s->bc == ORC_AMR_SPHERE) orc_amr_sphere_bc(s, C_PSI, C_SRC);
    double resh2max = h2absmax(s, 0, C_SRC);
    double srcmax, srcsq;
    leaf_norms(s, C_SRC, &srcmax, &srcsq);
    int nmax = p->maxit;
    if (fabs(resh2max) <= 2.2250738585072014e-308) nmax = 0;
    double prev = HUGE_VAL;
    int it = 0, status = ORC_NOT_CONVERGED;
    for (;;) {
        poisson_resid(s, 0, C_RHO, C_PSI, C_SRC, 0);
        double rmax, rsq;
        leaf_norms(s, C_RHO, &rmax, &rsq);
        st->rel_linf = srcmax > 0 ? rmax / srcmax : 0.0;
        st->rel_l2 = srcsq > 0 ? sqrt(rsq / srcsq) : 0.0;
        st->rel_h2 = resh2max > 0 ? h2absmax(s, 0, C_RHO) / resh2max : 0.0;
        if (it < ORC_HIST) st->hist[it] = st->rel_linf;
        if (p->mode == ORC_STOP_TRUE_RESIDUAL && st->rel_linf < p->tol) { status = ORC_OK; break; }
        if (it >= nmax) break;
        fmg_lin(s);
        fmg_add(s, 0, C_PSI, C_U);
        ++it;
        double ratio = s->resh2maxg / resh2max;
        st->ratio = ratio;
        if (p->mode == ORC_STOP_REFERENCE) {
            if (ratio < p->tol) { status = ORC_OK; break; }
            if (ratio > prev * 0.5) break;   /* stall test, fmg.F90:188 */
            prev = ratio;
        }
    }
    if (nmax == 0) status = ORC_OK;
    if (p->mode == ORC_STOP_REFERENCE) {
        poisson_resid(s, 0, C_RHO, C_PSI, C_SRC, 0);
        double rmax, rsq;
        leaf_norms(s, C_RHO, &rmax, &rsq);
        st->rel_linf = srcmax > 0 ? rmax / srcmax : 0.0;
        st->rel_l2 = srcsq > 0 ? sqrt(rsq / srcsq) : 0.0;
        st->rel_h2 = resh2max > 0 ? h2absmax(s, 0, C_RHO) / resh2max : 0.0;
    }
    st->iterations = it;
    st->status = status;
    /* fmg.F90:205-216: restore U from PSI (all blocks), refresh covered blocks, g from fluxes */
    fmg_copy(s, 0, C_U, C_PSI, 1);
    c2p(s, 0, C_U, 0);
    for (int b = 0; b < s->nblk; ++b) {
        const double *u = arr(s, 0, C_U, b);
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) psi[(size_t)b * n3 + ((size_t)k * n + j) * n + i] = u[GI(n, i, j, k)];
    }
    if (g) psi2g(s, C_U, g);
    return status;
}
