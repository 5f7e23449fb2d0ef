/*
 * ORACLE -- test infrastructure, not product code.
 *
 * CPU FP64 restatement of the reference's uniform-grid (base-grid) multigrid solver, used only as
 * the checker: by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg.  The
 * product path (code_only_amd/csrc) never links, loads or calls anything in this directory.
 *
 * Layout follows the reference: every level array carries one ghost layer (Ngh = 1), index range
 * -1..n in each axis, i (x) fastest -- the memory order of Fortran u(-1:n,-1:n,-1:n).
 *
 * Arithmetic is written in the reference's left-to-right operand order and this file is compiled
 * with -ffp-contract=off so that no multiply-add is fused; the GPU FP64 path of the product is
 * built the same way and is compared with this file bit for bit.
 *
 * Reference functions restated here (all paths relative to the reference repository root):
 *   orc_boundary_u         fmg_boundary_phys_periodic.F90:71-104   (mg_boundary_u, periodic wrap)
 *   orc_relax              mg_poisson.F90:32-79                    (mg_poisson_relax, RB-GS)
 *   orc_resid              mg_poisson.F90:5-28                     (mg_poisson_resid)
 *   orc_rstrct_linear      restriction.F90:127-151                 (rstrct_linear)
 *   orc_rstrct_cubic       restriction.F90:38-119                  (rstrct_cubic)
 *   orc_interp_trilinear   interpolation.F90:222-247               (interp_trilinear)
 *   orc_interp_tricubic    interpolation.F90:40-218                (interp_tricubic)
 *   orc_mg_interp          mg.F90:608-633                          (mg_interp)
 *   orc_mg_lin             mg.F90:106-218                          (mg_lin)
 *   orc_mg_cycle           mg.F90:42-94                            (mg_cycle)
 *   orc_exact_periodic     no reference counterpart: the coarsest-level solve of grids with
 *                          n = 3 * 2^a (see orc_mg_setup), whose odd 3^3 level neither halves
 *                          further nor admits a red-black ordering under periodic wrap
 *   orc_gradient           fmg_poisson.F90:29-31                   (centred g = -grad phi)
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define GI(n, i, j, k) ((size_t)((k) + 1) * (size_t)((n) + 2) * (size_t)((n) + 2) + \
                        (size_t)((j) + 1) * (size_t)((n) + 2) + (size_t)((i) + 1))

size_t orc_ghosted_len(int n) { return (size_t)(n + 2) * (size_t)(n + 2) * (size_t)(n + 2); }

/* Boundary mode of the uniform grid (the base grid of the MG tier):
 *   ORC_BC_PERIODIC  -- fmg_boundary_phys_periodic.F90: mg_boundary_u wraps the ghosts; the base
 *                       blocks have no physical boundary, so fill0 / extrap do nothing;
 *   ORC_BC_DIRICHLET -- fmg_boundary_phys_sphere.F90:71-75: mg_boundary_u does nothing; every face
 *                       touches the physical boundary, so mg_boundary_fill0 zeroes the ghosts
 *                       (mg.F90:735-750) and mg_boundary_extrap mirrors them (mg.F90:754-771). */
static int g_bc = ORC_BC_PERIODIC;
void orc_set_bc(int bc) { g_bc = bc; }

static void wrap_u(double *u, int n);

/* mg_boundary_u of the selected boundary module. */
void orc_boundary_u(double *u, int n)
{
    if (g_bc == ORC_BC_PERIODIC) wrap_u(u, n);
}

/* mg.F90:735-750 with every face touching the physical boundary (Dirichlet mode only). */
void orc_boundary_fill0(double *u, int n)
{
    if (g_bc != ORC_BC_DIRICHLET) return;
    for (int k = -1; k <= n; ++k)
        for (int j = -1; j <= n; ++j) { u[GI(n, -1, j, k)] = 0.0; u[GI(n, n, j, k)] = 0.0; }
    for (int k = -1; k <= n; ++k)
        for (int i = -1; i <= n; ++i) { u[GI(n, i, -1, k)] = 0.0; u[GI(n, i, n, k)] = 0.0; }
    for (int j = -1; j <= n; ++j)
        for (int i = -1; i <= n; ++i) { u[GI(n, i, j, -1)] = 0.0; u[GI(n, i, j, n)] = 0.0; }
}

/* mg.F90:754-771: u(ghost) = -u(first interior), x faces, then y, then z (Dirichlet mode only). */
void orc_boundary_extrap(double *u, int n)
{
    if (g_bc != ORC_BC_DIRICHLET) return;
    for (int k = -1; k <= n; ++k)
        for (int j = -1; j <= n; ++j) { u[GI(n, -1, j, k)] = -u[GI(n, 0, j, k)]; u[GI(n, n, j, k)] = -u[GI(n, n - 1, j, k)]; }
    for (int k = -1; k <= n; ++k)
        for (int i = -1; i <= n; ++i) { u[GI(n, i, -1, k)] = -u[GI(n, i, 0, k)]; u[GI(n, i, n, k)] = -u[GI(n, i, n - 1, k)]; }
    for (int j = -1; j <= n; ++j)
        for (int i = -1; i <= n; ++i) { u[GI(n, i, j, -1)] = -u[GI(n, i, j, 0)]; u[GI(n, i, j, n)] = -u[GI(n, i, j, n - 1)]; }
}

/* fmg_boundary_phys_periodic.F90:71-104 -- ghost wrap i, then j, then k (edges/corners wrap too). */
static void wrap_u(double *u, int n)
{
    for (int k = -1; k <= n; ++k)
        for (int j = -1; j <= n; ++j) {
            u[GI(n, -1, j, k)] = u[GI(n, n - 1, j, k)];
            u[GI(n, n, j, k)] = u[GI(n, 0, j, k)];
        }
    for (int k = -1; k <= n; ++k)
        for (int i = -1; i <= n; ++i) {
            u[GI(n, i, -1, k)] = u[GI(n, i, n - 1, k)];
            u[GI(n, i, n, k)] = u[GI(n, i, 0, k)];
        }
    for (int j = -1; j <= n; ++j)
        for (int i = -1; i <= n; ++i) {
            u[GI(n, i, j, -1)] = u[GI(n, i, j, n - 1)];
            u[GI(n, i, j, n)] = u[GI(n, i, j, 0)];
        }
}

/* mg_poisson.F90:32-79.  Returns Resmaxg = max|resh2| / h^2 over both colours. */
double orc_relax(double *u, const double *rhs, int n, double h)
{
    const double sixth = 1.0 / 6.0;
    const double h2 = h * h;
    double resh2max = 0.0;
    for (int ipass = 1; ipass <= 2; ++ipass) {
        orc_boundary_u(u, n);
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j)
                for (int i = (j + k + ipass) % 2; i < n; i += 2) {
                    double resh2 = u[GI(n, i + 1, j, k)] + u[GI(n, i - 1, j, k)];
                    resh2 = resh2 + u[GI(n, i, j + 1, k)];
                    resh2 = resh2 + u[GI(n, i, j - 1, k)];
                    resh2 = resh2 + u[GI(n, i, j, k + 1)];
                    resh2 = resh2 + u[GI(n, i, j, k - 1)];
                    resh2 = resh2 - 6.0 * u[GI(n, i, j, k)];
                    resh2 = resh2 - h2 * rhs[GI(n, i, j, k)];
                    u[GI(n, i, j, k)] = u[GI(n, i, j, k)] + sixth * resh2;
                    double a = fabs(resh2);
                    if (a > resh2max) resh2max = a;
                }
    }
    return resh2max / h2;
}

/* mg_poisson.F90:5-28: res = rhs - (sum6 - 6u)/h^2, after refreshing u's ghosts. */
void orc_resid(double *res, double *u, const double *rhs, int n, double h)
{
    const double h2i = 1.0 / (h * h);
    orc_boundary_u(u, n);
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                double s = u[GI(n, i + 1, j, k)] + u[GI(n, i - 1, j, k)];
                s = s + u[GI(n, i, j + 1, k)];
                s = s + u[GI(n, i, j - 1, k)];
                s = s + u[GI(n, i, j, k + 1)];
                s = s + u[GI(n, i, j, k - 1)];
                s = s - 6.0 * u[GI(n, i, j, k)];
                res[GI(n, i, j, k)] = -(h2i * s) + rhs[GI(n, i, j, k)];
            }
}

/* restriction.F90:127-151: 8-cell average, operand order as written there. */
void orc_rstrct_linear(double *uc, const double *uf, int nc)
{
    const int nf = 2 * nc;
    const double rdv = 1.0 / 8.0;
    for (int kc = 0; kc < nc; ++kc)
        for (int jc = 0; jc < nc; ++jc)
            for (int ic = 0; ic < nc; ++ic) {
                int i = 2 * ic, j = 2 * jc, k = 2 * kc;
                double s = uf[GI(nf, i, j, k)] + uf[GI(nf, i + 1, j, k)];
                s = s + uf[GI(nf, i, j + 1, k)];
                s = s + uf[GI(nf, i, j, k + 1)];
                s = s + uf[GI(nf, i + 1, j + 1, k)];
                s = s + uf[GI(nf, i + 1, j, k + 1)];
                s = s + uf[GI(nf, i, j + 1, k + 1)];
                s = s + uf[GI(nf, i + 1, j + 1, k + 1)];
                uc[GI(nc, ic, jc, kc)] = s * rdv;
            }
}

/* restriction.F90:38-63: 1D cubic restriction weights at x = 1.5 through nodes 0..3. */
void orc_rstrct_cubic_weights(double s1d[4])
{
    const double ic = 1.5;
    const int ii[4] = {0, 1, 2, 3};
    s1d[0] = (ic - ii[1]) * (ic - ii[2]) * (ic - ii[3]) / (double)((ii[0] - ii[1]) * (ii[0] - ii[2]) * (ii[0] - ii[3]));
    s1d[1] = (ic - ii[0]) * (ic - ii[2]) * (ic - ii[3]) / (double)((ii[1] - ii[0]) * (ii[1] - ii[2]) * (ii[1] - ii[3]));
    s1d[2] = (ic - ii[0]) * (ic - ii[1]) * (ic - ii[3]) / (double)((ii[2] - ii[0]) * (ii[2] - ii[1]) * (ii[2] - ii[3]));
    s1d[3] = (ic - ii[0]) * (ic - ii[1]) * (ic - ii[2]) / (double)((ii[3] - ii[0]) * (ii[3] - ii[1]) * (ii[3] - ii[2]));
}

/* restriction.F90:71-119: tensor cubic restriction; uf's ghosts must be filled. */
void orc_rstrct_cubic(double *uc, const double *uf, int nc)
{
    const int nf = 2 * nc;
    double s1d[4], st[4][4][4];
    orc_rstrct_cubic_weights(s1d);
    for (int a = 0; a < 4; ++a)        /* St(a,b,c) = (s1d s1d^T)(a,b) * s1d(c) */
        for (int b = 0; b < 4; ++b)
            for (int c = 0; c < 4; ++c) st[c][b][a] = (s1d[a] * s1d[b]) * s1d[c];
    for (int kc = 0; kc < nc; ++kc)
        for (int jc = 0; jc < nc; ++jc)
            for (int ic = 0; ic < nc; ++ic) {
                int i0 = 2 * ic - 1, j0 = 2 * jc - 1, k0 = 2 * kc - 1;
                double s = 0.0;
                int first = 1;
                for (int c = 0; c < 4; ++c)
                    for (int b = 0; b < 4; ++b)
                        for (int a = 0; a < 4; ++a) {
                            double t = uf[GI(nf, i0 + a, j0 + b, k0 + c)] * st[c][b][a];
                            s = first ? t : s + t;
                            first = 0;
                        }
                uc[GI(nc, ic, jc, kc)] = s;
            }
}

/* config.h:249 IJKC with origin 0: floor(i/2). */
static inline int ijkc(int i) { return i / 2 + ((i < 0) ? ((i % 2) ? -1 : 0) : 0); }
/* interpolation.F90:55 NEAR(nf, nc) = nc + 2*modulo(nf,2) - 1 */
static inline int near_(int nf, int nc) { return nc + 2 * (((nf % 2) + 2) % 2) - 1; }

/* interpolation.F90:222-247 */
void orc_interp_trilinear(double *uf, const double *uc, int nf)
{
    const int nc = nf / 2 > 0 ? nf / 2 : 1;
    for (int kf = 0; kf < nf; ++kf)
        for (int jf = 0; jf < nf; ++jf)
            for (int i = 0; i < nf; ++i) {
                int ic = ijkc(i), jc = ijkc(jf), kc = ijkc(kf);
                int ic1 = near_(i, ic), jc1 = near_(jf, jc), kc1 = near_(kf, kc);
                double b = uc[GI(nc, ic1, jc, kc)] + uc[GI(nc, ic, jc1, kc)];
                b = b + uc[GI(nc, ic, jc, kc1)];
                double e = uc[GI(nc, ic, jc1, kc1)] + uc[GI(nc, ic1, jc, kc1)];
                e = e + uc[GI(nc, ic1, jc1, kc)];
                double s = 27.0 * uc[GI(nc, ic, jc, kc)] + 9.0 * b;
                s = s + 3.0 * e;
                s = s + uc[GI(nc, ic1, jc1, kc1)];
                uf[GI(nf, i, jf, kf)] = s / 64.0;
            }
}

/* One axis of interp_tricubic (interpolation.F90:106-127): the four coarse indices (shifted at
 * the array edges, lo..hi = ghosted coarse bounds) and the Lagrange weights for fine index f. */
void orc_cubic_stencil(int f, int lo, int hi, int idx[4], double a[4])
{
    int cp = ijkc(f);
    int cn = near_(f, cp);
    int c1 = cp < cn ? cp : cn;
    int c2 = cp < cn ? cn : cp;
    int c0 = 2 * c1 - c2;
    int c3 = 2 * c2 - c1;
    int nshift = 0;
    if (c3 > hi) nshift = c3 - hi;
    if (c0 < lo) nshift = c0 - lo;
    c0 -= nshift; c1 -= nshift; c2 -= nshift; c3 -= nshift;
    double xc = cp + 0.25 * (double)(2 * (((f % 2) + 2) % 2) - 1);
    a[0] = (xc - c1) * (xc - c2) * (xc - c3) / (double)((c0 - c1) * (c0 - c2) * (c0 - c3));
    a[1] = (xc - c0) * (xc - c2) * (xc - c3) / (double)((c1 - c0) * (c1 - c2) * (c1 - c3));
    a[2] = (xc - c0) * (xc - c1) * (xc - c3) / (double)((c2 - c0) * (c2 - c1) * (c2 - c3));
    a[3] = (xc - c0) * (xc - c1) * (xc - c2) / (double)((c3 - c0) * (c3 - c1) * (c3 - c2));
    idx[0] = c0; idx[1] = c1; idx[2] = c2; idx[3] = c3;
}

/* interpolation.F90:40-218 on a ghosted uniform coarse array (bounds -1..nc): x pass into bufx,
 * y pass into bufxy, z pass into uf.  Arrays narrower than 4 with ghosts fall back to trilinear
 * (interpolation.F90:87-96). */
void orc_interp_tricubic(double *uf, const double *uc, int nf)
{
    const int nc = nf / 2 > 0 ? nf / 2 : 1;
    const int lo = -1, hi = nc;
    const int ncg = nc + 2;
    if (ncg < 4) { orc_interp_trilinear(uf, uc, nf); return; }
    double *bufx = (double *)malloc(sizeof(double) * (size_t)ncg * ncg * nf);
    double *bufxy = (double *)malloc(sizeof(double) * (size_t)ncg * nf * nf);
    int idx[4]; double a[4];
    /* x: bufx(if, jc, kc), jc,kc in lo..hi */
    for (int kc = lo; kc <= hi; ++kc)
        for (int jc = lo; jc <= hi; ++jc)
            for (int f = 0; f < nf; ++f) {
                orc_cubic_stencil(f, lo, hi, idx, a);
                double s = a[0] * uc[GI(nc, idx[0], jc, kc)] + a[1] * uc[GI(nc, idx[1], jc, kc)];
                s = s + a[2] * uc[GI(nc, idx[2], jc, kc)];
                s = s + a[3] * uc[GI(nc, idx[3], jc, kc)];
                bufx[((size_t)(kc - lo) * ncg + (jc - lo)) * nf + f] = s;
            }
    /* y: bufxy(if, jf, kc) */
    for (int kc = lo; kc <= hi; ++kc)
        for (int jf = 0; jf < nf; ++jf) {
            orc_cubic_stencil(jf, lo, hi, idx, a);
            for (int f = 0; f < nf; ++f) {
                const double *bx = bufx + (size_t)(kc - lo) * ncg * nf + f;
                double s = a[0] * bx[(size_t)(idx[0] - lo) * nf] + a[1] * bx[(size_t)(idx[1] - lo) * nf];
                s = s + a[2] * bx[(size_t)(idx[2] - lo) * nf];
                s = s + a[3] * bx[(size_t)(idx[3] - lo) * nf];
                bufxy[((size_t)(kc - lo) * nf + jf) * nf + f] = s;
            }
        }
    /* z: uf(if, jf, kf) */
    for (int kf = 0; kf < nf; ++kf) {
        orc_cubic_stencil(kf, lo, hi, idx, a);
        for (int jf = 0; jf < nf; ++jf)
            for (int f = 0; f < nf; ++f) {
                const double *bxy = bufxy + (size_t)jf * nf + f;
                size_t pl = (size_t)nf * nf;
                double s = a[0] * bxy[(size_t)(idx[0] - lo) * pl] + a[1] * bxy[(size_t)(idx[1] - lo) * pl];
                s = s + a[2] * bxy[(size_t)(idx[2] - lo) * pl];
                s = s + a[3] * bxy[(size_t)(idx[3] - lo) * pl];
                uf[GI(nf, f, jf, kf)] = s;
            }
    }
    free(bufx);
    free(bufxy);
}

/* ------------------------------------------------------------------------------------------ */
/* mg_lin / mg_cycle                                                                          */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
    int nlev;              /* levels 0..nlev-1, n_l = n >> l, last level is 1^3 or 3^3 */
    int n[32];
    double h[32];
    double *U[32], *RHS[32], *RES[32], *RHO[32];
    double resmaxg;        /* Resmaxg of the most recent relax call (mg.F90:12) */
} orc_mg;

static double *zalloc(int n) { return (double *)calloc(orc_ghosted_len(n), sizeof(double)); }

static int orc_mg_setup(orc_mg *m, int n, double h)
{
    memset(m, 0, sizeof(*m));
    /* mg_data.F90:57-75: MG_LevelMax = log2(boxsize) (+0.5 rounding); level l has n/2^l cells.
     * That formula only describes n = 2^a (for n = 96 it yields a 1-cell level from a 3-cell one
     * and then a 0-cell level).  Extension for the batched uniform problems: n = 3 * 2^a halves
     * while even, down to a 3^3 coarsest level, which orc_exact_periodic solves exactly. */
    int lmax = 0;
    while ((n >> lmax) % 2 == 0) ++lmax;
    const int nc = n >> lmax;
    if (nc != 1 && nc != 3) return -1;
    if (nc == 3 && g_bc != ORC_BC_PERIODIC) return -1;
    if (lmax + 1 > 32) return -1;
    m->nlev = lmax + 1;
    for (int l = 0; l <= lmax; ++l) {
        m->n[l] = n >> l;
        m->h[l] = h * (double)(1 << l);          /* mg_data.F90:554-559 */
        m->U[l] = zalloc(m->n[l]);
        m->RHS[l] = zalloc(m->n[l]);
        m->RES[l] = zalloc(m->n[l]);
        m->RHO[l] = zalloc(m->n[l]);
    }
    return 0;
}

/* Exact solve of (sum6 - 6u)/h^2 = rhs on a periodic n^3 grid (the mean of u set to zero).  The
 * 1-D periodic second difference is diagonalised by the Hartley basis
 *     Q[m][k] = (cos(2 pi mk/n) + sin(2 pi mk/n)) / sqrt(n),   lam[m] = (2 cos(2 pi m/n) - 2) / h^2,
 * which is symmetric and orthonormal (Q Q = I), so u = Q3 (Lambda^+ (Q3 rhs)) with Q3 applied as
 * three axis passes, each output line accumulated k = 0 .. n-1 in order; Lambda = (lam_i + lam_j)
 * + lam_k, and the zero mode is set to zero.  u's ghosts are refreshed afterwards. */
void orc_exact_basis(int n, double h, double *Q, double *lam)
{
    const double pi = 3.14159265358979323846;
    for (int m = 0; m < n; ++m) {
        for (int k = 0; k < n; ++k) {
            const double a = 2.0 * pi * (double)((m * k) % n) / (double)n;
            Q[m * n + k] = (cos(a) + sin(a)) / sqrt((double)n);
        }
        lam[m] = (2.0 * cos(2.0 * pi * (double)m / (double)n) - 2.0) / (h * h);
    }
}

static void exact_pass(double *g, const double *Q, int n, int axis)
{
    double x[64], y[64];
    for (int b = 0; b < n; ++b)
        for (int a = 0; a < n; ++a) {
            for (int k = 0; k < n; ++k) {
                const size_t q = axis == 0 ? ((size_t)b * n + a) * n + k
                               : axis == 1 ? ((size_t)b * n + k) * n + a : ((size_t)k * n + b) * n + a;
                x[k] = g[q];
            }
            for (int m = 0; m < n; ++m) {
                double acc = 0.0;
                for (int k = 0; k < n; ++k) acc = acc + Q[m * n + k] * x[k];
                y[m] = acc;
            }
            for (int m = 0; m < n; ++m) {
                const size_t q = axis == 0 ? ((size_t)b * n + a) * n + m
                               : axis == 1 ? ((size_t)b * n + m) * n + a : ((size_t)m * n + b) * n + a;
                g[q] = y[m];
            }
        }
}

int orc_exact_periodic(double *u, const double *rhs, int n, double h)
{
    if (n < 1 || n > 64) return -1;
    double *Q = (double *)malloc(sizeof(double) * n * n), *lam = (double *)malloc(sizeof(double) * n);
    double *g = (double *)malloc(sizeof(double) * n * n * n);
    orc_exact_basis(n, h, Q, lam);
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) g[((size_t)k * n + j) * n + i] = rhs[GI(n, i, j, k)];
    for (int a = 0; a < 3; ++a) exact_pass(g, Q, n, a);
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const double den = (lam[i] + lam[j]) + lam[k];
                double *v = g + ((size_t)k * n + j) * n + i;
                *v = den != 0.0 ? *v / den : 0.0;
            }
    for (int a = 0; a < 3; ++a) exact_pass(g, Q, n, a);
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) u[GI(n, i, j, k)] = g[((size_t)k * n + j) * n + i];
    orc_boundary_u(u, n);
    free(Q);
    free(lam);
    free(g);
    return 0;
}

static void orc_mg_free(orc_mg *m)
{
    for (int l = 0; l < m->nlev; ++l) { free(m->U[l]); free(m->RHS[l]); free(m->RES[l]); free(m->RHO[l]); }
}

static void fill0(double *u, int n) { memset(u, 0, sizeof(double) * orc_ghosted_len(n)); }

/* mg.F90:608-633: coarse ghosts extrapolated and refreshed, cubic or linear prolongation into the
 * fine interior, then fine and coarse ghosts zeroed (Dirichlet) and refreshed (periodic). */
void orc_mg_interp(double *uf, double *uc, int nf, int cubic)
{
    int nc = nf / 2;
    orc_boundary_extrap(uc, nc);
    orc_boundary_u(uc, nc);
    if (cubic) orc_interp_tricubic(uf, uc, nf);
    else orc_interp_trilinear(uf, uc, nf);
    orc_boundary_fill0(uf, nf);
    orc_boundary_fill0(uc, nc);
    orc_boundary_u(uf, nf);
    orc_boundary_u(uc, nc);
}

static void orc_mg_relax(orc_mg *m, int l, double *u, const double *rhs)
{
    m->resmaxg = orc_relax(u, rhs, m->n[l], m->h[l]);
}

/* mg.F90:540-546 (two sweeps); a 3^3 coarsest level is solved exactly instead. */
static void orc_mg_slvsml(orc_mg *m, int l, double *u, const double *rhs)
{
    if (m->n[l] == 3) {
        orc_exact_periodic(u, rhs, 3, m->h[l]);
        return;
    }
    for (int it = 0; it < 2; ++it) orc_mg_relax(m, l, u, rhs);
}

/* mg.F90:106-218 (Ncycle = 2, Npre = Npost = 2): FMG from the coarsest level, V-cycles per level. */
static void orc_mg_lin(orc_mg *m, int ncycle, int npre, int npost)
{
    const int lmax = m->nlev - 1;
    for (int l = 1; l <= lmax; ++l) orc_rstrct_linear(m->RHO[l], m->RHO[l - 1], m->n[l]);
    fill0(m->U[lmax], m->n[lmax]);
    orc_mg_slvsml(m, lmax, m->U[lmax], m->RHO[lmax]);
    for (int ml = lmax - 1; ml >= 0; --ml) {
        orc_mg_interp(m->U[ml], m->U[ml + 1], m->n[ml], 1);
        memcpy(m->RHS[ml], m->RHO[ml], sizeof(double) * orc_ghosted_len(m->n[ml]));
        for (int jc = 0; jc < ncycle; ++jc) {
            for (int v = ml; v <= lmax - 1; ++v) {
                for (int p = 0; p < npre; ++p) orc_mg_relax(m, v, m->U[v], m->RHS[v]);
                orc_resid(m->RES[v], m->U[v], m->RHS[v], m->n[v], m->h[v]);
                orc_rstrct_linear(m->RHS[v + 1], m->RES[v], m->n[v + 1]);
                fill0(m->U[v + 1], m->n[v + 1]);
            }
            orc_mg_slvsml(m, lmax, m->U[lmax], m->RHS[lmax]);
            for (int v = lmax - 1; v >= ml; --v) {
                /* mg_addint (mg.F90:727-731): RES = trilinear(U coarse); U += RES (whole array) */
                orc_mg_interp(m->RES[v], m->U[v + 1], m->n[v], 0);
                size_t len = orc_ghosted_len(m->n[v]);
                for (size_t q = 0; q < len; ++q) m->U[v][q] = m->U[v][q] + m->RES[v][q];
                for (int p = 0; p < npost; ++p) orc_mg_relax(m, v, m->U[v], m->RHS[v]);
            }
        }
    }
}

static void interior_norms(const double *a, int n, double *amax, double *l2)
{
    double mx = 0.0, s = 0.0;
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                double v = a[GI(n, i, j, k)];
                double av = fabs(v);
                if (av > mx) mx = av;
                s += v * v;
            }
    *amax = mx;
    *l2 = sqrt(s);
}

/* mg.F90:42-94.  psi: ghosted initial guess (in) / solution (out); src: ghosted source.
 * mode ORC_STOP_REFERENCE: stop when Resmaxg/rhomax < tol (mg.F90:91) after at most maxit sweeps.
 * mode ORC_STOP_TRUE_RESIDUAL: stop when the post-update residual max|src - L psi| / max|src| < tol. */
int orc_mg_cycle(double *psi, const double *src, int n, double h, const orc_params *p, orc_stats *st)
{
    orc_mg m;
    if (orc_mg_setup(&m, n, h) != 0) return -1;
    size_t len = orc_ghosted_len(n);
    double *PSI = psi;
    double *SRC = (double *)malloc(sizeof(double) * len);
    memcpy(SRC, src, sizeof(double) * len);
    orc_boundary_fill0(PSI, n);   /* mg.F90:67-68 */
    orc_boundary_u(PSI, n);
    double rhomax = 0.0, srcmax, srcl2;
    interior_norms(SRC, n, &srcmax, &srcl2);
    int it = 0, status = 1;
    memset(st, 0, sizeof(*st));
    for (;;) {
        orc_resid(m.RHO[0], PSI, SRC, n, h);
        double rmax, rl2;
        interior_norms(m.RHO[0], n, &rmax, &rl2);
        st->rel_linf = srcmax > 0 ? rmax / srcmax : 0.0;
        st->rel_l2 = srcl2 > 0 ? rl2 / srcl2 : 0.0;
        if (it < ORC_HIST) st->hist[it] = st->rel_linf;
        if (it == 0) rhomax = rmax;
        if (p->mode == ORC_STOP_TRUE_RESIDUAL && st->rel_linf < p->tol) { status = 0; break; }
        if (it >= p->maxit) break;
        orc_mg_lin(&m, p->ncycle, p->npre, p->npost);
        for (size_t q = 0; q < len; ++q) PSI[q] = PSI[q] + m.U[0][q];
        ++it;
        st->resmaxg = m.resmaxg;
        if (p->mode == ORC_STOP_REFERENCE && (rhomax > 0 ? m.resmaxg / rhomax : 0.0) < p->tol) {
            status = 0;
            orc_resid(m.RHO[0], PSI, SRC, n, h);
            double rmax2, rl22;
            interior_norms(m.RHO[0], n, &rmax2, &rl22);
            st->rel_linf = srcmax > 0 ? rmax2 / srcmax : 0.0;
            st->rel_l2 = srcl2 > 0 ? rl22 / srcl2 : 0.0;
            break;
        }
    }
    st->iterations = it;
    st->status = status;
    free(SRC);
    orc_mg_free(&m);
    return status;
}

/* fmg_poisson.F90:29-31 (covered-block form): g_d = -(u(+1) - u(-1)) * hi * 0.5 on a periodic grid.
 * g is dense n^3 per component (component-major: gx, gy, gz). */
void orc_gradient(double *g, double *u, int n, double h)
{
    const double hi = 1.0 / h;
    size_t n3 = (size_t)n * n * n;
    orc_boundary_u(u, n);
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                size_t q = ((size_t)k * n + j) * n + i;
                g[q] = -(u[GI(n, i + 1, j, k)] - u[GI(n, i - 1, j, k)]) * hi * 0.5;
                g[n3 + q] = -(u[GI(n, i, j + 1, k)] - u[GI(n, i, j - 1, k)]) * hi * 0.5;
                g[2 * n3 + q] = -(u[GI(n, i, j, k + 1)] - u[GI(n, i, j, k - 1)]) * hi * 0.5;
            }
}

/* Dense <-> ghosted helpers for the tests. */
void orc_to_ghosted(double *ug, const double *dense, int n)
{
    memset(ug, 0, sizeof(double) * orc_ghosted_len(n));
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            memcpy(ug + GI(n, 0, j, k), dense + ((size_t)k * n + j) * n, sizeof(double) * n);
    orc_boundary_u(ug, n);
}

void orc_from_ghosted(double *dense, const double *ug, int n)
{
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            memcpy(dense + ((size_t)k * n + j) * n, ug + GI(n, 0, j, k), sizeof(double) * n);
}

/* Dense-layout wrappers so the tests can call single operators on n^3 arrays. */
double orc_relax_dense(double *u, const double *rhs, int n, double h)
{
    size_t len = orc_ghosted_len(n);
    double *ug = (double *)malloc(sizeof(double) * len), *rg = (double *)malloc(sizeof(double) * len);
    orc_to_ghosted(ug, u, n); orc_to_ghosted(rg, rhs, n);
    double r = orc_relax(ug, rg, n, h);
    orc_from_ghosted(u, ug, n);
    free(ug); free(rg);
    return r;
}

void orc_resid_dense(double *res, const double *u, const double *rhs, int n, double h)
{
    size_t len = orc_ghosted_len(n);
    double *ug = (double *)malloc(sizeof(double) * len), *rg = (double *)malloc(sizeof(double) * len);
    double *qg = (double *)calloc(len, sizeof(double));
    orc_to_ghosted(ug, u, n); orc_to_ghosted(rg, rhs, n);
    orc_resid(qg, ug, rg, n, h);
    orc_from_ghosted(res, qg, n);
    free(ug); free(rg); free(qg);
}

void orc_rstrct_dense(double *uc, const double *uf, int nc, int cubic)
{
    int nf = 2 * nc;
    double *fg = (double *)malloc(sizeof(double) * orc_ghosted_len(nf));
    double *cg = (double *)calloc(orc_ghosted_len(nc), sizeof(double));
    orc_to_ghosted(fg, uf, nf);
    if (cubic) orc_rstrct_cubic(cg, fg, nc); else orc_rstrct_linear(cg, fg, nc);
    orc_from_ghosted(uc, cg, nc);
    free(fg); free(cg);
}

void orc_interp_dense(double *uf, const double *uc, int nf, int cubic)
{
    int nc = nf / 2;
    double *cg = (double *)malloc(sizeof(double) * orc_ghosted_len(nc));
    double *fg = (double *)calloc(orc_ghosted_len(nf), sizeof(double));
    orc_to_ghosted(cg, uc, nc);
    orc_mg_interp(fg, cg, nf, cubic);
    orc_from_ghosted(uf, fg, nf);
    free(cg); free(fg);
}

int orc_mg_cycle_dense(double *psi, const double *src, int n, double h, const orc_params *p, orc_stats *st)
{
    size_t len = orc_ghosted_len(n);
    double *pg = (double *)malloc(sizeof(double) * len), *sg = (double *)malloc(sizeof(double) * len);
    orc_to_ghosted(pg, psi, n); orc_to_ghosted(sg, src, n);
    int r = orc_mg_cycle(pg, sg, n, h, p, st);
    orc_from_ghosted(psi, pg, n);
    free(pg); free(sg);
    return r;
}

void orc_gradient_dense(double *g, const double *u, int n, double h)
{
    double *ug = (double *)malloc(sizeof(double) * orc_ghosted_len(n));
    orc_to_ghosted(ug, u, n);
    orc_gradient(g, ug, n, h);
    free(ug);
}

int orc_exact_periodic_dense(double *u, const double *rhs, int n, double h)
{
    size_t len = orc_ghosted_len(n);
    double *ug = (double *)calloc(len, sizeof(double)), *rg = (double *)calloc(len, sizeof(double));
    orc_to_ghosted(rg, rhs, n);
    int r = orc_exact_periodic(ug, rg, n, h);
    orc_from_ghosted(u, ug, n);
    free(ug); free(rg);
    return r;
}
