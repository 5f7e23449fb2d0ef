// The base-grid multigrid that the AMR coarse solve hands its base level to (vmg_slvsml,
// vmg.F90:387-432 -> mg_cycle, mg.F90:42-94), as one single-workgroup kernel: the gathered base
// grid is small (32^3 for a 16^3-block base at the 2^3-cell FMG level), so the whole defect-
// correction loop with its FMG cycles runs in one launch with workgroup barriers between phases,
// instead of several hundred tiny launches.  Arrays are ghosted (n+2)^3, i fastest.
//
// Boundary modes: periodic (fmg_boundary_phys_periodic.F90: mg_boundary_u wraps the ghosts) or
// zero-Dirichlet (fmg_boundary_phys_sphere.F90: mg_boundary_u does nothing; mg_boundary_fill0
// zeroes and mg_boundary_extrap mirrors the ghosts, mg.F90:735-771).  Operand order follows
// mg_poisson.F90, restriction.F90 and interpolation.F90 exactly, so results match a sequential run.
#pragma once
#include "amr_kernels.h"

namespace sg {
namespace amr {

constexpr int kBaseMaxLev = 12;

struct BaseMG {
    int nlev;                       // levels 0..nlev-1, n[l] = n0 >> l, last is 1
    int n[kBaseMaxLev];
    double h[kBaseMaxLev];
    double *U[kBaseMaxLev], *RHS[kBaseMaxLev], *RES[kBaseMaxLev], *RHO[kBaseMaxLev];
    double *psi, *src;              // level-0 potential (in/out) and source
    int lds_off[kBaseMaxLev][4];    // LDS offset of level l's U, RHS, RES, RHO, or -1 when in global memory
    int dirichlet;
    double tol;                     // Resmaxg / rhomax stopping ratio
    int maxit;
    int *iters;                     // out: iterations taken
};

__device__ __forceinline__ int bgi(int n, int i, int j, int k)
{
    return ((k + 1) * (n + 2) + (j + 1)) * (n + 2) + (i + 1);
}

__device__ __forceinline__ int ilog2_pow2(int n) { return 31 - __clz(n); }

__device__ __forceinline__ void wg_sync() { __syncthreads(); }

// wrap (periodic) ghost layers x, then y, then z -- edges and corners wrap too
__device__ inline void bmg_bu(double *u, int n, int dirichlet)
{
    if (dirichlet) return;
    const int W = n + 2;
    for (int t = threadIdx.x; t < W * W; t += blockDim.x) {
        int j = t % W - 1, k = t / W - 1;
        u[bgi(n, -1, j, k)] = u[bgi(n, n - 1, j, k)];
        u[bgi(n, n, j, k)] = u[bgi(n, 0, j, k)];
    }
    wg_sync();
    for (int t = threadIdx.x; t < W * W; t += blockDim.x) {
        int i = t % W - 1, k = t / W - 1;
        u[bgi(n, i, -1, k)] = u[bgi(n, i, n - 1, k)];
        u[bgi(n, i, n, k)] = u[bgi(n, i, 0, k)];
    }
    wg_sync();
    for (int t = threadIdx.x; t < W * W; t += blockDim.x) {
        int i = t % W - 1, j = t / W - 1;
        u[bgi(n, i, j, -1)] = u[bgi(n, i, j, n - 1)];
        u[bgi(n, i, j, n)] = u[bgi(n, i, j, 0)];
    }
    wg_sync();
}

__device__ inline void bmg_fill0(double *u, int n, int dirichlet)
{
    if (!dirichlet) return;
    const int W = n + 2;
    for (int t = threadIdx.x; t < W * W; t += blockDim.x) {
        int a = t % W - 1, c = t / W - 1;
        u[bgi(n, -1, a, c)] = 0.0; u[bgi(n, n, a, c)] = 0.0;
        u[bgi(n, a, -1, c)] = 0.0; u[bgi(n, a, n, c)] = 0.0;
        u[bgi(n, a, c, -1)] = 0.0; u[bgi(n, a, c, n)] = 0.0;
    }
    wg_sync();
}

__device__ inline void bmg_extrap(double *u, int n, int dirichlet)
{
    if (!dirichlet) return;
    const int W = n + 2;
    for (int t = threadIdx.x; t < W * W; t += blockDim.x) {
        int j = t % W - 1, k = t / W - 1;
        u[bgi(n, -1, j, k)] = -u[bgi(n, 0, j, k)];
        u[bgi(n, n, j, k)] = -u[bgi(n, n - 1, j, k)];
    }
    wg_sync();
    for (int t = threadIdx.x; t < W * W; t += blockDim.x) {
        int i = t % W - 1, k = t / W - 1;
        u[bgi(n, i, -1, k)] = -u[bgi(n, i, 0, k)];
        u[bgi(n, i, n, k)] = -u[bgi(n, i, n - 1, k)];
    }
    wg_sync();
    for (int t = threadIdx.x; t < W * W; t += blockDim.x) {
        int i = t % W - 1, j = t / W - 1;
        u[bgi(n, i, j, -1)] = -u[bgi(n, i, j, 0)];
        u[bgi(n, i, j, n)] = -u[bgi(n, i, j, n - 1)];
    }
    wg_sync();
}

// workgroup max of non-negative per-thread values, returned to every thread
__device__ inline double wg_max(double v)
{
    __shared__ double part[16];
    __shared__ double res;
    v = wave_max(v);
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave, nw = (blockDim.x + kWave - 1) / kWave;
    if (lane == 0) part[wid] = v;
    wg_sync();
    if (threadIdx.x == 0) {
        double a = part[0];
        for (int w = 1; w < nw; ++w) a = fmax(a, part[w]);
        res = a;
    }
    wg_sync();
    double r = res;
    wg_sync();
    return r;
}

// mg_poisson_relax (mg_poisson.F90:32-79); returns Resmaxg = max|resh2| / h^2 when want_max
// (the workgroup reduction costs three barriers, and only the last sweep's value is used).
// Each thread loads the stencils of up to four cells of the colour before it updates any of them:
// cells of one colour are never face neighbours, so the loads see exactly the values a sequential
// sweep reads, and four cells' latencies overlap.
__device__ inline double bmg_relax(double *u, const double *__restrict__ rhs, int n, double h, int dirichlet,
                                  bool want_max)
{
    const double sixth = 1.0 / 6.0, h2 = h * h;
    const int hn = (n + 1) / 2, ncell = n * n * hn, lg = ilog2_pow2(n), lgh = ilog2_pow2(hn);
    double m = 0.0;
    for (int ipass = 1; ipass <= 2; ++ipass) {
        bmg_bu(u, n, dirichlet);
        for (int t0 = threadIdx.x; t0 < ncell; t0 += 4 * blockDim.x) {
            double v[4][8];
            int q[4];
            bool ok[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                int t = t0 + c * blockDim.x;
                int ih = t & (hn - 1), j = (t >> lgh) & (n - 1), k = t >> (lgh + lg);
                int i = 2 * ih + ((j + k + ipass) & 1);
                ok[c] = t < ncell && i < n;
                if (!ok[c]) { i = 0; j = 0; k = 0; }
                q[c] = bgi(n, i, j, k);
                v[c][0] = u[bgi(n, i + 1, j, k)];
                v[c][1] = u[bgi(n, i - 1, j, k)];
                v[c][2] = u[bgi(n, i, j + 1, k)];
                v[c][3] = u[bgi(n, i, j - 1, k)];
                v[c][4] = u[bgi(n, i, j, k + 1)];
                v[c][5] = u[bgi(n, i, j, k - 1)];
                v[c][6] = u[q[c]];
                v[c][7] = rhs[q[c]];
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                if (!ok[c]) continue;
                double r = v[c][0] + v[c][1];
                r = r + v[c][2];
                r = r + v[c][3];
                r = r + v[c][4];
                r = r + v[c][5];
                r = r - 6.0 * v[c][6];
                r = r - h2 * v[c][7];
                u[q[c]] = v[c][6] + sixth * r;
                m = fmax(m, fabs(r));
            }
        }
        wg_sync();
    }
    return want_max ? wg_max(m) / h2 : 0.0;
}

// mg_poisson_resid (mg_poisson.F90:5-28)
__device__ inline void bmg_resid(double *__restrict__ res, double *u, const double *__restrict__ rhs, int n, double h,
                                 int dirichlet)
{
    const double h2i = 1.0 / (h * h);
    const int lg = ilog2_pow2(n);
    bmg_bu(u, n, dirichlet);
    for (int t = threadIdx.x; t < n * n * n; t += blockDim.x) {
        int i = t & (n - 1), j = (t >> lg) & (n - 1), k = t >> (2 * lg);
        double s = u[bgi(n, i + 1, j, k)] + u[bgi(n, i - 1, j, k)];
        s = s + u[bgi(n, i, j + 1, k)];
        s = s + u[bgi(n, i, j - 1, k)];
        s = s + u[bgi(n, i, j, k + 1)];
        s = s + u[bgi(n, i, j, k - 1)];
        s = s - 6.0 * u[bgi(n, i, j, k)];
        res[bgi(n, i, j, k)] = -(h2i * s) + rhs[bgi(n, i, j, k)];
    }
    wg_sync();
}

__device__ inline void bmg_rstrct(double *__restrict__ uc, const double *__restrict__ uf, int nc)
{
    const int nf = 2 * nc, lg = ilog2_pow2(nc);
    for (int t = threadIdx.x; t < nc * nc * nc; t += blockDim.x) {
        int ic = t & (nc - 1), jc = (t >> lg) & (nc - 1), kc = t >> (2 * lg);
        int i = 2 * ic, j = 2 * jc, k = 2 * kc;
        double s = uf[bgi(nf, i, j, k)] + uf[bgi(nf, i + 1, j, k)];
        s = s + uf[bgi(nf, i, j + 1, k)];
        s = s + uf[bgi(nf, i, j, k + 1)];
        s = s + uf[bgi(nf, i + 1, j + 1, k)];
        s = s + uf[bgi(nf, i + 1, j, k + 1)];
        s = s + uf[bgi(nf, i, j + 1, k + 1)];
        s = s + uf[bgi(nf, i + 1, j + 1, k + 1)];
        uc[bgi(nc, ic, jc, kc)] = s * (1.0 / 8.0);
    }
    wg_sync();
}

// mg_interp (mg.F90:608-633).  The cubic Lagrange stencils depend only on the fine index, so
// they are formed once per call into LDS (same function, same values) instead of per cell.
constexpr int kBaseMaxN = 64;
__device__ inline void bmg_interp(double *uf, double *uc, int nf, int cubic, int dirichlet)
{
    __shared__ int s_idx[kBaseMaxN][4];
    __shared__ double s_a[kBaseMaxN][4];
    const int nc = nf / 2 > 0 ? nf / 2 : 1;
    const bool use_cubic = cubic && nc + 2 >= 4;
    bmg_extrap(uc, nc, dirichlet);
    bmg_bu(uc, nc, dirichlet);
    if (use_cubic) {
        for (int f = threadIdx.x; f < nf; f += blockDim.x) cubic_stencil(f, -1, nc, s_idx[f], s_a[f]);
        wg_sync();
    }
    auto cget = [&](int a, int b, int c) { return uc[bgi(nc, a, b, c)]; };
    const int lg = ilog2_pow2(nf);
    for (int t = threadIdx.x; t < nf * nf * nf; t += blockDim.x) {
        int i = t & (nf - 1), j = (t >> lg) & (nf - 1), k = t >> (2 * lg);
        double v;
        if (use_cubic) {
            double bxy[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                double bx[4];
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    bx[b] = comb4(s_a[i], cget(s_idx[i][0], s_idx[j][b], s_idx[k][c]),
                                  cget(s_idx[i][1], s_idx[j][b], s_idx[k][c]), cget(s_idx[i][2], s_idx[j][b], s_idx[k][c]),
                                  cget(s_idx[i][3], s_idx[j][b], s_idx[k][c]));
                bxy[c] = comb4(s_a[j], bx[0], bx[1], bx[2], bx[3]);
            }
            v = comb4(s_a[k], bxy[0], bxy[1], bxy[2], bxy[3]);
        } else {
            v = trilinear_at(i, j, k, cget);
        }
        uf[bgi(nf, i, j, k)] = v;
    }
    wg_sync();
    bmg_fill0(uf, nf, dirichlet);
    bmg_fill0(uc, nc, dirichlet);
    bmg_bu(uf, nf, dirichlet);
    bmg_bu(uc, nc, dirichlet);
}

__device__ inline void bmg_zero(double *u, int n)
{
    long len = (long)(n + 2) * (n + 2) * (n + 2);
    for (long t = threadIdx.x; t < len; t += blockDim.x) u[t] = 0.0;
    wg_sync();
}

__device__ inline void bmg_copy(double *o, const double *a, int n)
{
    long len = (long)(n + 2) * (n + 2) * (n + 2);
    for (long t = threadIdx.x; t < len; t += blockDim.x) o[t] = a[t];
    wg_sync();
}

__device__ inline void bmg_add(double *o, const double *a, int n)
{
    long len = (long)(n + 2) * (n + 2) * (n + 2);
    for (long t = threadIdx.x; t < len; t += blockDim.x) o[t] = o[t] + a[t];
    wg_sync();
}

// Level arrays: a = 0 U, 1 RHS, 2 RES, 3 RHO; levels with an LDS offset live in LDS.
__device__ __forceinline__ double *lv(const BaseMG &M, double *lds, int l, int a)
{
    const int o = M.lds_off[l][a];
    if (o >= 0) return lds + o;
    return a == 0 ? M.U[l] : (a == 1 ? M.RHS[l] : (a == 2 ? M.RES[l] : M.RHO[l]));
}

// mg_lin (mg.F90:106-218), Ncycle = 2, Npre = Npost = 2; returns the last sweep's Resmaxg
__device__ inline double bmg_lin(const BaseMG &M, double *lds)
{
    const int lmax = M.nlev - 1, bc = M.dirichlet;
    double resmaxg = 0.0;
    for (int l = 1; l <= lmax; ++l) bmg_rstrct(lv(M, lds, l, 3), lv(M, lds, l - 1, 3), M.n[l]);
    bmg_zero(lv(M, lds, lmax, 0), M.n[lmax]);
    for (int it = 0; it < 2; ++it) {
        double r = bmg_relax(lv(M, lds, lmax, 0), lv(M, lds, lmax, 3), M.n[lmax], M.h[lmax], bc, lmax == 0 && it == 1);
        if (lmax == 0 && it == 1) resmaxg = r;
    }
    for (int ml = lmax - 1; ml >= 0; --ml) {
        bmg_interp(lv(M, lds, ml, 0), lv(M, lds, ml + 1, 0), M.n[ml], 1, bc);
        bmg_copy(lv(M, lds, ml, 1), lv(M, lds, ml, 3), M.n[ml]);
        for (int jc = 0; jc < 2; ++jc) {
            for (int v = ml; v <= lmax - 1; ++v) {
                for (int p = 0; p < 2; ++p) bmg_relax(lv(M, lds, v, 0), lv(M, lds, v, 1), M.n[v], M.h[v], bc, false);
                bmg_resid(lv(M, lds, v, 2), lv(M, lds, v, 0), lv(M, lds, v, 1), M.n[v], M.h[v], bc);
                bmg_rstrct(lv(M, lds, v + 1, 1), lv(M, lds, v, 2), M.n[v + 1]);
                bmg_zero(lv(M, lds, v + 1, 0), M.n[v + 1]);
            }
            for (int it = 0; it < 2; ++it)
                bmg_relax(lv(M, lds, lmax, 0), lv(M, lds, lmax, 1), M.n[lmax], M.h[lmax], bc, false);
            for (int v = lmax - 1; v >= ml; --v) {
                bmg_interp(lv(M, lds, v, 2), lv(M, lds, v + 1, 0), M.n[v], 0, bc);
                bmg_add(lv(M, lds, v, 0), lv(M, lds, v, 2), M.n[v]);
                for (int p = 0; p < 2; ++p) {
                    bool last = ml == 0 && jc == 1 && v == 0 && p == 1;
                    double r = bmg_relax(lv(M, lds, v, 0), lv(M, lds, v, 1), M.n[v], M.h[v], bc, last);
                    if (last) resmaxg = r;
                }
            }
        }
    }
    return resmaxg;
}

// Levels below the finest keep their arrays in LDS as far as kBaseLdsBytes allows (all four on
// levels of at most 8^3 cells, U and RHS on 16^3); the finest level stays in global memory.
constexpr long kBaseLdsBytes = 150 * 1024;
constexpr int kBaseThreads = 1024;

// mg_cycle (mg.F90:42-94) with the reference's stopping test Resmaxg / rhomax < tol.  The final
// residual that mg_cycle's caller never reads is not recomputed.
__global__ void __launch_bounds__(kBaseThreads) k_base_mg(BaseMG M)
{
    extern __shared__ double lds[];
    const int n = M.n[0], bc = M.dirichlet;
    double *rho0 = lv(M, lds, 0, 3);
    bmg_fill0(M.psi, n, bc);
    bmg_bu(M.psi, n, bc);
    double rhomax = 0.0;
    int it = 0;
    for (;;) {
        bmg_resid(rho0, M.psi, M.src, n, M.h[0], bc);
        if (it == 0) {
            double m = 0.0;
            const int lg = ilog2_pow2(n);
            for (int t = threadIdx.x; t < n * n * n; t += blockDim.x) {
                int i = t & (n - 1), j = (t >> lg) & (n - 1), k = t >> (2 * lg);
                m = fmax(m, fabs(rho0[bgi(n, i, j, k)]));
            }
            rhomax = wg_max(m);
        }
        if (it >= M.maxit) break;
        double resmaxg = bmg_lin(M, lds);
        bmg_add(M.psi, lv(M, lds, 0, 0), n);
        ++it;
        if ((rhomax > 0 ? resmaxg / rhomax : 0.0) < M.tol) break;
    }
    if (threadIdx.x == 0 && M.iters) *M.iters = it;
}

// ---- the same mg_cycle spread over the chip: the finest level's steps as multi-workgroup launches
//      (one thread per cell), the stretches on the coarser levels (levels >= 2 in LDS, level 1 in
//      global memory, which crosses launches) as single-workgroup launches over the functions
//      above; same operations in the same order on every value ----

// ghost layers of level array u (n cells per side) in one launch: mode 0 wraps (periodic
// mg_boundary_u: after the x, y, z passes each ghost holds the interior cell with every coordinate
// wrapped), 1 extrapolates (mg_boundary_extrap's x, y, z passes leave (-1)^g times the clamped
// interior cell, g the number of ghost coordinates), 2 zeroes (mg_boundary_fill0)
__global__ void k_bm_ghosts(double *u, int n, int mode)
{
    const int W = n + 2;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)W * W * W) return;
    const int i = (int)(t % W) - 1, j = (int)((t / W) % W) - 1, k = (int)(t / ((long)W * W)) - 1;
    const bool gi_ = i < 0 || i >= n, gj = j < 0 || j >= n, gk = k < 0 || k >= n;
    if (!(gi_ || gj || gk)) return;
    if (mode == 2) { u[t] = 0.0; return; }
    if (mode == 0) {
        u[t] = u[bgi(n, (i + n) % n, (j + n) % n, (k + n) % n)];
        return;
    }
    const int ci = i < 0 ? 0 : (i >= n ? n - 1 : i), cj = j < 0 ? 0 : (j >= n ? n - 1 : j), ck = k < 0 ? 0 : (k >= n ? n - 1 : k);
    const int g = (gi_ ? 1 : 0) + (gj ? 1 : 0) + (gk ? 1 : 0);
    const double v = u[bgi(n, ci, cj, ck)];
    u[t] = (g & 1) ? -v : v;
}

// one colour of bmg_relax; |r| maxima into *resmax (then / h^2 = Resmaxg)
__global__ void k_bm_relax(double *u, const double *__restrict__ rhs, int n, double h, int ipass, double *resmax)
{
    const double sixth = 1.0 / 6.0, h2 = h * h;
    const int hn = (n + 1) / 2, ncell = n * n * hn, lg = ilog2_pow2(n), lgh = ilog2_pow2(hn);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0.0;
    if (t < ncell) {
        const int ih = t & (hn - 1), j = (t >> lgh) & (n - 1), k = t >> (lgh + lg);
        const int i = 2 * ih + ((j + k + ipass) & 1);
        if (i < n) {
            const int q = bgi(n, i, j, k);
            double r = u[bgi(n, i + 1, j, k)] + u[bgi(n, i - 1, j, k)];
            r = r + u[bgi(n, i, j + 1, k)];
            r = r + u[bgi(n, i, j - 1, k)];
            r = r + u[bgi(n, i, j, k + 1)];
            r = r + u[bgi(n, i, j, k - 1)];
            const double uc = u[q];
            r = r - 6.0 * uc;
            r = r - h2 * rhs[q];
            u[q] = uc + sixth * r;
            m = fabs(r);
        }
    }
    if (resmax) block_atomic_max(resmax, m);
}

// bmg_resid without its boundary call; with amax, max |res| over the cells into *amax
__global__ void k_bm_resid(double *__restrict__ res, const double *__restrict__ u, const double *__restrict__ rhs, int n,
                           double h, double *amax)
{
    const double h2i = 1.0 / (h * h);
    const int lg = ilog2_pow2(n);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0.0;
    if (t < n * n * n) {
        const int i = t & (n - 1), j = (t >> lg) & (n - 1), k = t >> (2 * lg);
        double s = u[bgi(n, i + 1, j, k)] + u[bgi(n, i - 1, j, k)];
        s = s + u[bgi(n, i, j + 1, k)];
        s = s + u[bgi(n, i, j - 1, k)];
        s = s + u[bgi(n, i, j, k + 1)];
        s = s + u[bgi(n, i, j, k - 1)];
        s = s - 6.0 * u[bgi(n, i, j, k)];
        const double r = -(h2i * s) + rhs[bgi(n, i, j, k)];
        res[bgi(n, i, j, k)] = r;
        m = fabs(r);
    }
    if (amax) block_atomic_max(amax, m);
}

__global__ void k_bm_rstrct(double *__restrict__ uc, const double *__restrict__ uf, int nc)
{
    const int nf = 2 * nc, lg = ilog2_pow2(nc);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nc * nc * nc) return;
    const int ic = t & (nc - 1), jc = (t >> lg) & (nc - 1), kc = t >> (2 * lg);
    const int i = 2 * ic, j = 2 * jc, k = 2 * kc;
    double s = uf[bgi(nf, i, j, k)] + uf[bgi(nf, i + 1, j, k)];
    s = s + uf[bgi(nf, i, j + 1, k)];
    s = s + uf[bgi(nf, i, j, k + 1)];
    s = s + uf[bgi(nf, i + 1, j + 1, k)];
    s = s + uf[bgi(nf, i + 1, j, k + 1)];
    s = s + uf[bgi(nf, i, j + 1, k + 1)];
    s = s + uf[bgi(nf, i + 1, j + 1, k + 1)];
    uc[bgi(nc, ic, jc, kc)] = s * (1.0 / 8.0);
}

// the interpolation of bmg_interp (its boundary calls are separate launches of k_bm_ghosts)
__global__ void k_bm_interp(double *uf, const double *uc, int nf, int cubic)
{
    const int nc = nf / 2 > 0 ? nf / 2 : 1;
    const bool use_cubic = cubic && nc + 2 >= 4;
    const int lg = ilog2_pow2(nf);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nf * nf * nf) return;
    const int i = t & (nf - 1), j = (t >> lg) & (nf - 1), k = t >> (2 * lg);
    auto cget = [&](int a, int b, int c) { return uc[bgi(nc, a, b, c)]; };
    double v;
    if (use_cubic) {
        int ix[4], jx[4], kx[4];
        double ax[4], ay[4], az[4];
        cubic_stencil(i, -1, nc, ix, ax);
        cubic_stencil(j, -1, nc, jx, ay);
        cubic_stencil(k, -1, nc, kx, az);
        double bxy[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            double bx[4];
#pragma unroll
            for (int b = 0; b < 4; ++b)
                bx[b] = comb4(ax, cget(ix[0], jx[b], kx[c]), cget(ix[1], jx[b], kx[c]), cget(ix[2], jx[b], kx[c]),
                              cget(ix[3], jx[b], kx[c]));
            bxy[c] = comb4(ay, bx[0], bx[1], bx[2], bx[3]);
        }
        v = comb4(az, bxy[0], bxy[1], bxy[2], bxy[3]);
    } else {
        v = trilinear_at(i, j, k, cget);
    }
    uf[bgi(nf, i, j, k)] = v;
}

// whole ghosted arrays: op 0 o = 0, 1 o = a, 2 o = o + a
__global__ void k_bm_blas(double *o, const double *a, long len, int op)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= len) return;
    o[t] = op == 0 ? 0.0 : (op == 1 ? a[t] : o[t] + a[t]);
}

// mg_lin's stretches on the levels >= 1 in one workgroup (level 1 in global memory):
// part 0: the restrictions of RHO to levels 2..lmax, the bottom relaxation and the FMG stages
// ml = lmax-1 .. 1 (mg.F90:130-215 before the finest stage); part 1: one V-cycle's visit of the
// levels below the finest: zero U(1), down from level 1, bottom, up to level 1
__global__ void __launch_bounds__(kBaseThreads) k_bm_coarse(BaseMG M, int part)
{
    extern __shared__ double lds[];
    const int lmax = M.nlev - 1, bc = M.dirichlet;
    if (part == 0) {
        for (int l = 2; l <= lmax; ++l) bmg_rstrct(lv(M, lds, l, 3), lv(M, lds, l - 1, 3), M.n[l]);
        bmg_zero(lv(M, lds, lmax, 0), M.n[lmax]);
        for (int it = 0; it < 2; ++it) bmg_relax(lv(M, lds, lmax, 0), lv(M, lds, lmax, 3), M.n[lmax], M.h[lmax], bc, false);
        for (int ml = lmax - 1; ml >= 1; --ml) {
            bmg_interp(lv(M, lds, ml, 0), lv(M, lds, ml + 1, 0), M.n[ml], 1, bc);
            bmg_copy(lv(M, lds, ml, 1), lv(M, lds, ml, 3), M.n[ml]);
            for (int jc = 0; jc < 2; ++jc) {
                for (int v = ml; v <= lmax - 1; ++v) {
                    for (int p = 0; p < 2; ++p) bmg_relax(lv(M, lds, v, 0), lv(M, lds, v, 1), M.n[v], M.h[v], bc, false);
                    bmg_resid(lv(M, lds, v, 2), lv(M, lds, v, 0), lv(M, lds, v, 1), M.n[v], M.h[v], bc);
                    bmg_rstrct(lv(M, lds, v + 1, 1), lv(M, lds, v, 2), M.n[v + 1]);
                    bmg_zero(lv(M, lds, v + 1, 0), M.n[v + 1]);
                }
                for (int it = 0; it < 2; ++it)
                    bmg_relax(lv(M, lds, lmax, 0), lv(M, lds, lmax, 1), M.n[lmax], M.h[lmax], bc, false);
                for (int v = lmax - 1; v >= ml; --v) {
                    bmg_interp(lv(M, lds, v, 2), lv(M, lds, v + 1, 0), M.n[v], 0, bc);
                    bmg_add(lv(M, lds, v, 0), lv(M, lds, v, 2), M.n[v]);
                    for (int p = 0; p < 2; ++p) bmg_relax(lv(M, lds, v, 0), lv(M, lds, v, 1), M.n[v], M.h[v], bc, false);
                }
            }
        }
        return;
    }
    bmg_zero(lv(M, lds, 1, 0), M.n[1]);
    for (int v = 1; v <= lmax - 1; ++v) {
        for (int p = 0; p < 2; ++p) bmg_relax(lv(M, lds, v, 0), lv(M, lds, v, 1), M.n[v], M.h[v], bc, false);
        bmg_resid(lv(M, lds, v, 2), lv(M, lds, v, 0), lv(M, lds, v, 1), M.n[v], M.h[v], bc);
        bmg_rstrct(lv(M, lds, v + 1, 1), lv(M, lds, v, 2), M.n[v + 1]);
        bmg_zero(lv(M, lds, v + 1, 0), M.n[v + 1]);
    }
    for (int it = 0; it < 2; ++it) bmg_relax(lv(M, lds, lmax, 0), lv(M, lds, lmax, 1), M.n[lmax], M.h[lmax], bc, false);
    for (int v = lmax - 1; v >= 1; --v) {
        bmg_interp(lv(M, lds, v, 2), lv(M, lds, v + 1, 0), M.n[v], 0, bc);
        bmg_add(lv(M, lds, v, 0), lv(M, lds, v, 2), M.n[v]);
        for (int p = 0; p < 2; ++p) bmg_relax(lv(M, lds, v, 0), lv(M, lds, v, 1), M.n[v], M.h[v], bc, false);
    }
}

} // namespace amr
} // namespace sg
