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

} // namespace amr
} // namespace sg
