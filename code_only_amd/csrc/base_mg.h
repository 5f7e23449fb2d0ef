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
    int dirichlet;
    double tol;                     // Resmaxg / rhomax stopping ratio
    int maxit;
    int *iters;                     // out: iterations taken
};

__device__ __forceinline__ long bgi(int n, int i, int j, int k)
{
    return ((long)(k + 1) * (n + 2) + (j + 1)) * (n + 2) + (i + 1);
}

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

// mg_poisson_relax (mg_poisson.F90:32-79); returns Resmaxg = max|resh2| / h^2
__device__ inline double bmg_relax(double *u, const double *rhs, int n, double h, int dirichlet)
{
    const double sixth = 1.0 / 6.0, h2 = h * h;
    const int hn = (n + 1) / 2;
    double m = 0.0;
    for (int ipass = 1; ipass <= 2; ++ipass) {
        bmg_bu(u, n, dirichlet);
        for (int t = threadIdx.x; t < n * n * hn; t += blockDim.x) {
            int ih = t % hn, j = (t / hn) % n, k = t / (hn * n);
            int i = 2 * ih + ((j + k + ipass) & 1);
            if (i >= n) continue;
            double r = u[bgi(n, i + 1, j, k)] + u[bgi(n, i - 1, j, k)];
            r = r + u[bgi(n, i, j + 1, k)];
            r = r + u[bgi(n, i, j - 1, k)];
            r = r + u[bgi(n, i, j, k + 1)];
            r = r + u[bgi(n, i, j, k - 1)];
            r = r - 6.0 * u[bgi(n, i, j, k)];
            r = r - h2 * rhs[bgi(n, i, j, k)];
            u[bgi(n, i, j, k)] = u[bgi(n, i, j, k)] + sixth * r;
            m = fmax(m, fabs(r));
        }
        wg_sync();
    }
    return wg_max(m) / h2;
}

// mg_poisson_resid (mg_poisson.F90:5-28)
__device__ inline void bmg_resid(double *res, double *u, const double *rhs, int n, double h, int dirichlet)
{
    const double h2i = 1.0 / (h * h);
    bmg_bu(u, n, dirichlet);
    for (int t = threadIdx.x; t < n * n * n; t += blockDim.x) {
        int i = t % n, j = (t / n) % n, k = t / (n * n);
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

__device__ inline void bmg_rstrct(double *uc, const double *uf, int nc)
{
    const int nf = 2 * nc;
    for (int t = threadIdx.x; t < nc * nc * nc; t += blockDim.x) {
        int ic = t % nc, jc = (t / nc) % nc, kc = t / (nc * nc);
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

// mg_interp (mg.F90:608-633)
__device__ inline void bmg_interp(double *uf, double *uc, int nf, int cubic, int dirichlet)
{
    const int nc = nf / 2 > 0 ? nf / 2 : 1;
    bmg_extrap(uc, nc, dirichlet);
    bmg_bu(uc, nc, dirichlet);
    auto cget = [&](int a, int b, int c) { return uc[bgi(nc, a, b, c)]; };
    for (int t = threadIdx.x; t < nf * nf * nf; t += blockDim.x) {
        int i = t % nf, j = (t / nf) % nf, k = t / (nf * nf);
        uf[bgi(nf, i, j, k)] = (cubic && nc + 2 >= 4) ? tricubic_at(i, j, k, nc, cget) : trilinear_at(i, j, k, cget);
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

// mg_lin (mg.F90:106-218), Ncycle = 2, Npre = Npost = 2; returns the last Resmaxg
__device__ inline double bmg_lin(const BaseMG &M)
{
    const int lmax = M.nlev - 1, bc = M.dirichlet;
    double resmaxg = 0.0;
    for (int l = 1; l <= lmax; ++l) bmg_rstrct(M.RHO[l], M.RHO[l - 1], M.n[l]);
    bmg_zero(M.U[lmax], M.n[lmax]);
    for (int it = 0; it < 2; ++it) resmaxg = bmg_relax(M.U[lmax], M.RHO[lmax], M.n[lmax], M.h[lmax], bc);
    for (int ml = lmax - 1; ml >= 0; --ml) {
        bmg_interp(M.U[ml], M.U[ml + 1], M.n[ml], 1, bc);
        bmg_copy(M.RHS[ml], M.RHO[ml], M.n[ml]);
        for (int jc = 0; jc < 2; ++jc) {
            for (int v = ml; v <= lmax - 1; ++v) {
                for (int p = 0; p < 2; ++p) resmaxg = bmg_relax(M.U[v], M.RHS[v], M.n[v], M.h[v], bc);
                bmg_resid(M.RES[v], M.U[v], M.RHS[v], M.n[v], M.h[v], bc);
                bmg_rstrct(M.RHS[v + 1], M.RES[v], M.n[v + 1]);
                bmg_zero(M.U[v + 1], M.n[v + 1]);
            }
            for (int it = 0; it < 2; ++it) resmaxg = bmg_relax(M.U[lmax], M.RHS[lmax], M.n[lmax], M.h[lmax], bc);
            for (int v = lmax - 1; v >= ml; --v) {
                bmg_interp(M.RES[v], M.U[v + 1], M.n[v], 0, bc);
                bmg_add(M.U[v], M.RES[v], M.n[v]);
                for (int p = 0; p < 2; ++p) resmaxg = bmg_relax(M.U[v], M.RHS[v], M.n[v], M.h[v], bc);
            }
        }
    }
    return resmaxg;
}

// mg_cycle (mg.F90:42-94) with the reference's stopping test Resmaxg / rhomax < tol.
__global__ void __launch_bounds__(1024) k_base_mg(BaseMG M)
{
    const int n = M.n[0], bc = M.dirichlet;
    bmg_fill0(M.psi, n, bc);
    bmg_bu(M.psi, n, bc);
    double rhomax = 0.0;
    int it = 0;
    for (;;) {
        bmg_resid(M.RHO[0], M.psi, M.src, n, M.h[0], bc);
        double m = 0.0;
        for (int t = threadIdx.x; t < n * n * n; t += blockDim.x) {
            int i = t % n, j = (t / n) % n, k = t / (n * n);
            m = fmax(m, fabs(M.RHO[0][bgi(n, i, j, k)]));
        }
        m = wg_max(m);
        if (it == 0) rhomax = m;
        if (it >= M.maxit) break;
        double resmaxg = bmg_lin(M);
        bmg_add(M.psi, M.U[0], n);
        ++it;
        if ((rhomax > 0 ? resmaxg / rhomax : 0.0) < M.tol) {
            bmg_resid(M.RHO[0], M.psi, M.src, n, M.h[0], bc);
            break;
        }
    }
    if (threadIdx.x == 0 && M.iters) *M.iters = it;
}

} // namespace amr
} // namespace sg
