// Kernels of the batched uniform solve: many independent periodic n^3 problems (n = 2^a or
// 3 * 2^a, each with its own h) advanced together, one launch per step of the multigrid schedule.
//
// A launch carries a list of items (problem p, level l, first workgroup, workgroup count), sorted
// by first workgroup; every workgroup finds its item by binary search, so one launch covers the
// same step of all problems whatever their sizes.  Per-(problem, level) data -- array pointers by
// code, n, h -- sit in a device table indexed p * kBLev + l.  A per-problem activity flag lets a
// problem that has converged drop out of every later launch without changing the launch sequence
// (which can then be replayed as a graph).
//
// Per-cell arithmetic is that of the single-grid kernels (relax_value, resid_value,
// restrict_value, trilinear/tricubic values, the fused sweep and cubic tile bodies, the coarse
// kernel's level operations), so a power-of-two problem solved in a batch has the same values as
// solved alone.  Index arithmetic is general (division by n through a multiply-high reciprocal):
// n need not be a power of two.
#pragma once
#include "exact_planes.h"
#include "fast_kernels.h"

namespace sg {
namespace batch {

constexpr int kMaxProb = 4096;
constexpr int kBLev = 12;        // levels per problem (n <= 1024, so n^3 < 2^31)
constexpr int kThreads = 256;    // workgroup of the cell-per-thread kernels

// Codes of the per-level arrays.  FP64 plain path: U, RHS (shared with RHO), RES.  FP32 mixed
// path: U, U1 (second buffer of the fused sweeps), RHS (shared with RHO).
enum { CU = 0, CU1 = 1, CRHS = 2, CRES = 3, kCodes = 4 };

// t / d for 0 <= t < 2^31 as (umulhi(t, m) + t) >> s, m = floor(2^32 (2^s - d) / d) + 1,
// s = ceil(log2 d) (Granlund & Montgomery, "Division by invariant integers using multiplication").
struct FastDiv {
    unsigned d, m;
    int s;
};

inline FastDiv make_div(unsigned d)
{
    FastDiv f{d, 0u, 0};
    while ((1u << f.s) < d) ++f.s;
    f.m = (unsigned)((((unsigned long long)1 << 32) * (((unsigned long long)1 << f.s) - d)) / d + 1);
    return f;
}

__device__ __forceinline__ unsigned fdiv(unsigned t, const FastDiv &f) { return (__umulhi(t, f.m) + t) >> f.s; }

template <typename T> struct PL {
    T *a[kCodes];
    int n;        // cells per side
    int lc;       // mixed path: first level held by the coarse kernel (problem-wide)
    FastDiv dn;   // division by n
    FastDiv dnh;  // division by n / 2 (n even)
    T h2, h2i;
};

struct Item {
    int p, l, blk0, nblk;
};

template <typename T> struct BL {
    const Item *items;
    int nitems;
    const PL<T> *pl;
    const int *active;
};

__device__ __forceinline__ int find_item(const Item *it, int n, int b)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (it[mid].blk0 <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Workgroup b of a grid of 8 * ceil(total / 8) workgroups, dealt round-robin over the 8 XCDs:
// XCD x works through the contiguous range [x * per, (x + 1) * per) of the logical workgroups, so
// neighbouring tiles of a problem share one L2.  Returns -1 for the padding.
__device__ __forceinline__ int xcd_block(int total)
{
    const int per = (int)(gridDim.x / 8);
    const int b = (int)blockIdx.x;
    const int v = (b % 8) * per + b / 8;
    return v < total ? v : -1;
}

template <typename T> struct Cell {
    const Item *it;
    const PL<T> *q;
    int lb;  // workgroup index inside the item
};

template <typename T> __device__ __forceinline__ bool locate(const BL<T> &L, int b, Cell<T> &c)
{
    if (b < 0) return false;
    const Item &it = L.items[find_item(L.items, L.nitems, b)];
    if (!L.active[it.p]) return false;
    c.it = &it;
    c.q = L.pl + it.p * kBLev + it.l;
    c.lb = b - it.blk0;
    return true;
}

__device__ __forceinline__ int wrp(int i, int n) { return i + 1 == n ? 0 : i + 1; }
__device__ __forceinline__ int wrm(int i, int n) { return i == 0 ? n - 1 : i - 1; }

// mg_poisson.F90:55-74, one colour of level l of every item (ipass 1: i+j+k odd, 2: even; a 1^3
// level's single cell belongs to pass 2).  n even or 1.
template <typename T>
__global__ void __launch_bounds__(kThreads) kb_relax(BL<T> L, int cu, int crhs, int ipass)
{
    Cell<T> c;
    if (!locate(L, (int)blockIdx.x, c)) return;
    const PL<T> &q = *c.q;
    T *u = q.a[cu];
    const T *f = q.a[crhs];
    const int n = q.n;
    const unsigned t = (unsigned)c.lb * kThreads + threadIdx.x;
    T s;
    if (n == 1) {
        if (ipass == 2 && t == 0) u[0] = relax_value<T>(u[0], u[0], u[0], u[0], u[0], u[0], u[0], f[0], q.h2, &s);
        return;
    }
    const unsigned half = (unsigned)n * n * (n / 2);
    if (t >= half) return;
    const unsigned r1 = fdiv(t, q.dnh);
    const int ih = (int)(t - r1 * (unsigned)(n / 2));
    const unsigned k = fdiv(r1, q.dn);
    const int j = (int)(r1 - k * (unsigned)n);
    const int i = 2 * ih + ((j + (int)k + ipass) & 1);
    const int row = ((int)k * n + j) * n, c0 = row + i;
    u[c0] = relax_value<T>(u[c0], u[row + wrp(i, n)], u[row + wrm(i, n)], u[((int)k * n + wrp(j, n)) * n + i],
                           u[((int)k * n + wrm(j, n)) * n + i], u[(wrp((int)k, n) * n + j) * n + i],
                           u[(wrm((int)k, n) * n + j) * n + i], f[c0], q.h2, &s);
}

__device__ __forceinline__ void decode(unsigned t, const FastDiv &dn, int n, int &i, int &j, int &k)
{
    const unsigned r1 = fdiv(t, dn);
    i = (int)(t - r1 * (unsigned)n);
    const unsigned r2 = fdiv(r1, dn);
    j = (int)(r1 - r2 * (unsigned)n);
    k = (int)r2;
}

// mg_poisson.F90:17-27 on level l: res = -(h2i (sum6 - 6u)) + rhs.
template <typename T>
__global__ void __launch_bounds__(kThreads) kb_resid(BL<T> L, int cres, int cu, int crhs)
{
    Cell<T> c;
    if (!locate(L, (int)blockIdx.x, c)) return;
    const PL<T> &q = *c.q;
    const int n = q.n;
    const unsigned t = (unsigned)c.lb * kThreads + threadIdx.x;
    if (t >= (unsigned)n * n * n) return;
    int i, j, k;
    decode(t, q.dn, n, i, j, k);
    const T *u = q.a[cu];
    const int row = (k * n + j) * n;
    q.a[cres][t] = resid_value<T>(u[t], u[row + wrp(i, n)], u[row + wrm(i, n)], u[(k * n + wrp(j, n)) * n + i],
                                  u[(k * n + wrm(j, n)) * n + i], u[(wrp(k, n) * n + j) * n + i],
                                  u[(wrm(k, n) * n + j) * n + i], q.a[crhs][t], q.h2i);
}

// restriction.F90:132-147: item level l is the coarse level, its source level l - 1.
template <typename T>
__global__ void __launch_bounds__(kThreads) kb_restrict(BL<T> L, int cdst, int csrc)
{
    Cell<T> c;
    if (!locate(L, (int)blockIdx.x, c)) return;
    const PL<T> &q = *c.q, &qf = *(c.q - 1);
    const int nc = q.n, nf = qf.n;
    const unsigned t = (unsigned)c.lb * kThreads + threadIdx.x;
    if (t >= (unsigned)nc * nc * nc) return;
    int ic, jc, kc;
    decode(t, q.dn, nc, ic, jc, kc);
    const T *uf = qf.a[csrc];
    const int i = 2 * ic, j = 2 * jc, k = 2 * kc;
    auto F = [&](int a, int b, int d) { return uf[((k + d) * nf + (j + b)) * nf + (i + a)]; };
    q.a[cdst][t] = restrict_value<T>(F(0, 0, 0), F(1, 0, 0), F(0, 1, 0), F(0, 0, 1), F(1, 1, 0), F(1, 0, 1),
                                     F(0, 1, 1), F(1, 1, 1));
}

// mg_interp (mg.F90:608-633) from level l + 1 into level l: tricubic when CUBIC (the caller
// chooses it only for coarse grids with nc + 2 >= 4), trilinear otherwise; ADD: uf += P uc
// (mg_addint, mg.F90:727-731).
template <typename T, bool CUBIC, bool ADD>
__global__ void __launch_bounds__(kThreads) kb_interp(BL<T> L, int cdst, int csrc)
{
    Cell<T> c;
    if (!locate(L, (int)blockIdx.x, c)) return;
    const PL<T> &q = *c.q, &qc = *(c.q + 1);
    const int nf = q.n, nc = qc.n;
    const unsigned t = (unsigned)c.lb * kThreads + threadIdx.x;
    if (t >= (unsigned)nf * nf * nf) return;
    int i, j, k;
    decode(t, q.dn, nf, i, j, k);
    const T *uc = qc.a[csrc];
    auto C = [&](int a, int b, int d) { return uc[(d * nc + b) * nc + a]; };
    const T v = CUBIC ? tricubic_point<T>(C, i, j, k, nc) : trilinear_point<T>(C, i, j, k, nc);
    T *uf = q.a[cdst];
    uf[t] = ADD ? uf[t] + v : v;
}

template <typename T>
__global__ void __launch_bounds__(kThreads) kb_fill0(BL<T> L, int code)
{
    Cell<T> c;
    if (!locate(L, (int)blockIdx.x, c)) return;
    const PL<T> &q = *c.q;
    const unsigned t = (unsigned)c.lb * kThreads + threadIdx.x;
    if (t < (unsigned)q.n * q.n * q.n) q.a[code][t] = T(0);
}

// Exact periodic solve of a 3^3 coarsest level (one thread per item): u = Q3 (Lambda^+ (Q3 rhs)),
// Q the symmetric orthonormal Hartley basis, Lambda = (lam_i + lam_j) + lam_k with the zero mode
// dropped; every output of an axis pass is accumulated k = 0, 1, 2 in order.  lam3: 3 eigenvalues
// per problem of that level.  The FP64 instantiation rounds as the CPU restatement does.
template <typename T> __device__ __forceinline__ void exact3_pass(T *g, const T *Q, int axis)
{
    for (int b = 0; b < 3; ++b)
        for (int a = 0; a < 3; ++a) {
            T x[3], y[3];
            for (int k = 0; k < 3; ++k) {
                const int q = axis == 0 ? (b * 3 + a) * 3 + k : axis == 1 ? (b * 3 + k) * 3 + a : (k * 3 + b) * 3 + a;
                x[k] = g[q];
            }
            for (int m = 0; m < 3; ++m) {
                T acc = T(0);
                for (int k = 0; k < 3; ++k) acc = acc + Q[m * 3 + k] * x[k];
                y[m] = acc;
            }
            for (int m = 0; m < 3; ++m) {
                const int q = axis == 0 ? (b * 3 + a) * 3 + m : axis == 1 ? (b * 3 + m) * 3 + a : (m * 3 + b) * 3 + a;
                g[q] = y[m];
            }
        }
}

template <typename T> __device__ __forceinline__ void exact3(T *u, const T *rhs, const T *Q, const T *lam)
{
    T g[27];
    for (int c = 0; c < 27; ++c) g[c] = rhs[c];
    for (int a = 0; a < 3; ++a) exact3_pass<T>(g, Q, a);
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) {
                const T den = (lam[i] + lam[j]) + lam[k];
                T &v = g[(k * 3 + j) * 3 + i];
                v = den != T(0) ? v / den : T(0);
            }
    for (int a = 0; a < 3; ++a) exact3_pass<T>(g, Q, a);
    for (int c = 0; c < 27; ++c) u[c] = g[c];
}

template <typename T>
__global__ void kb_exact3(BL<T> L, int cu, int crhs, const double *Q3, const double *lam3)
{
    Cell<T> c;
    if (!locate(L, (int)blockIdx.x, c) || threadIdx.x != 0) return;
    T Q[9], lam[3];
    for (int e = 0; e < 9; ++e) Q[e] = (T)Q3[e];
    for (int e = 0; e < 3; ++e) lam[e] = (T)lam3[3 * c.it->p + e];
    exact3<T>(c.q->a[cu], c.q->a[crhs], Q, lam);
}

// ---- FP64 outer step (mg.F90:84-90 per problem) ---------------------------------------------

struct Outer {
    double *psi[kMaxProb];
    double *alt[kMaxProb];       // second phi buffer (ping-pong of the fused update)
    const double *src[kMaxProb];
    double h2i[kMaxProb];        // 1 / h^2 of the finest level, as the CPU restatement forms it
};

// Norm partials of the source of level-0 items: per workgroup max|f| and sum f^2.
__global__ void __launch_bounds__(kThreads) kb_src_norms(BL<double> L, const Outer *O, double *pmax, double *psum)
{
    Cell<double> c;
    if (!locate(L, (int)blockIdx.x, c)) return;
    const int n = c.q->n;
    const unsigned t = (unsigned)c.lb * kThreads + threadIdx.x;
    double mx = 0.0, sm = 0.0;
    if (t < (unsigned)n * n * n) {
        const double v = O->src[c.it->p][t];
        mx = fabs(v);
        sm = v * v;
    }
    block_max_sum(mx, sm);
    if (threadIdx.x == 0) { pmax[blockIdx.x] = mx; psum[blockIdx.x] = sm; }
}

// pout = pin + e (UPD) and r = f - L pout stored as Te into level 0's RHS, with per-workgroup norm
// partials of r.  pin / pout are psi / alt of the problem by the parity (0: psi -> alt).  Neighbour
// values are pin + e formed on the fly, exactly the values pout receives.  e: level 0 array `ce`
// (U when the problem has no fine level).
template <typename Te, bool UPD>
__global__ void __launch_bounds__(kThreads)
kb_outer_resid(BL<Te> L, const Outer *O, int parity, int ce, double *pmax, double *psum)
{
    Cell<Te> c;
    if (!locate(L, (int)blockIdx.x, c)) return;
    const PL<Te> &q = *c.q;
    const int p = c.it->p, n = q.n;
    const double *pin = parity ? O->alt[p] : O->psi[p];
    double *pout = parity ? O->psi[p] : O->alt[p];
    const Te *e = q.a[q.lc == 0 ? CU : ce];  // a problem held by the coarse kernel alone keeps e in U
    const unsigned t = (unsigned)c.lb * kThreads + threadIdx.x;
    double mx = 0.0, sm = 0.0;
    if (t < (unsigned)n * n * n) {
        int i, j, k;
        decode(t, q.dn, n, i, j, k);
        auto P = [&](int x) -> double { return UPD ? pin[x] + (double)e[x] : pin[x]; };
        const int row = (k * n + j) * n;
        const double pc = P((int)t);
        const double rv = resid_value<double>(pc, P(row + wrp(i, n)), P(row + wrm(i, n)), P((k * n + wrp(j, n)) * n + i),
                                              P((k * n + wrm(j, n)) * n + i), P((wrp(k, n) * n + j) * n + i),
                                              P((wrm(k, n) * n + j) * n + i), O->src[p][t], O->h2i[p]);
        q.a[CRHS][t] = (Te)rv;
        if (UPD) pout[t] = pc;
        mx = fabs(rv);
        sm = rv * rv;
    }
    block_max_sum(mx, sm);
    if (threadIdx.x == 0) { pmax[blockIdx.x] = mx; psum[blockIdx.x] = sm; }
}

// The same outer step marching along z: workgroup b of an item owns a kOzX x kOzY patch of (i, j)
// columns and zc planes of that problem (n % 32 == 0, n % zc == 0).  Each thread forms its cell
// of the next plane (and, on the patch edge, one x and/or y halo cell: the periodic neighbours)
// while the current plane sits in a double-buffered LDS plane with a one-cell halo, the z
// neighbours in registers: every value of pin + e is loaded once per patch instead of once per
// stencil use.  Same operands, in the same order, as kb_outer_resid, so the same values.
constexpr int kOzX = 32, kOzY = 8;

template <typename Te, bool UPD>
__global__ void __launch_bounds__(kOzX * kOzY)
kb_outer_resid_z(BL<Te> L, int total, const Outer *O, int parity, int ce, int zc, double *pmax, double *psum)
{
    constexpr int TX = kOzX, TY = kOzY, SP = TX + 2;
    __shared__ double sp[2][(TY + 2) * SP];
    Cell<Te> c;
    // XCD-aware numbering: neighbouring patches (consecutive in x) share an XCD, so the cache
    // lines an x-halo cell pulls in are the ones its neighbour loads anyway
    const int wb = xcd_block(total);
    if (!locate(L, wb, c)) return;   // uniform per workgroup
    const PL<Te> &q = *c.q;
    const int p = c.it->p, n = q.n;
    const double *pin = parity ? O->alt[p] : O->psi[p];
    double *pout = parity ? O->psi[p] : O->alt[p];
    const Te *e = q.a[q.lc == 0 ? CU : ce];
    const double *f = O->src[p];
    Te *r = q.a[CRHS];
    const int nbx = n / TX, nby = n / TY;
    const int bx = c.lb % nbx, by = (c.lb / nbx) % nby, bz = c.lb / (nbx * nby);
    const int tx = (int)threadIdx.x % TX, ty = (int)threadIdx.x / TX;
    const int i = bx * TX + tx, j = by * TY + ty, z0 = bz * zc;
    auto at = [&](int ii, int jj, int kk) { return (kk * n + jj) * n + ii; };
    auto wrapz = [&](int a) { return a < 0 ? a + n : (a >= n ? a - n : a); };
    auto P = [&](int x) -> double { return UPD ? pin[x] + (double)e[x] : pin[x]; };
    const bool hx = tx == 0 || tx == TX - 1, hy = ty == 0 || ty == TY - 1;
    const int ih = tx == 0 ? wrapz(i - 1) : wrapz(i + 1), jh = ty == 0 ? wrapz(j - 1) : wrapz(j + 1);
    const int sx = tx == 0 ? (ty + 1) * SP : (ty + 1) * SP + TX + 1;
    const int sy = ty == 0 ? tx + 1 : (TY + 1) * SP + tx + 1;
    const int sc = (ty + 1) * SP + tx + 1;
    double zm = P(at(i, j, wrapz(z0 - 1))), zcen = P(at(i, j, z0));
    double vx = hx ? P(at(ih, j, z0)) : 0.0, vy = hy ? P(at(i, jh, z0)) : 0.0;
    double fc = f[at(i, j, z0)];
    double mx = 0.0, sm = 0.0;
    for (int k = z0; k < z0 + zc; ++k) {
        double *s = sp[k & 1];
        s[sc] = zcen;
        if (hx) s[sx] = vx;
        if (hy) s[sy] = vy;
        const int kn = wrapz(k + 1);
        const double zp = P(at(i, j, kn));
        double vxn = 0.0, vyn = 0.0, fn = 0.0;
        if (k + 1 < z0 + zc) {
            if (hx) vxn = P(at(ih, j, kn));
            if (hy) vyn = P(at(i, jh, kn));
            fn = f[at(i, j, kn)];
        }
        const int cc = at(i, j, k);
        __syncthreads();
        const double rv = resid_value<double>(zcen, s[sc + 1], s[sc - 1], s[sc + SP], s[sc - SP], zp, zm, fc, O->h2i[p]);
        r[cc] = (Te)rv;
        if (UPD) pout[cc] = zcen;
        mx = fmax(mx, fabs(rv));
        sm += rv * rv;
        zm = zcen;
        zcen = zp;
        vx = vxn;
        vy = vyn;
        fc = fn;
    }
    block_max_sum(mx, sm);
    if (threadIdx.x == 0) { pmax[wb] = mx; psum[wb] = sm; }
}

// Per-problem second stage of a norm: workgroup w reduces item w's partials in order.
__global__ void __launch_bounds__(1024) kb_norm_final(const Item *items, const int *active, const double *pmax,
                                                     const double *psum, double *out)
{
    const Item it = items[blockIdx.x];
    if (!active[it.p]) return;
    double mx = 0.0, sm = 0.0;
    for (int q = threadIdx.x; q < it.nblk; q += blockDim.x) {
        mx = fmax(mx, pmax[it.blk0 + q]);
        sm += psum[it.blk0 + q];
    }
    block_max_sum(mx, sm);
    if (threadIdx.x == 0) { out[2 * it.p] = mx; out[2 * it.p + 1] = sm; }
}

// ---- fused FP32 inner solve -----------------------------------------------------------------

// The coarse array of level l + 1 that a level-l kernel reads: the coarse kernel's U when l + 1 is
// the problem's first coarse level, else the fused sweeps' current buffer `cur`.
__device__ __forceinline__ const float *coarse_src(const PL<float> &q, int l, int cur)
{
    const PL<float> &qc = *(&q + 1);
    return qc.a[(l + 1 == q.lc) ? CU : cur];
}

// ns fused sweeps of level l of every item (the fused sweep body of the single-grid solver).
template <class TL, int NS, bool ZERO, bool PROLONG, bool RESID, bool FS>
__global__ void __launch_bounds__(TL::NT) kb_sweep(BL<float> L, int total, int cin, int cout, int cuc)
{
    Cell<float> c;
    if (!locate(L, xcd_block(total), c)) return;
    const PL<float> &q = *c.q;
    const int l = c.it->l;
    sweep_tile<TL, NS, ZERO, PROLONG, RESID, FS>(q.a[cin], q.a[cout], q.a[CRHS], q.n, q.h2, q.h2i,
                                                 PROLONG ? coarse_src(q, l, cuc) : nullptr,
                                                 RESID ? (&q + 1)->a[CRHS] : nullptr, nullptr, c.lb);
}

// The streaming form of the same sweep (one launch = one sweep).
template <class SS, bool ZERO, bool PROLONG, bool RESID>
__global__ void __launch_bounds__(SS::NT) kb_stream(BL<float> L, int total, int cin, int cout, int cuc)
{
    Cell<float> c;
    if (!locate(L, xcd_block(total), c)) return;
    const PL<float> &q = *c.q;
    const int l = c.it->l;
    sweep_stream<SS, ZERO, PROLONG, RESID>(q.a[cin], q.a[cout], q.a[CRHS], q.n, q.h2, q.h2i,
                                           PROLONG ? coarse_src(q, l, cuc) : nullptr,
                                           RESID ? (&q + 1)->a[CRHS] : nullptr, nullptr, c.lb);
}

// FMG tricubic prolongation into U of level l from level l + 1 (tile body of the single-grid
// solver).
template <int FX, int FY, int FZ, int NT>
__global__ void __launch_bounds__(NT) kb_interp_cubic_tile(BL<float> L, int total, int cuc)
{
    Cell<float> c;
    if (!locate(L, xcd_block(total), c)) return;
    const PL<float> &q = *c.q;
    interp_cubic_tile<FX, FY, FZ, NT>(q.a[CU], coarse_src(q, c.it->l, cuc), q.n, c.lb);
}

// The same tile body from code csrc of level l + 1 into code cdst of level l, in either precision:
// mg_interp's FMG step (mg.F90:608-633) in the plain FP64 and the unfused FP32 schedules.
template <int FX, int FY, int FZ, int NT, typename T>
__global__ void __launch_bounds__(NT) kb_interp_cubic_tile_codes(BL<T> L, int total, int cdst, int csrc)
{
    Cell<T> c;
    if (!locate(L, xcd_block(total), c)) return;
    const PL<T> &q = *c.q;
    interp_cubic_tile<FX, FY, FZ, NT, T>(q.a[cdst], (c.q + 1)->a[csrc], q.n, c.lb);
}

// Levels of <= 16^3 cells of one problem, resident in one workgroup's LDS (the coarse kernel of
// the single-grid solver, for any even n and a coarsest level of 1^3 or 3^3).
struct CoarseDesc {
    int nlev;
    int n[8];
    float h2[8], h2i[8];
    int off[8];
    FastDiv dn[8], dnh[8];
    float q3[9], lam3[3];  // exact 3^3 solve (coarsest n == 3)
    int exact;             // held level 0 lies below the problem's level 0: exact solve available
    float qx[kCoarseMaxN * kCoarseMaxN], lamx[kCoarseMaxN];  // its Hartley basis and eigenvalues
    const float *qp, *invp; // plane-parallel exact solve of the problem's bottom level (n <= 64):
                            // Hartley basis and 1 / Lambda (device), or null
};

constexpr int kCoarseCellsB = kCoarseCells;  // 16^3 + 8^3 + ... + 1 >= 12^3 + 6^3 + 3^3

__device__ __forceinline__ void cb_relax(float *U, const float *R, int n, const FastDiv &dn, const FastDiv &dnh,
                                         float h2)
{
    float s;
    if (n == 1) {
        // the single cell belongs to the second pass
        if (threadIdx.x == 0) U[0] = relax_value<float>(U[0], U[0], U[0], U[0], U[0], U[0], U[0], R[0], h2, &s);
        __syncthreads();
        return;
    }
    const int nh = n / 2;
    for (int ipass = 1; ipass <= 2; ++ipass) {
        for (int t = threadIdx.x; t < n * n * nh; t += kCoarseThreads) {
            const unsigned r1 = fdiv((unsigned)t, dnh);
            const int ih = t - (int)r1 * nh;
            const unsigned kk = fdiv(r1, dn);
            const int j = (int)r1 - (int)kk * n, k = (int)kk;
            const int i = 2 * ih + ((j + k + ipass) & 1);
            const int row = (k * n + j) * n, c = row + i;
            U[c] = relax_value<float>(U[c], U[row + wrp(i, n)], U[row + wrm(i, n)], U[(k * n + wrp(j, n)) * n + i],
                                      U[(k * n + wrm(j, n)) * n + i], U[(wrp(k, n) * n + j) * n + i],
                                      U[(wrm(k, n) * n + j) * n + i], R[c], h2, &s);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ void cb_resid_restrict(const float *U, const float *R, int n, float h2i, float *Rc,
                                                  const FastDiv &dnc)
{
    const int nc = n >> 1;
    for (int q = threadIdx.x; q < nc * nc * nc; q += kCoarseThreads) {
        int cx, cy, cz;
        decode((unsigned)q, dnc, nc, cx, cy, cz);
        float r[8];
#pragma unroll
        for (int o = 0; o < 8; ++o) {
            const int i = 2 * cx + (o & 1), j = 2 * cy + ((o >> 1) & 1), k = 2 * cz + ((o >> 2) & 1);
            const int row = (k * n + j) * n, c = row + i;
            r[o] = resid_value<float>(U[c], U[row + wrp(i, n)], U[row + wrm(i, n)], U[(k * n + wrp(j, n)) * n + i],
                                      U[(k * n + wrm(j, n)) * n + i], U[(wrp(k, n) * n + j) * n + i],
                                      U[(wrm(k, n) * n + j) * n + i], R[c], h2i);
        }
        Rc[q] = restrict_value<float>(r[0], r[1], r[2], r[4], r[3], r[5], r[6], r[7]);
    }
    __syncthreads();
}

__device__ __forceinline__ void cb_restrict(float *Rc, const float *Rf, int nc, const FastDiv &dnc)
{
    const int nf = 2 * nc;
    for (int q = threadIdx.x; q < nc * nc * nc; q += kCoarseThreads) {
        int ic, jc, kc;
        decode((unsigned)q, dnc, nc, ic, jc, kc);
        const int i = 2 * ic, j = 2 * jc, k = 2 * kc;
        auto F = [&](int a, int b, int d) { return Rf[((k + d) * nf + (j + b)) * nf + (i + a)]; };
        Rc[q] = restrict_value<float>(F(0, 0, 0), F(1, 0, 0), F(0, 1, 0), F(0, 0, 1), F(1, 1, 0), F(1, 0, 1),
                                      F(0, 1, 1), F(1, 1, 1));
    }
    __syncthreads();
}

__device__ __forceinline__ void cb_fill0(float *U, int n)
{
    for (int c = threadIdx.x; c < n * n * n; c += kCoarseThreads) U[c] = 0.0f;
    __syncthreads();
}

__device__ __forceinline__ void cb_interp(float *Uf, const float *Uc, int nf, const FastDiv &dnf, bool cubic, bool add)
{
    const int nc = nf >> 1;
    auto C = [&](int a, int b, int d) { return Uc[(d * nc + b) * nc + a]; };
    for (int c = threadIdx.x; c < nf * nf * nf; c += kCoarseThreads) {
        int i, j, k;
        decode((unsigned)c, dnf, nf, i, j, k);
        float v;
        if (cubic && nc + 2 >= 4) v = tricubic_point<float>(C, i, j, k, nc);
        else v = trilinear_point<float>(C, i, j, k, nc);
        Uf[c] = add ? Uf[c] + v : v;
    }
    __syncthreads();
}

// mg_slvsml (mg.F90:540-546) on the coarsest held level: two sweeps at 1^3, the exact solve at 3^3.
__device__ __forceinline__ void cb_slvsml(float *U, const float *R, const CoarseDesc &D, int l)
{
    if (D.n[l] == 3) {
        if (threadIdx.x == 0) exact3<float>(U, R, D.q3, D.lam3);
        __syncthreads();
    } else {
        for (int it = 0; it < 2; ++it) cb_relax(U, R, D.n[l], D.dn[l], D.dnh[l], D.h2[l]);
    }
}

// One workgroup per item (level l = the problem's first coarse level).  mode 0: FMG over the held
// levels from level 0's RHS (mg.F90:142-215 restricted to them); mode 1: one V-cycle from held
// level 0 with zero initial guess (mg.F90:179-210).  exact: both modes solve held level 0 exactly
// (coarse_exact) for the problems whose descriptor allows it.
__global__ void __launch_bounds__(kCoarseThreads)
kb_coarse(BL<float> L, const CoarseDesc *descs, int mode, int ncycle, int npre, int npost, int exact)
{
    Cell<float> c;
    if (!locate(L, (int)blockIdx.x, c)) return;
    const CoarseDesc &D = descs[c.it->p];
    __shared__ float sU[kCoarseCellsB], sR[kCoarseCellsB];
    const int lmax = D.nlev - 1;
    const int n0 = D.n[0];
    const float *rhs_in = c.q->a[CRHS];
    if (exact && D.exact) {
        coarse_exact(sU, sR, D.qx, D.lamx, rhs_in, c.q->a[CU], n0);
        return;
    }
    for (int q = threadIdx.x; q < n0 * n0 * n0; q += kCoarseThreads) sR[q] = rhs_in[q];
    __syncthreads();
    auto U = [&](int l) { return sU + D.off[l]; };
    auto R = [&](int l) { return sR + D.off[l]; };
    auto vcycle = [&](int ml) {
        for (int v = ml; v <= lmax - 1; ++v) {
            for (int p = 0; p < npre; ++p) cb_relax(U(v), R(v), D.n[v], D.dn[v], D.dnh[v], D.h2[v]);
            cb_resid_restrict(U(v), R(v), D.n[v], D.h2i[v], R(v + 1), D.dn[v + 1]);
            cb_fill0(U(v + 1), D.n[v + 1]);
        }
        cb_slvsml(U(lmax), R(lmax), D, lmax);
        for (int v = lmax - 1; v >= ml; --v) {
            cb_interp(U(v), U(v + 1), D.n[v], D.dn[v], false, true);
            for (int p = 0; p < npost; ++p) cb_relax(U(v), R(v), D.n[v], D.dn[v], D.dnh[v], D.h2[v]);
        }
    };
    if (mode == 0) {
        for (int l = 1; l <= lmax; ++l) cb_restrict(R(l), R(l - 1), D.n[l], D.dn[l]);
        cb_fill0(U(lmax), D.n[lmax]);
        cb_slvsml(U(lmax), R(lmax), D, lmax);
        for (int ml = lmax - 1; ml >= 0; --ml) {
            cb_interp(U(ml), U(ml + 1), D.n[ml], D.dn[ml], true, false);
            for (int jc = 0; jc < ncycle; ++jc) vcycle(ml);
        }
    } else {
        cb_fill0(U(0), n0);
        if (lmax == 0) cb_slvsml(U(0), R(0), D, 0);
        else vcycle(0);
    }
    float *u_out = c.q->a[CU];
    for (int q = threadIdx.x; q < n0 * n0 * n0; q += kCoarseThreads) u_out[q] = sU[q];
}

// Plane-parallel exact solve (exact_planes.h, whole form) of each item's level: workgroup b of an
// item works on plane (k_exact_xy) or row (k_exact_z) b of that problem's level.
__global__ void __launch_bounds__(kPlaneThreads) kb_exact_xy(BL<float> L, int total, const CoarseDesc *descs, int cin,
                                                             int cout)
{
    Cell<float> c;
    if (!locate(L, xcd_block(total), c)) return;
    exact_xy(c.q->a[cin], c.q->a[cout], descs[c.it->p].qp, c.q->n, c.lb);
}

__global__ void __launch_bounds__(kPlaneThreads) kb_exact_z(BL<float> L, int total, const CoarseDesc *descs, int cw)
{
    Cell<float> c;
    if (!locate(L, xcd_block(total), c)) return;
    const CoarseDesc &D = descs[c.it->p];
    exact_z(c.q->a[cw], D.qp, D.invp, c.q->n, c.lb);
}

} // namespace batch
} // namespace sg
