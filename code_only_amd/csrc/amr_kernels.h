// Device kernels of the block-structured AMR solve (fmg.F90 / vmg.F90 / fmg_poisson.F90 /
// fmg_ghostcell.F90 / fmg_converge.F90 / fmg_reflux.F90 / fmg_boundary*.F90), FP64.
//
// Every block keeps, per FMG level, one array per code of (N+2)^3 doubles (N = 8, 4, 2 cells per
// side plus one ghost layer, i fastest): element (i, j, k), i, j, k in -1..N, sits at
// ((k+1)(N+2) + j+1)(N+2) + i+1.  All arrays of one code and FMG level are contiguous, block b at
// offset b (N+2)^3; blocks are numbered level by level, coarsest first.
//
// Each kernel performs one loop nest of the reference over all blocks it applies to, one thread
// per output element, with the reference's operand order (the library is built without
// multiply-add contraction), so results are bitwise those of a sequential execution.  Face fluxes
// are not stored: every consumer recomputes F(i) = (u(i+1) - u(i)) / h from the block's own values,
// except on coarse faces that meet a finer level, whose fluxes are the 2x2 averages of the finer
// block's face fluxes (fmg_reflux.F90), kept in a small per-face "override" array.
#pragma once
#include "sg_common.h"

namespace sg {
namespace amr {

enum { C_U = 0, C_RES = 1, C_RHS = 2, C_RHO = 3, C_SRC = 4, C_PSI = 5, NCODE = 6 };
enum { NBR_NONE = 0, NBR_SAME = 1, NBR_PARENT = 2 };

template <int N> struct Blk {
    static constexpr int W = N + 2;
    static constexpr int LEN = W * W * W;
    static constexpr int H = N / 2;
    __device__ static constexpr int gi(int i, int j, int k) { return ((k + 1) * W + (j + 1)) * W + (i + 1); }
    // element at coordinate a along d and (t1, t2) along the two other axes in x, y, z order
    __device__ static int rot(int d, int a, int t1, int t2)
    {
        return d == 0 ? gi(a, t1, t2) : (d == 1 ? gi(t1, a, t2) : gi(t1, t2, a));
    }
};

// Workgroup w of G runs on XCD w mod 8 (the dispatcher deals workgroups out round-robin): the
// block it takes under the map that gives each XCD a contiguous range of the G blocks, so that
// blocks next to each other in the list share that XCD's L2.  A bijection of [0, G).
__device__ __forceinline__ int xcd_block(int w, int G)
{
    const int x = w & 7, i = w >> 3, q = G >> 3, r = G & 7;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}

__device__ __forceinline__ int ijkc(int i) { return i >= 0 ? i / 2 : -((1 - i) / 2); }   // floor(i/2)
__device__ __forceinline__ int near_(int f, int c) { return c + 2 * (((f % 2) + 2) % 2) - 1; }

// interpolation.F90:106-127: cubic Lagrange stencil of fine index f over coarse bounds lo..hi.
__device__ __forceinline__ void cubic_stencil(int f, int lo, int hi, int idx[4], double a[4])
{
    int cp = ijkc(f), cn = near_(f, cp);
    int c1 = cp < cn ? cp : cn, c2 = cp < cn ? cn : cp;
    int c0 = 2 * c1 - c2, c3 = 2 * c2 - c1;
    int ns = 0;
    if (c3 > hi) ns = c3 - hi;
    if (c0 < lo) ns = c0 - lo;
    c0 -= ns; c1 -= ns; c2 -= ns; c3 -= ns;
    double xc = cp + 0.25 * (double)(2 * (((f % 2) + 2) % 2) - 1);
    a[0] = (xc - c1) * (xc - c2) * (xc - c3) / (double)((c0 - c1) * (c0 - c2) * (c0 - c3));
    a[1] = (xc - c0) * (xc - c2) * (xc - c3) / (double)((c1 - c0) * (c1 - c2) * (c1 - c3));
    a[2] = (xc - c0) * (xc - c1) * (xc - c3) / (double)((c2 - c0) * (c2 - c1) * (c2 - c3));
    a[3] = (xc - c0) * (xc - c1) * (xc - c2) / (double)((c3 - c0) * (c3 - c1) * (c3 - c2));
    idx[0] = c0; idx[1] = c1; idx[2] = c2; idx[3] = c3;
}

template <typename T> __device__ __forceinline__ T comb4(const double a[4], T v0, T v1, T v2, T v3)
{
    T s = (T)a[0] * v0 + (T)a[1] * v1;
    s = s + (T)a[2] * v2;
    s = s + (T)a[3] * v3;
    return s;
}

struct Cubic64 { double w[64]; };   // St(a, b, c) of restriction.F90:38-63 at w[(c*4 + b)*4 + a]
struct LevelScalars { double v[16]; };   // one value per AMR level, passed by value

// restriction.F90:127-151
template <int N, typename T> __device__ __forceinline__ T rstr_lin(const T *uf, int i, int j, int k)
{
    using B = Blk<N>;
    T s = uf[B::gi(i, j, k)] + uf[B::gi(i + 1, j, k)];
    s = s + uf[B::gi(i, j + 1, k)];
    s = s + uf[B::gi(i, j, k + 1)];
    s = s + uf[B::gi(i + 1, j + 1, k)];
    s = s + uf[B::gi(i + 1, j, k + 1)];
    s = s + uf[B::gi(i, j + 1, k + 1)];
    s = s + uf[B::gi(i + 1, j + 1, k + 1)];
    return s * (T)(1.0 / 8.0);
}

// The inner solve's FMG-level-0 source in compact colour-split form: per block (in the order of
// the active list) the 256 cells of the first colour (i + j + k odd), then the 256 of the second,
// each colour in the order of k_blk8's colour threads, i/2 + 4 j + 32 k.  A colour pass reads its
// 1 KB contiguously instead of picking its cells out of the ghosted 10^3 array.
__device__ __forceinline__ int cc_index(int i, int j, int k)
{
    return ((i + j + k + 1) & 1) * 256 + (i >> 1) + 4 * j + 32 * k;
}

// restriction.F90:127-151 from the compact source: the operand order of rstr_lin
__device__ __forceinline__ float rstr_lin_cc(const float *uc, int i, int j, int k)
{
    float s = uc[cc_index(i, j, k)] + uc[cc_index(i + 1, j, k)];
    s = s + uc[cc_index(i, j + 1, k)];
    s = s + uc[cc_index(i, j, k + 1)];
    s = s + uc[cc_index(i + 1, j + 1, k)];
    s = s + uc[cc_index(i + 1, j, k + 1)];
    s = s + uc[cc_index(i, j + 1, k + 1)];
    s = s + uc[cc_index(i + 1, j + 1, k + 1)];
    return s * (1.0f / 8.0f);
}

// The inner solve's FMG-level-0 correction U in the same compact form (active-list entry a at
// 512 a): the interiors only.  The ghost values its colour passes read are held apart in a ghost
// store (gs_cell): same-level faces written there by the neighbours, coarse-fine faces by the
// refresh, physical faces left zero, the inner solve's boundary value (fmg_boundary_fill0).
__device__ __forceinline__ void cc_ijk(int c, int &i, int &j, int &k)   // inverse of cc_index
{
    const int r = c & 255;
    j = (r >> 2) & 7;
    k = r >> 5;
    i = 2 * (r & 3) + (((c >> 8) + j + k + 1) & 1);
}

// cell (a along axis d, (t1, t2) along the other two in x, y, z order) in the compact form
__device__ __forceinline__ int cc_rot(int d, int a, int t1, int t2)
{
    return d == 0 ? cc_index(a, t1, t2) : (d == 1 ? cc_index(t1, a, t2) : cc_index(t1, t2, a));
}

// restriction.F90:127-151 over the compact interiors, the 2x2x2 cells at origin v = a 512 + i + 8 j
// + 64 k (block a's cells (i, j, k) .. (i + 1, j + 1, k + 1))
__device__ __forceinline__ float rstr_cc(const float *U, int v)
{
    const int r = v & 511;
    return rstr_lin_cc(U + (long)(v >> 9) * 512, r & 7, (r >> 3) & 7, r >> 6);
}

// The ghost store of the compact form: per active block a and face F, the 64 ghost cells at
// (a 6 + F) 64 in two halves by the cell's parity (i + j + k in the block's frame, the ghost
// coordinate -1 or 8 included), each half in (t1 >> 1) + 4 t2 order.  A colour pass reads the half
// its stencils need and its neighbours write the other half (the cells they change), so the pass
// reads one 128-byte line per face and nothing it reads is written during it.
__device__ __forceinline__ int gs_cell(int F, int t1, int t2)
{
    return ((((F & 1) ? 0 : 1) + t1 + t2) & 1) * 32 + (t1 >> 1) + 4 * t2;
}

// NC contiguous floats from / to p (16-, 8- or 4-byte accesses; p aligned to the access)
template <int NC> __device__ __forceinline__ void ld_run(float (&d)[NC], const float *p)
{
    if constexpr (NC % 4 == 0) {
#pragma unroll
        for (int q = 0; q < NC / 4; ++q) {
            const float4 v = ((const float4 *)p)[q];
            d[4 * q] = v.x;
            d[4 * q + 1] = v.y;
            d[4 * q + 2] = v.z;
            d[4 * q + 3] = v.w;
        }
    } else if constexpr (NC == 2) {
        const float2 v = *(const float2 *)p;
        d[0] = v.x;
        d[1] = v.y;
    } else {
#pragma unroll
        for (int q = 0; q < NC; ++q) d[q] = p[q];
    }
}

template <int NC> __device__ __forceinline__ void st_run(float *p, const float (&d)[NC])
{
    if constexpr (NC % 4 == 0) {
#pragma unroll
        for (int q = 0; q < NC / 4; ++q) ((float4 *)p)[q] = make_float4(d[4 * q], d[4 * q + 1], d[4 * q + 2], d[4 * q + 3]);
    } else if constexpr (NC == 2) {
        *(float2 *)p = make_float2(d[0], d[1]);
    } else {
#pragma unroll
        for (int q = 0; q < NC; ++q) p[q] = d[q];
    }
}

// interior cell (i, j, k) with value v into the ghost stores of the same-level neighbours whose
// ghost it is (pt: the block's six ptab entries)
template <typename T> __device__ __forceinline__ void push_faces(T *gs, const int *pt, int i, int j, int k, T v)
{
    const int hp = ((i + j + k) & 1) * 32;
    auto push = [&](int f, int t1, int t2) {
        const int tg = pt[f];
        if (tg >= 0) gs[(long)tg * 64 + hp + (t1 >> 1) + 4 * t2] = v;
    };
    if (i == 0) push(0, j, k);
    if (i == 7) push(1, j, k);
    if (j == 0) push(2, i, k);
    if (j == 7) push(3, i, k);
    if (k == 0) push(4, i, j);
    if (k == 7) push(5, i, j);
}

// restriction.F90:71-119 (stencil i1-1 .. i1+2 per axis)
template <int N, typename T> __device__ __forceinline__ T rstr_cub(const T *uf, int i1, int j1, int k1, const Cubic64 &st)
{
    using B = Blk<N>;
    T s = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                T t = uf[B::gi(i1 - 1 + a, j1 - 1 + b, k1 - 1 + c)] * (T)st.w[(c * 4 + b) * 4 + a];
                s = (a == 0 && b == 0 && c == 0) ? t : s + t;
            }
    return s;
}

// interpolation.F90:222-247 on a coarse array addressed through cget(ic, jc, kc)
template <typename G> __device__ __forceinline__ auto trilinear_at(int i, int j, int k, G cget) -> decltype(cget(0, 0, 0))
{
    using T = decltype(cget(0, 0, 0));
    int ic = ijkc(i), jc = ijkc(j), kc = ijkc(k);
    int ic1 = near_(i, ic), jc1 = near_(j, jc), kc1 = near_(k, kc);
    T b = cget(ic1, jc, kc) + cget(ic, jc1, kc);
    b = b + cget(ic, jc, kc1);
    T e = cget(ic, jc1, kc1) + cget(ic1, jc, kc1);
    e = e + cget(ic1, jc1, kc);
    T s = (T)27.0 * cget(ic, jc, kc) + (T)9.0 * b;
    s = s + (T)3.0 * e;
    s = s + cget(ic1, jc1, kc1);
    return s / (T)64.0;
}

// interpolation.F90:40-218 evaluated for one fine cell: x pass (bufx), y pass (bufxy), z pass,
// each intermediate formed exactly as the three-pass loop forms it.  Coarse bounds -1..nc.
template <typename G> __device__ __forceinline__ auto tricubic_at(int i, int j, int k, int nc, G cget) -> decltype(cget(0, 0, 0))
{
    using T = decltype(cget(0, 0, 0));
    int ix[4], jx[4], kx[4];
    double ax[4], ay[4], az[4];
    cubic_stencil(i, -1, nc, ix, ax);
    cubic_stencil(j, -1, nc, jx, ay);
    cubic_stencil(k, -1, nc, kx, az);
    T bxy[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        T bx[4];
#pragma unroll
        for (int b = 0; b < 4; ++b)
            bx[b] = comb4(ax, cget(ix[0], jx[b], kx[c]), cget(ix[1], jx[b], kx[c]), cget(ix[2], jx[b], kx[c]),
                          cget(ix[3], jx[b], kx[c]));
        bxy[c] = comb4(ay, bx[0], bx[1], bx[2], bx[3]);
    }
    return comb4(az, bxy[0], bxy[1], bxy[2], bxy[3]);
}

// ---- fmg_converge_c2p_lev: each child of [b0, b0+nb) restricted into its parent's octant ----
template <int N, bool CUBIC, typename T>
__global__ void k_c2p(T *A, int b0, int nb, const int *parent, const int *octant, Cubic64 st)
{
    using B = Blk<N>;
    constexpr int H = N / 2, HC = H * H * H;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nb * HC) return;
    int b = b0 + (int)(t / HC), c = (int)(t % HC);
    int ic = c % H, jc = (c / H) % H, kc = c / (H * H);
    const T *uf = A + (long)b * B::LEN;
    T *uc = A + (long)parent[b] * B::LEN;
    int o = octant[b];
    int oi = (o & 1) * H, oj = ((o >> 1) & 1) * H, ok = ((o >> 2) & 1) * H;
    uc[B::gi(oi + ic, oj + jc, ok + kc)] = CUBIC ? rstr_cub<N, T>(uf, 2 * ic, 2 * jc, 2 * kc, st)
                                                 : rstr_lin<N, T>(uf, 2 * ic, 2 * jc, 2 * kc);
}

// ---- one direction of fmg_ghostcell_fix for the blocks [b0, b0+nb) of one AMR level:
//      same-level faces copied (fmg_ghfix_samelev), parent-level faces interpolated
//      (fmg_ghfix_parentlev: bilinear or separable cubic tangentially, quadratic normally) ----
template <int N, typename T>
__global__ void k_ghost(T *A, int b0, int nb, int d, int cubic, int kinds, const int *nbr, const int *nkind,
                        const int *octant)
{
    using B = Blk<N>;
    constexpr int W = B::W, H = N / 2;
    // 4^3 and larger: workgroups over the XCDs in contiguous ranges (xcd_block), neighbour reads
    // then mostly hit the reading XCD's L2
    const long w = N >= 4 ? xcd_block((int)blockIdx.x, (int)gridDim.x) : (long)blockIdx.x;
    long t = w * blockDim.x + threadIdx.x;
    if (t >= (long)nb * 2 * W * W) return;
    int t1 = (int)(t % W) - 1, t2 = (int)((t / W) % W) - 1, lr = (int)((t / (W * W)) & 1);
    int b = b0 + (int)(t / (2 * W * W));
    int f = 2 * d + lr;
    // the three table loads issued together, ahead of the test that may make two of them moot
    const int kind = nkind[b * 6 + f], src = nbr[b * 6 + f], o = octant[b];
    if (kind == NBR_NONE || (kind == NBR_SAME && !(kinds & 1)) || (kind == NBR_PARENT && !(kinds & 2))) return;
    T *ud = A + (long)b * B::LEN;
    const T *us = A + (long)src * B::LEN;
    const int ag = lr ? N : -1;
    if (kind == NBR_SAME) {
        ud[B::rot(d, ag, t1, t2)] = us[B::rot(d, lr ? 0 : N - 1, t1, t2)];
        return;
    }
    const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
    const int e1 = ((o >> ta) & 1) * H, e2 = ((o >> tb) & 1) * H;
    const int an = lr ? 0 : N - 1;
    auto buf = [&](int c1, int c2) { return us[B::rot(d, an, c1 + e1, c2 + e2)]; };
    T uH;
    if (cubic && H + 2 >= 4) {
        int i1[4], i2[4];
        double a1[4], a2[4];
        cubic_stencil(t1, -1, H, i1, a1);
        cubic_stencil(t2, -1, H, i2, a2);
        T bt[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            bt[q] = comb4(a1, buf(i1[0], i2[q]), buf(i1[1], i2[q]), buf(i1[2], i2[q]), buf(i1[3], i2[q]));
        uH = comb4(a2, bt[0], bt[1], bt[2], bt[3]);
    } else {
        int c1 = ijkc(t1), c2 = ijkc(t2);
        int c1n = near_(t1, c1), c2n = near_(t2, c2);
        T S = (T)9.0 * buf(c1, c2) + buf(c1n, c2n);
        S = S + (T)3.0 * buf(c1, c2n);
        S = S + (T)3.0 * buf(c1n, c2);
        uH = S / (T)16.0;
    }
    const int inner = lr ? -1 : 1;
    T u0 = ud[B::rot(d, ag + inner, t1, t2)];
    T uN = ud[B::rot(d, ag + 2 * inner, t1, t2)];
    T q = (T)10.0 * u0 + (T)8.0 * uH;
    q = q - (T)3.0 * uN;
    ud[B::rot(d, ag, t1, t2)] = q / (T)15.0;
}

// ---- all three directions of fmg_ghostcell_fix for one AMR level in one launch ----
// The staged fill (x, then y, then z; parent-level then same-level faces in each) leaves in every
// ghost cell the value of the last pass that writes it: the highest direction among the cell's
// ghost coordinates whose face is not a physical boundary.  S<D>(b, cell) is the cell's value after
// the passes of directions <= D of this level; a pass-D value needs only S<D-1> values (a
// same-level copy reads the neighbour's layer, a parent-level interpolation reads the coarser
// level, complete, and the block's own cells at the two layers next to the face).  The recursion
// reads from memory only interior cells and ghost cells whose every ghost face is a physical
// boundary -- cells no pass writes, since physical faces of neighbouring blocks lie on the same
// domain face -- so each thread computes its cell's final value without reading a value another
// thread writes.  Same values as the three staged launches, bit for bit; on MI355X slower than
// them (the recursion's dependent loads and divergence cost more than the two launches it saves),
// so the solves keep the staged launches and this form is reached through the operator interface.
template <int N, typename T> struct GhostLevel {
    using B = Blk<N>;
    static constexpr int H = N / 2;
    T *A;
    const int *nbr, *nkind, *octant;
    int cubic;

    template <int D> __device__ T S(int b, int i, int j, int k) const
    {
        if constexpr (D < 0) {
            return A[(long)b * B::LEN + B::gi(i, j, k)];
        } else {
            const int c = D == 0 ? i : (D == 1 ? j : k);
            if (c == -1 || c == N) {
                const int lr = c == N ? 1 : 0;
                const int f = b * 6 + 2 * D + lr;
                const int kind = nkind[f];
                if (kind == NBR_SAME) {
                    const int m = nbr[f], layer = lr ? 0 : N - 1;
                    if (D == 0) return S<D - 1>(m, layer, j, k);
                    if (D == 1) return S<D - 1>(m, i, layer, k);
                    return S<D - 1>(m, i, j, layer);
                }
                if (kind == NBR_PARENT) return P<D>(b, lr, i, j, k);
            }
            return S<D - 1>(b, i, j, k);
        }
    }

    // fmg_ghfix_parentlev value of ghost cell (i, j, k) on face (D, lr) of block b (the formulas
    // of k_ghost's parent branch)
    template <int D> __device__ T P(int b, int lr, int i, int j, int k) const
    {
        constexpr int ta = D == 0 ? 1 : 0, tb = D == 2 ? 1 : 2;
        const int t1 = D == 0 ? j : i, t2 = D == 2 ? j : k;
        const int o = octant[b];
        const int e1 = ((o >> ta) & 1) * H, e2 = ((o >> tb) & 1) * H;
        const int an = lr ? 0 : N - 1;
        const T *us = A + (long)nbr[b * 6 + 2 * D + lr] * B::LEN;
        auto buf = [&](int c1, int c2) { return us[B::rot(D, an, c1 + e1, c2 + e2)]; };
        T uH;
        if (cubic && H + 2 >= 4) {
            int i1[4], i2[4];
            double a1[4], a2[4];
            cubic_stencil(t1, -1, H, i1, a1);
            cubic_stencil(t2, -1, H, i2, a2);
            T bt[4];
#pragma unroll
            for (int q = 0; q < 4; ++q)
                bt[q] = comb4(a1, buf(i1[0], i2[q]), buf(i1[1], i2[q]), buf(i1[2], i2[q]), buf(i1[3], i2[q]));
            uH = comb4(a2, bt[0], bt[1], bt[2], bt[3]);
        } else {
            int c1 = ijkc(t1), c2 = ijkc(t2);
            int c1n = near_(t1, c1), c2n = near_(t2, c2);
            T Sv = (T)9.0 * buf(c1, c2) + buf(c1n, c2n);
            Sv = Sv + (T)3.0 * buf(c1, c2n);
            Sv = Sv + (T)3.0 * buf(c1n, c2);
            uH = Sv / (T)16.0;
        }
        const int ag = lr ? N : -1, inner = lr ? -1 : 1;
        auto own = [&](int a) {
            if (D == 0) return S<D - 1>(b, a, t1, t2);
            if (D == 1) return S<D - 1>(b, t1, a, t2);
            return S<D - 1>(b, t1, t2, a);
        };
        const T u0 = own(ag + inner), uN = own(ag + 2 * inner);
        T q = (T)10.0 * u0 + (T)8.0 * uH;
        q = q - (T)3.0 * uN;
        return q / (T)15.0;
    }
};

// ghost cells of a block, each once: the two x slabs whole, the y slabs without x ghosts, the z
// slabs without x and y ghosts
template <int N> struct GhostCells {
    static constexpr int W = N + 2, NX = 2 * W * W, NY = 2 * N * W, NZ = 2 * N * N, COUNT = NX + NY + NZ;
    __device__ static void at(int g, int &i, int &j, int &k)
    {
        if (g < NX) {
            const int r = g % (W * W);
            i = g < W * W ? -1 : N;
            j = r % W - 1;
            k = r / W - 1;
        } else if (g < NX + NY) {
            const int h = g - NX, r = h % (N * W);
            j = h < N * W ? -1 : N;
            i = r % N;
            k = r / N - 1;
        } else {
            const int h = g - NX - NY, r = h % (N * N);
            k = h < N * N ? -1 : N;
            i = r % N;
            j = r / N;
        }
    }
};

// vmg_ghostcell (vmg.F90:728-742) of one AMR level in one launch: every cell it writes is either a
// copy of one stored cell (the same-level passes x, y, z only copy, so the value a ghost cell ends
// with is the cell the copies lead back to) or, on a physical face, zero.  pairs[t] = {destination,
// source or -1 for zero}, resolved by the host (amr.hip, same_fill_pairs), which also checks that no
// source is a destination.
template <typename T> __device__ __forceinline__ void fill_pair(T *A, const int2 *pairs, long t)
{
    const int2 p = pairs[t];
    A[p.x] = p.y >= 0 ? A[p.y] : (T)0;
}

template <typename T> __global__ void k_fill_pairs(T *A, const int2 *pairs, int n)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    fill_pair(A, pairs, t);
}

template <int N, typename T>
__global__ void k_ghost_level(T *A, int b0, int nb, int cubic, const int *nbr, const int *nkind, const int *octant)
{
    using B = Blk<N>;
    using GC = GhostCells<N>;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nb * GC::COUNT) return;
    const int b = b0 + (int)(t / GC::COUNT);
    int i, j, k;
    GC::at((int)(t % GC::COUNT), i, j, k);
    const int c = B::gi(i, j, k);
    const int gc[3] = {i, j, k};
    bool written = false;
#pragma unroll
    for (int d = 0; d < 3; ++d)
        if (gc[d] == -1 || gc[d] == N) written = written || nkind[b * 6 + 2 * d + (gc[d] == N ? 1 : 0)] != NBR_NONE;
    if (!written) return;
    const GhostLevel<N, T> G{A, nbr, nkind, octant, cubic};
    A[(long)b * B::LEN + c] = G.template S<2>(b, i, j, k);
}

// ---- fused refresh of the mixed solve (fmg_poisson_flux without its staged launches) ----
// The values a refresh (fmg_converge_c2p, then the staged fmg_ghostcell_fix) leaves in the cells
// that a colour pass or a residual reads, evaluated directly from the blocks' current interiors.
// Which stored cell each such value comes from depends on the hierarchy only, so the host resolves
// it once per context (amr.hip, build_prep_tables) into "sources": s >= 0 reads A[s] (a leaf's
// interior cell, or a ghost cell on a physical face, which no refresh pass writes); s < 0 is a
// covered block's interior cell, the linear restriction (restriction.F90:127-151) of its leaf
// child's 2x2x2 cells whose first cell sits at A[-s - 1].
template <int N, typename T> __device__ __forceinline__ T rstr_at(const T *A, long o)
{
    constexpr long W = N + 2, W2 = W * W;
    const T *u = A + o;
    T s = u[0] + u[1];
    s = s + u[W];
    s = s + u[W2];
    s = s + u[1 + W];
    s = s + u[1 + W2];
    s = s + u[W + W2];
    s = s + u[1 + W + W2];
    return s * (T)(1.0 / 8.0);
}

template <int N, typename T> __device__ __forceinline__ T src_val(const T *A, int s)
{
    return s >= 0 ? A[s] : rstr_at<N, T>(A, -(long)s - 1);
}


// One launch per refresh.  Items [0, ncov): the covered blocks' face layers that same-level leaf
// neighbours read (cov[t] = {destination, restriction origin}).  Items [ncov, + ncf H^2): one coarse
// face cell (c1, c2) of coarse-fine face e each: the four fine ghost cells behind it
// (fmg_ghfix_parentlev: bilinear (9,3,3,1)/16 tangentially, then (10 u0 + 8 uH - 3 uN)/15 normally,
// the operand order of k_ghost) and the reflux override of the coarse neighbour there (the 2x2
// average of the fine face fluxes, the order of k_reflux).  rec[e] = {fine block, face + 8 octant,
// override slot, AMR level}; src + e (H+2)^2 holds the sources of the coarse neighbour's face layer
// at tangential coarse positions -1..H (the fine block's quarter of it, ghosts included).
// Reads only leaf interiors and physical ghost cells, which no item writes.
// CC: the compact level-0 form of the inner solve (N = 8, T = float; A the compact interiors by
// active-list position, the tables of build_compact_tables): covered items write their value to the
// ghost store gs at cov[t].x, the fine ghosts go there too (aidx: the blocks' active positions).
template <int N, typename T, bool CC = false>
__global__ void k_prep(T *A, T *ov, const int2 *cov, int ncov, const int4 *rec, const int *src, int ncf,
                       LevelScalars hil, T *gs = nullptr, const int *aidx = nullptr)
{
    using B = Blk<N>;
    constexpr int H = N / 2, S = H + 2;
    static_assert(!CC || (N == 8 && std::is_same<T, float>::value), "the compact form is FP32 8^3");
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ncov) {
        const int2 p = cov[t];
        if constexpr (CC) {
            gs[p.x] = rstr_cc(A, p.y);
            return;
        }
        A[p.x] = rstr_at<N, T>(A, p.y);
        return;
    }
    t -= ncov;
    if (t >= (long)ncf * H * H) return;
    const int c1 = (int)(t % H), c2 = (int)((t / H) % H), e = (int)(t / (H * H));
    const int *sr = src + (long)e * S * S;
    // the 3x3 coarse neighbourhood the four fine ghosts interpolate from
    T w[3][3];
#pragma unroll
    for (int y = 0; y < 3; ++y)
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            const int so = sr[(c2 + y) * S + (c1 + x)];
            if constexpr (CC) {
                w[y][x] = so >= 0 ? A[so] : (so == -1 ? (T)0 : rstr_cc(A, -so - 2));
            } else {
                w[y][x] = src_val<N, T>(A, so);
            }
        }
    const int4 r = rec[e];
    const int b = r.x, f = r.y & 7, o = r.y >> 3, d = f >> 1, lr = f & 1;
    const int ag = lr ? N : -1, inner = lr ? -1 : 1;
    const int ab = CC ? aidx[b] : b;   // the block's position in A
    T *u = A + (long)ab * (CC ? 512 : B::LEN);
    const T hi = (T)hil.v[r.w];
    T ffv[2][2];
#pragma unroll
    for (int q2 = 0; q2 < 2; ++q2)
#pragma unroll
        for (int q1 = 0; q1 < 2; ++q1) {
            const int s1 = 2 * c1 + q1, s2 = 2 * c2 + q2;
            // buf(cc, cc) is w[1][1]; the near neighbour is one coarse cell down (q = 0) or up (q = 1)
            const int n1 = 2 * q1, n2 = 2 * q2;
            T Sv = (T)9.0 * w[1][1] + w[n2][n1];
            Sv = Sv + (T)3.0 * w[n2][1];
            Sv = Sv + (T)3.0 * w[1][n1];
            const T uH = Sv / (T)16.0;
            const T u0 = u[CC ? cc_rot(d, ag + inner, s1, s2) : B::rot(d, ag + inner, s1, s2)];
            const T uN = u[CC ? cc_rot(d, ag + 2 * inner, s1, s2) : B::rot(d, ag + 2 * inner, s1, s2)];
            T q = (T)10.0 * u0 + (T)8.0 * uH;
            q = q - (T)3.0 * uN;
            const T g = q / (T)15.0;
            if constexpr (CC) gs[((long)ab * 6 + f) * 64 + gs_cell(f, s1, s2)] = g;
            else u[B::rot(d, ag, s1, s2)] = g;
            // face flux (u(af + 1) - u(af)) hi with af = -1 (left face) or N - 1 (right face)
            ffv[q1][q2] = lr ? (g - u0) * hi : (u0 - g) * hi;
        }
    T v = ffv[0][0] + ffv[0][1];
    v = v + ffv[1][0];
    v = v + ffv[1][1];
    const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
    const int off1 = ((o >> ta) & 1) * H, off2 = ((o >> tb) & 1) * H;
    ov[(long)r.z * N * N + (off1 + c1) + N * (off2 + c2)] = v * (T)0.25;
}

// ---- fmg_boundary_fill0 on a list of physical faces (entries b*6+f) ----
template <int N, typename T> struct Fill0FacesF {   // item t of nf (N+2)^2
    T *A;
    const int *faces;
    __device__ void operator()(long t) const
    {
        using B = Blk<N>;
        constexpr int W = B::W;
        int t1 = (int)(t % W) - 1, t2 = (int)((t / W) % W) - 1;
        int e = faces[t / (W * W)];
        int b = e / 6, f = e % 6, d = f / 2;
        A[(long)b * B::LEN + B::rot(d, (f & 1) ? N : -1, t1, t2)] = (T)0;
    }
};

template <int N, typename T> __global__ void k_fill0_faces(T *A, const int *faces, int nf)
{
    constexpr int W = Blk<N>::W;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nf * W * W) return;
    Fill0FacesF<N, T>{A, faces}(t);
}

// ---- fmg_boundary_extrap, the faces of direction d only ----
template <int N, typename T> __global__ void k_extrap_faces(T *A, const int *faces, int nf, int d)
{
    using B = Blk<N>;
    constexpr int W = B::W;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nf * W * W) return;
    int t1 = (int)(t % W) - 1, t2 = (int)((t / W) % W) - 1;
    int e = faces[t / (W * W)];
    int b = e / 6, f = e % 6;
    if (f / 2 != d) return;
    T *u = A + (long)b * B::LEN;
    u[B::rot(d, (f & 1) ? N : -1, t1, t2)] = -u[B::rot(d, (f & 1) ? N - 1 : 0, t1, t2)];
}

// fmg_boundary_extrap's three directions in one launch, the values the passes x, y, z leave: a pass
// sets a physical-face ghost cell to minus the cell next to it inside, so the final value of a cell
// is plus or minus the cell reached by moving each of its ghost coordinates that lies on a physical
// face to the adjacent layer, last pass first; that cell is on no physical face, so no item writes
// it.  Items: the (N+2)^2 cells of each physical face (nkind names a block's other physical faces);
// a cell on two physical faces is written by both with the same value.
template <int N, typename T> __global__ void k_extrap_all(T *A, const int *faces, int nf, const int *nkind)
{
    using B = Blk<N>;
    constexpr int W = B::W;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nf * W * W) return;
    const int t1 = (int)(t % W) - 1, t2 = (int)((t / W) % W) - 1;
    const int e = faces[t / (W * W)];
    const int b = e / 6, f = e % 6, d = f / 2;
    int c[3];
    c[d] = (f & 1) ? N : -1;
    c[d == 0 ? 1 : 0] = t1;
    c[d == 2 ? 1 : 2] = t2;
    const int dst = B::gi(c[0], c[1], c[2]);
    bool neg = false;
#pragma unroll
    for (int D = 2; D >= 0; --D)
        if ((c[D] == -1 || c[D] == N) && nkind[b * 6 + 2 * D + (c[D] == N ? 1 : 0)] == NBR_NONE) {
            c[D] = c[D] == N ? N - 1 : 0;
            neg = !neg;
        }
    T *u = A + (long)b * B::LEN;
    const T v = u[B::gi(c[0], c[1], c[2])];
    u[dst] = neg ? -v : v;
}

// ---- fmg_reflux_do: fine face fluxes averaged 2x2 into the coarse neighbour's override ----
template <int N, typename T>
__global__ void k_reflux(const T *A, T *ov, const int *cf, const int *cf_slot, int ncf,
                         const double *hi_lev, const int *level, const int *octant)
{
    using B = Blk<N>;
    constexpr int H = N / 2;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)ncf * H * H) return;
    int c1 = (int)(t % H), c2 = (int)((t / H) % H), e = (int)(t / (H * H));
    int b = cf[e] / 6, f = cf[e] % 6, d = f / 2, lr = f & 1;
    const T hi = (T)hi_lev[level[b]];
    const T *u = A + (long)b * B::LEN;
    const int af = lr ? N - 1 : -1;
    auto ff = [&](int a1, int a2) { return (u[B::rot(d, af + 1, a1, a2)] - u[B::rot(d, af, a1, a2)]) * hi; };
    const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
    const int o = octant[b];
    const int off1 = ((o >> ta) & 1) * H, off2 = ((o >> tb) & 1) * H;
    int f1 = 2 * c1, f2 = 2 * c2;
    T v = ff(f1, f2) + ff(f1, f2 + 1);
    v = v + ff(f1 + 1, f2);
    v = v + ff(f1 + 1, f2 + 1);
    ov[(long)cf_slot[e] * N * N + (off1 + c1) + N * (off2 + c2)] = v * (T)0.25;
}

// Face-flux view of one block: fluxes from the block's values, or the override on reflux faces.
// With nb (the neighbour table of a block) set, same-level faces read the neighbour block's
// boundary cell directly instead of the block's own ghost cell ("halo-free" form): the value a
// same-level ghost copy (fmg_ghfix_samelev) would have placed there.
template <int N, typename T> struct Flux {
    const T *u;
    const int *slot;     // ovslot + 6 b, or nullptr
    const T *ov;
    T hi;
    const T *A = nullptr;       // halo-free form: all blocks of the level
    const int *nb = nullptr;    //                 nbr + 6 b
    const int *kind = nullptr;  //                 nkind + 6 b
    __device__ T ov_or(int f, int t1, int t2, T computed) const
    {
        if (slot) {
            int s = slot[f];
            if (s >= 0) return ov[(long)s * N * N + t1 + N * t2];
        }
        return computed;
    }
    // value across face f (d = f / 2) at tangential (t1, t2): own ghost, or the neighbour's cell
    __device__ T across(int f, int t1, int t2) const
    {
        using B = Blk<N>;
        const int d = f >> 1, lr = f & 1;
        if (nb && kind[f] == NBR_SAME) return A[(long)nb[f] * B::LEN + B::rot(d, lr ? 0 : N - 1, t1, t2)];
        return u[B::rot(d, lr ? N : -1, t1, t2)];
    }
    // sum of flux differences at (i, j, k) in the reference's operand order
    __device__ T div(int i, int j, int k) const
    {
        using B = Blk<N>;
        const T uc = u[B::gi(i, j, k)];
        const T xp = (i == N - 1) ? across(1, j, k) : u[B::gi(i + 1, j, k)];
        const T xm = (i == 0) ? across(0, j, k) : u[B::gi(i - 1, j, k)];
        const T yp = (j == N - 1) ? across(3, i, k) : u[B::gi(i, j + 1, k)];
        const T ym = (j == 0) ? across(2, i, k) : u[B::gi(i, j - 1, k)];
        const T zp = (k == N - 1) ? across(5, i, j) : u[B::gi(i, j, k + 1)];
        const T zm = (k == 0) ? across(4, i, j) : u[B::gi(i, j, k - 1)];
        T fxp = (i == N - 1) ? ov_or(1, j, k, (xp - uc) * hi) : (xp - uc) * hi;
        T fxm = (i == 0) ? ov_or(0, j, k, (uc - xm) * hi) : (uc - xm) * hi;
        T fyp = (j == N - 1) ? ov_or(3, i, k, (yp - uc) * hi) : (yp - uc) * hi;
        T fym = (j == 0) ? ov_or(2, i, k, (uc - ym) * hi) : (uc - ym) * hi;
        T fzp = (k == N - 1) ? ov_or(5, i, j, (zp - uc) * hi) : (zp - uc) * hi;
        T fzm = (k == 0) ? ov_or(4, i, j, (uc - zm) * hi) : (uc - zm) * hi;
        T v = fxp - fxm;
        v = v + fyp;
        v = v - fym;
        v = v + fzp;
        v = v - fzm;
        return v;
    }
};

__device__ __forceinline__ int pick_block(const int *list, int b0, int i) { return list ? list[i] : b0 + i; }

// Max reductions of the smoother go to kMaxSlots addresses (workgroup index modulo kMaxSlots), so
// that thousands of workgroups do not serialise on one atomic; k_max_slots folds them.
constexpr int kMaxSlots = 64;

__global__ void k_zero_slots(double *slots)
{
    if (threadIdx.x < kMaxSlots) slots[threadIdx.x] = 0.0;
}

__global__ void k_max_slots(const double *slots, double *out)
{
    double v = threadIdx.x < kMaxSlots ? slots[threadIdx.x] : 0.0;
    v = wave_max(v);
    if (threadIdx.x == 0) *out = v;
}

// ---- per-block records of the relaxation and residual kernels ----
// 16 ints per block: block number, AMR level, the same-level neighbour across each face (-1 where
// the face meets a coarser level or the domain boundary), the reflux override slot of each face
// (-1 where there is none), two unused.  One 64-byte load gives a workgroup or thread everything
// it needs before its data loads; the cell widths come in by value (LevelScalars).
constexpr int kRec = 16;
struct BlockRec {
    int b, l, nb[6], slot[6];
    __device__ void load(const int *rec, long a)
    {
        const int4 *p = (const int4 *)(rec + kRec * a);
        const int4 r0 = p[0], r1 = p[1], r2 = p[2], r3 = p[3];
        b = r0.x; l = r0.y;
        nb[0] = r0.z; nb[1] = r0.w; nb[2] = r1.x; nb[3] = r1.y; nb[4] = r1.z; nb[5] = r1.w;
        slot[0] = r2.x; slot[1] = r2.y; slot[2] = r2.z; slot[3] = r2.w; slot[4] = r3.x; slot[5] = r3.y;
    }
};

// Face-flux view of one block from its record: fluxes from the block's values, or the reflux
// override on faces that have one (when ov is set).  With A set ("halo-free" form), same-level
// faces read the neighbour block's boundary cell directly instead of the block's own ghost cell:
// the value a same-level ghost copy (fmg_ghfix_samelev) would place there.
template <int N, typename T> struct FluxR {
    const T *u;
    const T *A;
    const T *ov;
    T hi;
    const BlockRec &r;
    __device__ T ov_or(int f, int t1, int t2, T computed) const
    {
        if (ov && r.slot[f] >= 0) return ov[(long)r.slot[f] * N * N + t1 + N * t2];
        return computed;
    }
    __device__ T across(int f, int t1, int t2) const
    {
        using B = Blk<N>;
        const int d = f >> 1, lr = f & 1;
        if (A && r.nb[f] >= 0) return A[(long)r.nb[f] * B::LEN + B::rot(d, lr ? 0 : N - 1, t1, t2)];
        return u[B::rot(d, lr ? N : -1, t1, t2)];
    }
    // sum of flux differences at (i, j, k) in the reference's operand order
    __device__ T div(int i, int j, int k) const
    {
        using B = Blk<N>;
        const T uc = u[B::gi(i, j, k)];
        const T xp = (i == N - 1) ? across(1, j, k) : u[B::gi(i + 1, j, k)];
        const T xm = (i == 0) ? across(0, j, k) : u[B::gi(i - 1, j, k)];
        const T yp = (j == N - 1) ? across(3, i, k) : u[B::gi(i, j + 1, k)];
        const T ym = (j == 0) ? across(2, i, k) : u[B::gi(i, j - 1, k)];
        const T zp = (k == N - 1) ? across(5, i, j) : u[B::gi(i, j, k + 1)];
        const T zm = (k == 0) ? across(4, i, j) : u[B::gi(i, j, k - 1)];
        T fxp = (i == N - 1) ? ov_or(1, j, k, (xp - uc) * hi) : (xp - uc) * hi;
        T fxm = (i == 0) ? ov_or(0, j, k, (uc - xm) * hi) : (uc - xm) * hi;
        T fyp = (j == N - 1) ? ov_or(3, i, k, (yp - uc) * hi) : (yp - uc) * hi;
        T fym = (j == 0) ? ov_or(2, i, k, (uc - ym) * hi) : (uc - ym) * hi;
        T fzp = (k == N - 1) ? ov_or(5, i, j, (zp - uc) * hi) : (zp - uc) * hi;
        T fzm = (k == 0) ? ov_or(4, i, j, (uc - zm) * hi) : (uc - zm) * hi;
        T v = fxp - fxm;
        v = v + fyp;
        v = v - fym;
        v = v + fzp;
        v = v - fzm;
        return v;
    }
};

// ---- fmg_poisson_relax / vmg_poisson_relax: one colour pass over the blocks of nrec records ----
// (ov: the reflux overrides, or nullptr in the VMG, which has none; halo: read same-level
// neighbours directly)
template <int N, typename T>
__global__ void k_relax(T *Au, const T *Arhs, const int *rec, int nrec, int ipass, LevelScalars h_lev,
                        LevelScalars hi_lev, const T *ov, double *resmax, int halo)
{
    using B = Blk<N>;
    constexpr int HN = (N + 1) / 2, PER = N * N * HN;
    // halo-free 4^3 passes: blocks in contiguous ranges per XCD (xcd_block), the neighbour reads
    // then mostly hit the reading XCD's L2 (the few workgroups of a 2^3 pass gain nothing)
    const long w = (halo && N >= 4) ? xcd_block((int)blockIdx.x, (int)gridDim.x) : (long)blockIdx.x;
    long t = w * blockDim.x + threadIdx.x;
    double a = 0.0;
    if (t < (long)nrec * PER) {
        int ih = (int)(t % HN), j = (int)((t / HN) % N), k = (int)((t / (HN * N)) % N);
        BlockRec R;
        R.load(rec, t / PER);
        int i = 2 * ih + ((j + k + ipass) & 1);
        if (i < N) {
            const T h = (T)h_lev.v[R.l];
            T *u = Au + (long)R.b * B::LEN;
            FluxR<N, T> F{u, halo ? Au : nullptr, ov, (T)hi_lev.v[R.l], R};
            T r = F.div(i, j, k) - Arhs[(long)R.b * B::LEN + B::gi(i, j, k)] * h;
            T resh2 = r * h;
            u[B::gi(i, j, k)] = u[B::gi(i, j, k)] + (T)(1.0 / 6.0) * resh2;
            a = (double)fabs(resh2);
        }
    }
    if (resmax) block_atomic_max(resmax + (blockIdx.x & (kMaxSlots - 1)), a);
}

// ---- fmg_poisson_resid: res = rhs - div F / h on the blocks of nrec records ----
template <int N, typename T>
__global__ void k_resid(T *Ares, const T *Au, const T *Arhs, const int *rec, int nrec, LevelScalars hi_lev,
                        const T *ov, int halo)
{
    using B = Blk<N>;
    constexpr int PER = N * N * N;
    // halo-free 4^3 passes: blocks in contiguous ranges per XCD, as in k_relax
    const long w = (halo && N >= 4) ? xcd_block((int)blockIdx.x, (int)gridDim.x) : (long)blockIdx.x;
    long t = w * blockDim.x + threadIdx.x;
    if (t >= (long)nrec * PER) return;
    int i = (int)(t % N), j = (int)((t / N) % N), k = (int)((t / (N * N)) % N);
    BlockRec R;
    R.load(rec, t / PER);
    const T hi = (T)hi_lev.v[R.l];
    FluxR<N, T> F{Au + (long)R.b * B::LEN, halo ? Au : nullptr, ov, hi, R};
    const long q = (long)R.b * B::LEN + B::gi(i, j, k);
    Ares[q] = Arhs[q] - F.div(i, j, k) * hi;
}

// ---- fmg_rstrct (linear): coarse level with NC cells from the fine level with 2 NC ----
template <int NC, typename T> struct RstrctF {   // item t of nb NC^3
    T *Ac;
    const T *Af;
    const int *list;
    __device__ void operator()(long t) const
    {
        constexpr int PER = NC * NC * NC;
        int i = (int)(t % NC), j = (int)((t / NC) % NC), k = (int)((t / (NC * NC)) % NC);
        int b = list[t / PER];
        Ac[(long)b * Blk<NC>::LEN + Blk<NC>::gi(i, j, k)] =
            rstr_lin<2 * NC, T>(Af + (long)b * Blk<2 * NC>::LEN, 2 * i, 2 * j, 2 * k);
    }
};

template <int NC, typename T> __global__ void k_rstrct(T *Ac, const T *Af, const int *list, int nb)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nb * NC * NC * NC) return;
    RstrctF<NC, T>{Ac, Af, list}(t);
}

// fmg_rstrct from FMG level 0 to 1 with the level-0 source in compact form (src + a 512 for list
// entry a)
struct RstrctCCF {   // item t of nb 64
    float *Ac;
    const float *src;
    const int *list;
    __device__ void operator()(long t) const
    {
        const int i = (int)(t & 3), j = (int)((t >> 2) & 3), k = (int)((t >> 4) & 3), a = (int)(t >> 6);
        Ac[(long)list[a] * Blk<4>::LEN + Blk<4>::gi(i, j, k)] = rstr_lin_cc(src + (long)a * 512, 2 * i, 2 * j, 2 * k);
    }
};

__global__ void k_rstrct_cc(float *Ac, const float *src, const int *list, int nb)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nb * 64) return;
    RstrctCCF{Ac, src, list}(t);
}

// ---- fmg_interp_cubic / fmg_interp_linear: fine interior of each listed block from its own
//      coarser FMG level (ghosts included) ----
template <int NF, bool CUBIC, typename T> __global__ void k_interp(T *Af, const T *Ac, const int *list, int nb)
{
    constexpr int NC = NF / 2, PER = NF * NF * NF;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nb * PER) return;
    int i = (int)(t % NF), j = (int)((t / NF) % NF), k = (int)((t / (NF * NF)) % NF);
    int b = list[t / PER];
    const T *uc = Ac + (long)b * Blk<NC>::LEN;
    auto cget = [&](int a, int bb, int c) -> T { return uc[Blk<NC>::gi(a, bb, c)]; };
    T v = (CUBIC && NC + 2 >= 4) ? tricubic_at(i, j, k, NC, cget) : trilinear_at(i, j, k, cget);
    Af[(long)b * Blk<NF>::LEN + Blk<NF>::gi(i, j, k)] = v;
}

// The tricubic case of k_interp with one 256-thread workgroup per listed block: the coarse block
// with its ghosts is staged in LDS and the three separable passes of interpolation.F90:40-218
// (bufx, bufxy, then z) run over it, so each intermediate is formed once instead of once per fine
// cell that uses it.  Every intermediate is the value tricubic_at forms, so the results are the
// same bit for bit.
// (ucc: the fine blocks' interiors go there in compact form instead, list entry a at 512 a, and
// their boundary cells into the same-level neighbours' ghost stores gs through ptab)
template <int NF, typename T> __global__ void __launch_bounds__(256) k_interp_cubic_blk(T *Af, const T *Ac, const int *list,
                                                                                     int nb, T *ucc = nullptr,
                                                                                     T *gs = nullptr,
                                                                                     const int *ptab = nullptr)
{
    constexpr int NC = NF / 2, W = NC + 2;   // coarse cells -1 .. NC per axis
    static_assert(NC + 2 >= 4, "the tricubic stencil needs four coarse cells");
    __shared__ T sc[W * W * W];
    __shared__ T bx[W * W * NF];    // [kc][jc][i]
    __shared__ T bxy[W * NF * NF];  // [kc][j][i]
    __shared__ int sidx[NF][4];     // stencil index + 1 (ghosted offset)
    __shared__ double swt[NF][4];
    if ((int)blockIdx.x >= nb) return;
    const int b = list[blockIdx.x];
    const T *uc = Ac + (long)b * Blk<NC>::LEN;
    for (int q = (int)threadIdx.x; q < W * W * W; q += (int)blockDim.x) sc[q] = uc[q];
    if ((int)threadIdx.x < NF) {
        int idx[4];
        double a[4];
        cubic_stencil((int)threadIdx.x, -1, NC, idx, a);
        for (int e = 0; e < 4; ++e) {
            sidx[threadIdx.x][e] = idx[e] + 1;
            swt[threadIdx.x][e] = a[e];
        }
    }
    __syncthreads();
    for (int q = (int)threadIdx.x; q < W * W * NF; q += (int)blockDim.x) {
        const int i = q % NF, r = q / NF;  // row r = kc * W + jc (ghosted)
        const T *row = sc + r * W;
        bx[q] = comb4<T>(swt[i], row[sidx[i][0]], row[sidx[i][1]], row[sidx[i][2]], row[sidx[i][3]]);
    }
    __syncthreads();
    for (int q = (int)threadIdx.x; q < W * NF * NF; q += (int)blockDim.x) {
        const int i = q % NF, j = (q / NF) % NF, kc = q / (NF * NF);
        const T *col = bx + kc * W * NF + i;
        bxy[q] = comb4<T>(swt[j], col[sidx[j][0] * NF], col[sidx[j][1] * NF], col[sidx[j][2] * NF], col[sidx[j][3] * NF]);
    }
    __syncthreads();
    T *uf = Af + (long)b * Blk<NF>::LEN;
    for (int q = (int)threadIdx.x; q < NF * NF * NF; q += (int)blockDim.x) {
        const int i = q % NF, j = (q / NF) % NF, k = q / (NF * NF);
        const T *p = bxy + j * NF + i;
        const T v = comb4<T>(swt[k], p[sidx[k][0] * NF * NF], p[sidx[k][1] * NF * NF], p[sidx[k][2] * NF * NF],
                             p[sidx[k][3] * NF * NF]);
        if (NF == 8 && ucc) {
            ucc[(long)blockIdx.x * 512 + cc_index(i, j, k)] = v;
            push_faces(gs, ptab + blockIdx.x * 6, i, j, k, v);
        } else {
            uf[Blk<NF>::gi(i, j, k)] = v;
        }
    }
}

// ---- whole-array block operations (ghosts included) ----
template <typename T> struct Fill0BlocksF {   // item t of nb len
    T *A;
    const int *list;
    int b0, len;
    __device__ void operator()(long t) const { A[(long)pick_block(list, b0, (int)(t / len)) * len + t % len] = (T)0; }
};

template <typename T> struct CopyBlocksF {   // item t of nb len
    T *O;
    const T *I;
    const int *list;
    int b0, len;
    __device__ void operator()(long t) const
    {
        long q = (long)pick_block(list, b0, (int)(t / len)) * len + t % len;
        O[q] = I[q];
    }
};

// two independent steps in one launch: items [0, n1) of f1, then [n1, n1 + n2) of f2 (neither may
// read what the other writes)
template <class F1, class F2> __global__ void k_pair(F1 f1, long n1, F2 f2, long n2)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n1) f1(t);
    else if (t < n1 + n2) f2(t - n1);
}

template <typename T> __global__ void k_blocks_fill0(T *A, const int *list, int b0, int nb, int len)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nb * len) return;
    Fill0BlocksF<T>{A, list, b0, len}(t);
}

template <typename T> __global__ void k_blocks_copy(T *O, const T *I, const int *list, int b0, int nb, int len)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nb * len) return;
    CopyBlocksF<T>{O, I, list, b0, len}(t);
}

template <typename T> __global__ void k_blocks_add(T *A, const T *D, const int *list, int b0, int nb, int len)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nb * len) return;
    long q = (long)pick_block(list, b0, (int)(t / len)) * len + t % len;
    A[q] = A[q] + D[q];
}

template <typename T> __global__ void k_blocks_sub(T *O, const T *A, const T *C, int b0, int nb, int len)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nb * len) return;
    long q = (long)(b0 + t / len) * len + t % len;
    O[q] = A[q] - C[q];
}

// ---- vmg_poisson_tau, fine part: tau = 0, then tau = -(div F)/h inside, all blocks of a level ----
template <int N, typename T> __global__ void k_tau_fine(T *Atau, const T *Au, int b0, int nb, double hi_d)
{
    using B = Blk<N>;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nb * B::LEN) return;
    int b = b0 + (int)(t / B::LEN), q = (int)(t % B::LEN);
    int i = q % B::W - 1, j = (q / B::W) % B::W - 1, k = q / (B::W * B::W) - 1;
    const T hi = (T)hi_d;
    T v = 0;
    if (i >= 0 && i < N && j >= 0 && j < N && k >= 0 && k < N) {
        Flux<N, T> F{Au + (long)b * B::LEN, nullptr, nullptr, hi};
        v = -(F.div(i, j, k) * hi);
    }
    Atau[(long)b * B::LEN + q] = v;
}

// ---- vmg_poisson_tau, coarse part: tau += (div F)/h on the listed (covered) blocks ----
template <int N, typename T> __global__ void k_tau_add(T *Atau, const T *Au, const int *list, int nb, double hi_d)
{
    using B = Blk<N>;
    constexpr int PER = N * N * N;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nb * PER) return;
    int i = (int)(t % N), j = (int)((t / N) % N), k = (int)((t / (N * N)) % N);
    int b = list[t / PER];
    const T hi = (T)hi_d;
    Flux<N, T> F{Au + (long)b * B::LEN, nullptr, nullptr, hi};
    long q = (long)b * B::LEN + B::gi(i, j, k);
    Atau[q] = Atau[q] + F.div(i, j, k) * hi;
}

// ---- vmg_tau_crop: a child face meeting the parent level zeroes the parent's boundary plane ----
template <int N, typename T> __global__ void k_tau_crop(T *Atau, const int *cf, int ncf, const int *parent)
{
    using B = Blk<N>;
    constexpr int W = B::W;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)ncf * W * W) return;
    int t1 = (int)(t % W) - 1, t2 = (int)((t / W) % W) - 1, e = cf[t / (W * W)];
    int b = e / 6, f = e % 6;
    Atau[(long)parent[b] * B::LEN + B::rot(f / 2, (f & 1) ? N - 1 : 0, t1, t2)] = (T)0;
}

// ---- merged steps of vmg_fas at FMG level 2 (2^3 cells per block) for the mixed solve's fused
//      mode: each does, in one launch, what consecutive launches of the staged sequence do, with
//      the same operations in the same order on every value ----
namespace vmgm {
using B2 = Blk<2>;
__device__ __forceinline__ void cell_of(int c, int &i, int &j, int &k)
{
    i = c % B2::W - 1;
    j = (c / B2::W) % B2::W - 1;
    k = c / (B2::W * B2::W) - 1;
}
__device__ __forceinline__ bool interior(int i, int j, int k) { return i >= 0 && i < 2 && j >= 0 && j < 2 && k >= 0 && k < 2; }
} // namespace vmgm

// vmg_gridboundary (vmg.F90:708-724) of the blocks [b0, b0+nb) in one launch: the parent-level
// faces' passes x, y, z (linear, the formulas of k_ghost's parent branch) and then, with fill0, the
// zero on physical faces.  A parent-level pass reads only the coarser level and the block's own
// cells, never a same-level neighbour, so each block runs its passes on its own: 32 threads per
// block (the two face slabs of the pass direction), a workgroup barrier between passes.
template <typename T>
__device__ __forceinline__ void parent_fill_block(T *A, int b0, int nb, int fill0, const int *nbr, const int *nkind,
                                                  const int *octant)
{
    using B = Blk<2>;
    constexpr int N = 2, W = B::W, H = 1;
    const int bl = (int)threadIdx.x >> 5, idx = (int)threadIdx.x & 31;
    const int b = b0 + (int)blockIdx.x * 8 + bl;
    const bool mine = b < b0 + nb;
    const int lr = idx >> 4, t1 = (idx & 3) - 1, t2 = ((idx >> 2) & 3) - 1;
    const int bb = mine ? b : b0;
    T *ud = A + (long)bb * B::LEN;
    // the block's tables for its three passes, loaded together ahead of them
    int kd[3], src[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        kd[d] = nkind[bb * 6 + 2 * d + lr];
        src[d] = nbr[bb * 6 + 2 * d + lr];
    }
    const int o = octant[bb];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (mine && kd[d] == NBR_PARENT) {
            const T *us = A + (long)src[d] * B::LEN;
            const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
            const int e1 = ((o >> ta) & 1) * H, e2 = ((o >> tb) & 1) * H;
            const int an = lr ? 0 : N - 1;
            auto buf = [&](int c1, int c2) { return us[B::rot(d, an, c1 + e1, c2 + e2)]; };
            const int c1 = ijkc(t1), c2 = ijkc(t2);
            const int c1n = near_(t1, c1), c2n = near_(t2, c2);
            T S = (T)9.0 * buf(c1, c2) + buf(c1n, c2n);
            S = S + (T)3.0 * buf(c1, c2n);
            S = S + (T)3.0 * buf(c1n, c2);
            const T uH = S / (T)16.0;
            const int ag = lr ? N : -1, inner = lr ? -1 : 1;
            const T u0 = ud[B::rot(d, ag + inner, t1, t2)];
            const T uN = ud[B::rot(d, ag + 2 * inner, t1, t2)];
            T q = (T)10.0 * u0 + (T)8.0 * uH;
            q = q - (T)3.0 * uN;
            ud[B::rot(d, ag, t1, t2)] = q / (T)15.0;
        }
        __syncthreads();
    }
    if (!fill0 || !mine) return;
    for (int g = idx; g < 6 * W * W; g += 32) {
        const int f = g / (W * W), r = g % (W * W);
        if (nkind[b * 6 + f] == NBR_NONE) ud[B::rot(f >> 1, (f & 1) ? N : -1, r % W - 1, r / W - 1)] = (T)0;
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_parent_fill(T *A, int b0, int nb, int fill0, const int *nbr, const int *nkind,
                                                     const int *octant)
{
    parent_fill_block(A, b0, nb, fill0, nbr, nkind, octant);
}

// fmg_converge_c2p_lev of two arrays at once (children [b0, b0+nb) into their parents)
template <typename T> __global__ void k_c2p_two(T *A1, T *A2, int b0, int nb, const int *parent, const int *octant)
{
    using B = Blk<2>;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nb) return;
    const int b = b0 + (int)t, o = octant[b];
    const long dst = (long)parent[b] * B::LEN + B::gi(o & 1, (o >> 1) & 1, (o >> 2) & 1);
    A1[dst] = rstr_lin<2, T>(A1 + (long)b * B::LEN, 0, 0, 0);
    A2[dst] = rstr_lin<2, T>(A2 + (long)b * B::LEN, 0, 0, 0);
}

// vmg_rstrct(v-1, U, cubic) + vmg_copy(v-1, IRUF <- U) + vmg_rstrct(v-1, RHS) over the blocks
// [b0, b0+nb) of level v-1: a covered block's interior cell gets its child's cubic restriction of
// U (in U and in PSI, the IRUF slot) and linear restriction of RHS; every other cell copies U to PSI.
template <typename T>
__global__ void k_vmg_rstr(T *U, T *PSI, T *RHS, int b0, int nb, const int *child, const int *covered, Cubic64 st)
{
    using B = Blk<2>;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nb * B::LEN) return;
    const int b = b0 + (int)(t / B::LEN), c = (int)(t % B::LEN);
    int i, j, k;
    vmgm::cell_of(c, i, j, k);
    const long q = (long)b * B::LEN + c;
    const bool in = vmgm::interior(i, j, k);
    const int cov = covered[b], ch = in ? child[8 * b + i + 2 * j + 4 * k] : -1;   // loaded together
    if (in && cov) {
        const long cb = (long)ch * B::LEN;
        const T v = rstr_cub<2, T>(U + cb, 0, 0, 0, st);
        U[q] = v;
        PSI[q] = v;
        RHS[q] = rstr_lin<2, T>(RHS + cb, 0, 0, 0);
        return;
    }
    PSI[q] = U[q];
}

// vmg_poisson_tau's fine part for level v (tau = 0, then -(div F)/h inside, into RES of the blocks
// [b0, b0+nb)) + fmg_converge_c2p_lev of RES into the parents.  The thread of a block's first
// interior cell also forms the restriction, from the eight values of its block (recomputed in the
// same way, so equal to the stored ones).
template <typename T> __device__ __forceinline__ void tau_fine_c2p(T *RES, const T *U, int b0, double hi_d, const int *parent,
                                                                  const int *octant, long t)
{
    using B = Blk<2>;
    const int b = b0 + (int)(t / B::LEN), c = (int)(t % B::LEN);
    int i, j, k;
    vmgm::cell_of(c, i, j, k);
    const T hi = (T)hi_d;
    const long base = (long)b * B::LEN;
    if (!vmgm::interior(i, j, k)) {
        RES[base + c] = (T)0;
        return;
    }
    Flux<2, T> F{U + base, nullptr, nullptr, hi};
    RES[base + c] = -(F.div(i, j, k) * hi);
    if (c != B::gi(0, 0, 0)) return;
    T f[2][2][2];
#pragma unroll
    for (int kk = 0; kk < 2; ++kk)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int ii = 0; ii < 2; ++ii) f[kk][jj][ii] = -(F.div(ii, jj, kk) * hi);
    T sum = f[0][0][0] + f[0][0][1];
    sum = sum + f[0][1][0];
    sum = sum + f[1][0][0];
    sum = sum + f[0][1][1];
    sum = sum + f[1][0][1];
    sum = sum + f[1][1][0];
    sum = sum + f[1][1][1];
    const int o = octant[b];
    RES[(long)parent[b] * B::LEN + B::gi(o & 1, (o >> 1) & 1, (o >> 2) & 1)] = sum * (T)(1.0 / 8.0);
}

template <typename T> __global__ void k_tau_fine_c2p(T *RES, const T *U, int b0, int nb, double hi_d, const int *parent,
                                                    const int *octant)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nb * Blk<2>::LEN) return;
    tau_fine_c2p(RES, U, b0, hi_d, parent, octant, t);
}

// k_tau_fine_c2p on AMR level l (items [0, nb 64)) and vmg_ghostcell gathers (k_fill_pairs, items
// [nb 64, + np)) in one launch: of the coarser level, whose U the first does not read, and of level
// l itself where the gather leaves every cell the first reads (interiors, face ghosts) as it is.
template <typename T> __global__ void k_tau_fine_c2p_fill(T *RES, T *U, int b0, int nb, double hi_d, const int *parent,
                                                         const int *octant, const int2 *pairs, int np)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nt = (long)nb * Blk<2>::LEN;
    if (t < nt) tau_fine_c2p(RES, (const T *)U, b0, hi_d, parent, octant, t);
    else if (t < nt + np) fill_pair(U, pairs, t - nt);
}

// vmg_poisson_tau's coarse part (tau += (div F)/h inside) + vmg_tau_crop (the planes of the
// parent next to a child face that meets the parent's level set to zero; crop[e] is that mask of
// the 64 cells of covered block list[e]) + vmg_add_onov (RHS += tau) on the covered blocks of one level.
template <typename T> __device__ __forceinline__ void tau_finish_cell(T *RES, T *RHS, const T *U, const int *list,
                                                                     double hi_d, const unsigned long long *crop, long t)
{
    using B = Blk<2>;
    const int e = (int)(t / B::LEN), c = (int)(t % B::LEN), b = list[e];
    int i, j, k;
    vmgm::cell_of(c, i, j, k);
    const long q = (long)b * B::LEN + c;
    T tau = RES[q];
    if (vmgm::interior(i, j, k)) {
        Flux<2, T> F{U + (long)b * B::LEN, nullptr, nullptr, (T)hi_d};
        tau = tau + F.div(i, j, k) * (T)hi_d;
    }
    if ((crop[e] >> c) & 1ULL) tau = (T)0;
    RES[q] = tau;
    RHS[q] = RHS[q] + tau;
}

template <typename T> __global__ void k_tau_finish(T *RES, T *RHS, const T *U, const int *list, int nc, double hi_d,
                                                  const unsigned long long *crop)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nc * Blk<2>::LEN) return;
    tau_finish_cell(RES, RHS, U, list, hi_d, crop, t);
}

// k_tau_finish on the covered blocks of AMR level l (workgroups from (nb + 7) / 8 on) and the next
// VMG step's vmg_gridboundary of level l (k_parent_fill, the first (nb + 7) / 8 workgroups) in one
// launch.  A covered block never meets a coarser level, so the parent-level passes write no cell of
// a covered block but its physical-face ghosts, which hold zero already (the vmg_ghostcell before
// the finish zeroed them) and get zero again; the finish writes only RES and RHS.
template <typename T>
__global__ void __launch_bounds__(256) k_tau_finish_pfill(T *RES, T *RHS, T *U, const int *list, int nc, double hi_d,
                                                          const unsigned long long *crop, int b0, int nb, int fill0,
                                                          const int *nbr, const int *nkind, const int *octant)
{
    const int g1 = (nb + 7) / 8;
    if ((int)blockIdx.x < g1) {   // uniform per workgroup: the parent fill's barriers see whole workgroups
        parent_fill_block(U, b0, nb, fill0, nbr, nkind, octant);
        return;
    }
    const long t = (long)(blockIdx.x - g1) * blockDim.x + threadIdx.x;
    if (t < (long)nc * Blk<2>::LEN) tau_finish_cell(RES, RHS, (const T *)U, list, hi_d, crop, t);
}

// vmg_sub (RES = U - PSI over whole arrays of one level) + vmg_ghostcell of RES there: map[t] is -2
// where the fill leaves the cell, -1 where it zeroes it (physical face), else the cell it copies.
template <typename T> __global__ void k_sub_fill(T *RES, const T *U, const T *PSI, long base, int n, const int *map)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int m = map[t];
    const long q = base + t;
    RES[q] = m == -2 ? U[q] - PSI[q] : (m == -1 ? (T)0 : U[m] - PSI[m]);
}

// vmg_interp_p2c_lev (trilinear, child interiors from the parent's octant, its ghosts included) +
// vmg_add (U += RES over whole arrays) for the blocks [b0, b0+nb)
template <typename T> __device__ __forceinline__ void p2c_add_cell(T *RES, T *U, int b, int c, const int *parent,
                                                                  const int *octant)
{
    using B = Blk<2>;
    int i, j, k;
    vmgm::cell_of(c, i, j, k);
    const long q = (long)b * B::LEN + c;
    if (vmgm::interior(i, j, k)) {
        const T *uc = RES + (long)parent[b] * B::LEN;
        const int o = octant[b];
        const int oi = o & 1, oj = (o >> 1) & 1, ok = (o >> 2) & 1;
        auto cget = [&](int a, int bb, int cc) -> T { return uc[B::gi(oi + a, oj + bb, ok + cc)]; };
        const T r = trilinear_at(i, j, k, cget);
        RES[q] = r;
        U[q] = U[q] + r;
        return;
    }
    U[q] = U[q] + RES[q];
}

template <typename T> __global__ void k_p2c_add(T *RES, T *U, int b0, int nb, const int *parent, const int *octant)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nb * Blk<2>::LEN) return;
    p2c_add_cell(RES, U, b0 + (int)(t / Blk<2>::LEN), (int)(t % Blk<2>::LEN), parent, octant);
}

// k_p2c_add, then vmg_gridboundary (k_parent_fill), for the blocks [b0, b0+nb) in one launch: 8
// blocks per workgroup as in k_parent_fill.  A block's add writes only its own cells (from the
// parent level's RES), its parent-level passes read only its own cells and the parent level's U,
// so a workgroup barrier between the two is all the order they need.
template <typename T>
__global__ void __launch_bounds__(256) k_p2c_add_pfill(T *RES, T *U, int b0, int nb, int fill0, const int *nbr,
                                                       const int *nkind, const int *parent, const int *octant)
{
    constexpr int LEN = Blk<2>::LEN;
    const int bl = (int)threadIdx.x >> 5, idx = (int)threadIdx.x & 31;
    const int b = b0 + (int)blockIdx.x * 8 + bl;
    if (b < b0 + nb)
        for (int c = idx; c < LEN; c += 32) p2c_add_cell(RES, U, b, c, parent, octant);
    __syncthreads();
    parent_fill_block(U, b0, nb, fill0, nbr, nkind, octant);
}

// ---- vmg_interp_p2c_lev (trilinear): child interiors of [b0, b0+nb) from the parent's octant ----
template <int N, typename T> __global__ void k_p2c(T *Af, const T *Ac, int b0, int nb, const int *parent, const int *octant)
{
    using B = Blk<N>;
    constexpr int H = N / 2, PER = N * N * N;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nb * PER) return;
    int i = (int)(t % N), j = (int)((t / N) % N), k = (int)((t / (N * N)) % N);
    int b = b0 + (int)(t / PER);
    const T *uc = Ac + (long)parent[b] * B::LEN;
    int o = octant[b];
    int oi = (o & 1) * H, oj = ((o >> 1) & 1) * H, ok = ((o >> 2) & 1) * H;
    auto cget = [&](int a, int bb, int c) -> T { return uc[B::gi(oi + a, oj + bb, ok + c)]; };
    Af[(long)b * B::LEN + B::gi(i, j, k)] = trilinear_at(i, j, k, cget);
}

// ---- fmg_psi2g: g = -(F(i) + F(i-1))/2 on leaves, centred difference on covered blocks ----
template <int N>
__global__ void k_psi2g(double *g, const double *Au, int nblk, const double *hi_lev, const int *level,
                        const int *skip, const int *ovslot, const double *ov)
{
    using B = Blk<N>;
    constexpr int PER = N * N * N;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nblk * PER) return;
    int i = (int)(t % N), j = (int)((t / N) % N), k = (int)((t / (N * N)) % N);
    int b = (int)(t / PER);
    const double hi = hi_lev[level[b]];
    const double *u = Au + (long)b * B::LEN;
    double *gx = g + (long)b * 3 * PER, *gy = gx + PER, *gz = gy + PER;
    long q = (long)(k * N + j) * N + i;
    if (skip[b]) {
        gx[q] = -(u[B::gi(i + 1, j, k)] - u[B::gi(i - 1, j, k)]) * hi * 0.5;
        gy[q] = -(u[B::gi(i, j + 1, k)] - u[B::gi(i, j - 1, k)]) * hi * 0.5;
        gz[q] = -(u[B::gi(i, j, k + 1)] - u[B::gi(i, j, k - 1)]) * hi * 0.5;
        return;
    }
    Flux<N, double> F{u, ovslot + 6 * b, ov, hi};
    const double uc = u[B::gi(i, j, k)];
    double fxp = (u[B::gi(i + 1, j, k)] - uc) * hi, fxm = (uc - u[B::gi(i - 1, j, k)]) * hi;
    double fyp = (u[B::gi(i, j + 1, k)] - uc) * hi, fym = (uc - u[B::gi(i, j - 1, k)]) * hi;
    double fzp = (u[B::gi(i, j, k + 1)] - uc) * hi, fzm = (uc - u[B::gi(i, j, k - 1)]) * hi;
    if (i == N - 1) fxp = F.ov_or(1, j, k, fxp);
    if (i == 0) fxm = F.ov_or(0, j, k, fxm);
    if (j == N - 1) fyp = F.ov_or(3, i, k, fyp);
    if (j == 0) fym = F.ov_or(2, i, k, fym);
    if (k == N - 1) fzp = F.ov_or(5, i, j, fzp);
    if (k == 0) fzm = F.ov_or(4, i, j, fzm);
    gx[q] = -(fxp + fxm) * 0.5;
    gy[q] = -(fyp + fym) * 0.5;
    gz[q] = -(fzp + fzm) * 0.5;
}

// ---- norms over the listed blocks' interiors, two stages with a fixed order: each of G
//      workgroups reduces a grid-stride share into part[g], part[G + g], part[2G + g]
//      (max |a| h^2 as fmg_get_h2absmax, max |a|, sum a^2 dv), then one workgroup reduces those ----
template <int N>
__global__ void k_norms_part(const double *A, const int *list, int nb, const double *h_lev, const int *level,
                             double *part)
{
    using B = Blk<N>;
    constexpr int PER = N * N * N;
    double m2 = 0.0, m = 0.0, s = 0.0;
    const long total = (long)nb * PER, G = gridDim.x;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += G * blockDim.x) {
        int i = (int)(t % N), j = (int)((t / N) % N), k = (int)((t / (N * N)) % N);
        int b = list[t / PER];
        double h = h_lev[level[b]];
        double v = A[(long)b * B::LEN + B::gi(i, j, k)];
        double a = fabs(v);
        m = fmax(m, a);
        m2 = fmax(m2, a * (h * h));
        s += v * v * ((h * h) * h);
    }
    block_max_sum(m2, s);
    double dummy = 0.0;
    block_max_sum(m, dummy);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = m2;
        part[G + blockIdx.x] = m;
        part[2 * G + blockIdx.x] = s;
    }
}

__global__ void k_norms_final(const double *part, int G, double *out)
{
    double m2 = 0.0, m = 0.0, s = 0.0;
    for (int g = threadIdx.x; g < G; g += blockDim.x) {
        m2 = fmax(m2, part[g]);
        m = fmax(m, part[G + g]);
        s += part[2 * G + g];
    }
    block_max_sum(m2, s);
    double dummy = 0.0;
    block_max_sum(m, dummy);
    if (threadIdx.x == 0) { out[0] = m2; out[1] = m; out[2] = s; }
}

// ---- 8^3 blocks, one workgroup of NT threads (64, 128 or 256) per record: the block's ghosted
//      array is staged in LDS with coalesced loads, then one colour of fmg_poisson_relax / the
//      residual is formed from LDS (same operand order as k_relax / k_resid).  The record load is
//      the only load the data loads wait for; the right-hand side is fetched before the barrier.
//      Fewer threads per block let a CU hold more blocks at once (the kernel waits on load latency
//      more than on bandwidth); the AMR autotuner (amr.hip) picks NT per hierarchy. ----
__device__ __forceinline__ int rec_nb(const BlockRec &R, int f)
{
    // select chain: a runtime index into the record's array would put it in scratch memory
    int m = R.nb[0];
    m = f == 1 ? R.nb[1] : m;
    m = f == 2 ? R.nb[2] : m;
    m = f == 3 ? R.nb[3] : m;
    m = f == 4 ? R.nb[4] : m;
    m = f == 5 ? R.nb[5] : m;
    return m;
}

// The FP64 residual of the mixed solve can also write its FP32 copy (cast, in the compact
// colour-split form of cc_index, which rhsc then feeds the FP32 passes from; Aout may then be null,
// the FP64 residual not being read) and its block's norm partials (npart[g], npart[G + g], npart[2G + g] for workgroup g of G: the per-cell
// terms of k_norms_part, so the two max norms are the same values; the sum adds in another order).
//
// CC: the compact level-0 form of the inner solve (T = float, halo-free, rhsc set), everything by
// the workgroup's active-list position a: Au holds the interiors in compact form (512 a), the ghost
// cells come from the ghost store gs (gs_cell), and the residual goes to Aout in compact form.  A
// colour pass writes the boundary cells it changes into its same-level neighbours' ghost store
// (gs + 64 ptab[6 a + f]), the half of it that no concurrent pass reads.
template <bool RELAX, typename T, int NT, bool CC = false>
__global__ void __launch_bounds__(NT) k_blk8(T *Aout, T *Au, const T *Arhs, const int *rec, int ipass,
                                            LevelScalars h_lev, LevelScalars hi_lev, const T *ov, double *resmax,
                                            int halo, float *cast, double *npart, const float *rhsc, T *gs,
                                            const int *ptab, int xcd)
{
    using B = Blk<8>;
    static_assert(!CC || std::is_same<T, float>::value, "the compact form is FP32");
    constexpr int NC = RELAX ? 256 / NT : 512 / NT;   // cells per thread
    __shared__ __align__(16) T su[B::LEN];
    const int wa = xcd ? xcd_block((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x;
    BlockRec R;
    R.load(rec, wa);
    if constexpr (CC) {
        // the interiors (16-byte loads), then the ghost cells: for a colour pass only the parity its
        // stencils read
        const float4 *ug = (const float4 *)(Au + (long)wa * 512);
        for (int q4 = threadIdx.x; q4 < 128; q4 += NT) {
            const float4 v = ug[q4];
            const float e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                int i, j, k;
                cc_ijk(4 * q4 + u, i, j, k);
                su[B::gi(i, j, k)] = e[u];
            }
        }
        // (16-byte loads: four ghost cells along t1 of one t2 row per load)
        const float4 *gw = (const float4 *)(gs + (long)wa * 384);
        constexpr int NG4 = RELAX ? 48 : 96;
        for (int q4 = threadIdx.x; q4 < NG4; q4 += NT) {
            const int f = RELAX ? q4 >> 3 : q4 >> 4;
            const int par = RELAX ? (ipass + 1) & 1 : (q4 >> 3) & 1, t2 = q4 & 7;
            const int d = f >> 1, lr = f & 1, p1 = (par + (lr ? 0 : 1) + t2) & 1;
            const float4 v = gw[f * 16 + par * 8 + t2];
            su[B::rot(d, lr ? 8 : -1, p1, t2)] = v.x;
            su[B::rot(d, lr ? 8 : -1, 2 + p1, t2)] = v.y;
            su[B::rot(d, lr ? 8 : -1, 4 + p1, t2)] = v.z;
            su[B::rot(d, lr ? 8 : -1, 6 + p1, t2)] = v.w;
        }
    } else {
        // the ghosted block, 16 bytes per load (a block is 1000 values: 4000 or 8000 bytes)
        constexpr int VW = 16 / sizeof(T), NV = B::LEN / VW;
        static_assert(B::LEN % VW == 0, "the ghosted block splits into 16-byte loads");
        using V = typename std::conditional<std::is_same<T, float>::value, float4, double2>::type;
        const V *ug = (const V *)(Au + (long)R.b * B::LEN);
        for (int q = threadIdx.x; q < NV; q += NT) *(V *)(su + VW * q) = ug[q];
    }
    // this thread's cells: colour cells c = threadIdx.x + r NT (relax) or cells c (residual); in the
    // compact form the NC contiguous cells c = NC threadIdx.x + r (one row of one colour for NC = 4),
    // whose source and result move as one access
    int ci[NC], cj[NC], ck[NC];
    T rhs[NC];
    if constexpr (CC) ld_run<NC>(rhs, rhsc + (long)wa * 512 + (RELAX ? (ipass - 1) * 256 : 0) + NC * threadIdx.x);
#pragma unroll
    for (int r = 0; r < NC; ++r) {
        const int c = CC ? NC * (int)threadIdx.x + r : (int)threadIdx.x + r * NT;
        if (RELAX) {
            cj[r] = (c >> 2) & 7;
            ck[r] = c >> 5;
            ci[r] = 2 * (c & 3) + ((cj[r] + ck[r] + ipass) & 1);
        } else if (CC) {
            cc_ijk(c, ci[r], cj[r], ck[r]);
        } else {
            ci[r] = c & 7;
            cj[r] = (c >> 3) & 7;
            ck[r] = c >> 6;
        }
        if constexpr (CC) continue;
        if constexpr (std::is_same<T, float>::value) {
            if (rhsc) {
                rhs[r] = rhsc[(long)wa * 512 + (RELAX ? (ipass - 1) * 256 + c : cc_index(ci[r], cj[r], ck[r]))];
                continue;
            }
        }
        rhs[r] = Arhs[(long)R.b * B::LEN + B::gi(ci[r], cj[r], ck[r])];
    }
    if (!CC && halo) {
        // halo-free form: the same-level faces' ghost cells taken from the neighbours' boundary
        // layers (what fmg_ghfix_samelev would copy there); only the face cells the 7-point
        // operator reads, not edges or corners.  The neighbour's cells next to this block's
        // colour-ipass cells have the other colour, so a concurrent pass does not change them.
        // The barrier orders these LDS writes after the block copy above, which writes the same
        // ghost cells with their stored values.
        __syncthreads();
        for (int q = threadIdx.x; q < 6 * 64; q += NT) {
            const int f = q >> 6, t1 = q & 7, t2 = (q >> 3) & 7, d = f >> 1, lr = f & 1;
            const int m = rec_nb(R, f);
            if (m < 0) continue;
            su[B::rot(d, lr ? 8 : -1, t1, t2)] = Au[(long)m * B::LEN + B::rot(d, lr ? 0 : 7, t1, t2)];
        }
    }
    __syncthreads();
    const T h = (T)h_lev.v[R.l], hi = (T)hi_lev.v[R.l];
    FluxR<8, T> F{su, nullptr, ov, hi, R};
    double a = 0.0, m2 = 0.0, sq = 0.0;
    const double hd = h_lev.v[R.l];
    [[maybe_unused]] T vout[NC];   // the compact form's results, stored as one run after the loop
#pragma unroll
    for (int r = 0; r < NC; ++r) {
        const long q = (long)R.b * B::LEN + B::gi(ci[r], cj[r], ck[r]);
        if (RELAX) {
            T res = F.div(ci[r], cj[r], ck[r]) - rhs[r] * h;
            T resh2 = res * h;
            const T v = su[B::gi(ci[r], cj[r], ck[r])] + (T)(1.0 / 6.0) * resh2;
            if constexpr (CC) {
                vout[r] = v;
                // x faces here; the y and z faces take the thread's run of cells as one store below
                const int i = ci[r], hp = ((i + cj[r] + ck[r]) & 1) * 32;
                if (i == 0 && ptab[wa * 6] >= 0) gs[(long)ptab[wa * 6] * 64 + hp + (cj[r] >> 1) + 4 * ck[r]] = v;
                if (i == 7 && ptab[wa * 6 + 1] >= 0) gs[(long)ptab[wa * 6 + 1] * 64 + hp + (cj[r] >> 1) + 4 * ck[r]] = v;
            } else {
                Au[q] = v;
            }
            a = fmax(a, (double)fabs(resh2));
        } else {
            const T v = rhs[r] - F.div(ci[r], cj[r], ck[r]) * hi;
            if constexpr (CC) vout[r] = v;
            else if (Aout) Aout[q] = v;
            if constexpr (std::is_same<T, double>::value) {
                if (cast) cast[(long)wa * 512 + cc_index(ci[r], cj[r], ck[r])] = (float)v;
                if (npart) {
                    const double av = fabs(v);
                    a = fmax(a, av);
                    m2 = fmax(m2, av * (hd * hd));
                    sq += v * v * ((hd * hd) * hd);
                }
            }
        }
    }
    if constexpr (CC) {
        st_run<NC>((RELAX ? Au + (ipass - 1) * 256 : Aout) + (long)wa * 512 + NC * threadIdx.x, vout);
        if constexpr (RELAX) {
            // the run is NC cells of one row (fixed j, k; i = ci[0], ci[0] + 2, ...) of one parity:
            // on a y or z face they are NC contiguous cells of the neighbour's ghost store half
            const int j = cj[0], k = ck[0], hp = ((ci[0] + j + k) & 1) * 32, x0 = ci[0] >> 1;
            auto run = [&](int f, int t2) {
                const int tg = ptab[wa * 6 + f];
                if (tg >= 0) st_run<NC>(gs + (long)tg * 64 + hp + x0 + 4 * t2, vout);
            };
            if (j == 0) run(2, k);
            if (j == 7) run(3, k);
            if (k == 0) run(4, j);
            if (k == 7) run(5, j);
        }
    }
    if (!RELAX && npart) {
        const int G = gridDim.x;
        block_max_sum(m2, sq);
        double dummy = 0.0;
        block_max_sum(a, dummy);
        if (threadIdx.x == 0) {
            npart[wa] = m2;
            npart[G + wa] = a;
            npart[2 * G + wa] = sq;
        }
    }
    if (RELAX && resmax) block_atomic_max(resmax + (blockIdx.x & (kMaxSlots - 1)), a);
}

// fmg_addint at FMG level 0 in the compact form: U += the trilinear interpolation of the listed
// blocks' FMG-level-1 correction (fmg_interp_linear, rounded to FP32 as the staged sequence stores
// it before fmg_add), one workgroup per block (U by list position), which also writes the new
// boundary cells into the same-level neighbours' ghost stores.  Only the interiors: the compact
// form has no ghost cells to add to.
__global__ void __launch_bounds__(256) k_addint_cc(float *U, const float *Ac, const int *list, float *gs,
                                                   const int *ptab)
{
    const int a = xcd_block((int)blockIdx.x, (int)gridDim.x), b = list[a];   // per XCD a contiguous range
    const float *uc = Ac + (long)b * Blk<4>::LEN;
    auto cget = [&](int x, int y, int z) -> float { return uc[Blk<4>::gi(x, y, z)]; };
    for (int q = threadIdx.x; q < 512; q += 256) {
        int i, j, k;
        cc_ijk(q, i, j, k);
        const float e = trilinear_at(i, j, k, cget);
        const float v = U[(long)a * 512 + q] + e;
        U[(long)a * 512 + q] = v;
        push_faces(gs, ptab + a * 6, i, j, k, v);
    }
}

// the FP64 update of the mixed solve from the compact FP32 correction: the listed blocks' interiors
// (threads in the FP64 array's x-fastest order, so its rows are read and written whole)
__global__ void k_add_cast_cc(double *A, const float *U, const int *list, int nb)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nb * 512) return;
    const int b = list[t >> 9], c = (int)(t & 511);
    const int i = c & 7, j = (c >> 3) & 7, k = c >> 6;
    const long q = (long)b * Blk<8>::LEN + Blk<8>::gi(i, j, k);
    A[q] = A[q] + (double)U[(t >> 9) * 512 + cc_index(i, j, k)];
}

// ---- single-workgroup helpers (blockDim.x threads, one workgroup) ----
// Pairwise halving over a[0..P) after zero-padding a[m..P): the fixed summation order of block
// partials shared with the CPU restatement.
__device__ inline void wg_tree_sum(double *a, int m, int P)
{
    for (int q = m + (int)threadIdx.x; q < P; q += blockDim.x) a[q] = 0.0;
    __syncthreads();
    for (int st = P / 2; st >= 1; st >>= 1) {
        for (int q = threadIdx.x; q < st; q += blockDim.x) a[q] = a[q] + a[q + st];
        __syncthreads();
    }
}

// vmg_zeroAverage (vmg.F90:890-915): average of the base blocks' interiors removed from them.
template <int N, typename T> __global__ void k_zero_average(T *A, int nbase, int P, double *scratch)
{
    using B = Blk<N>;
    for (int b = threadIdx.x; b < nbase; b += blockDim.x) {
        const T *a = A + (long)b * B::LEN;
        double t = 0.0;
        for (int k = 0; k < N; ++k)
            for (int j = 0; j < N; ++j)
                for (int i = 0; i < N; ++i) t = t + a[B::gi(i, j, k)];
        scratch[b] = t;
    }
    __syncthreads();
    wg_tree_sum(scratch, nbase, P);
    const double ave = scratch[0] / (double)(N * N * N * nbase);
    for (int q = threadIdx.x; q < nbase * N * N * N; q += blockDim.x) {
        int b = q / (N * N * N), c = q % (N * N * N);
        long e = (long)b * B::LEN + B::gi(c % N, (c / N) % N, c / (N * N));
        A[e] = (T)((double)A[e] - ave);
    }
}

// fmg_make_physvar mean (fmg_data_if.F90:53-78): per-block partials of mass and volume over leaves
template <int N>
__global__ void k_leaf_mass(const double *A, int nblk, const int *covered, const double *h_lev, const int *level,
                            double *pm, double *pv)
{
    using B = Blk<N>;
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblk) return;
    if (covered[b]) { pm[b] = 0.0; pv[b] = 0.0; return; }
    double h = h_lev[level[b]];
    double dv = (h * h) * h;
    const double *a = A + (long)b * B::LEN;
    double t = 0.0;
    for (int k = 0; k < N; ++k)
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i) t = t + a[B::gi(i, j, k)];
    pm[b] = t * dv;
    pv[b] = (double)(N * N * N) * dv;
}

__global__ void k_tree_mass(double *pm, double *pv, int nblk, int P, double *ave)
{
    wg_tree_sum(pm, nblk, P);
    wg_tree_sum(pv, nblk, P);
    if (threadIdx.x == 0) *ave = pm[0] / pv[0];
}

__global__ void k_sub_scalar(double *A, long len, const double *s)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= len) return;
    A[t] = A[t] - *s;
}

// ---- fmg_boundary_phys_sphere ----
struct Geo {
    const int *ijk, *level;
    const double *h0_lev;   // cell width per AMR level at FMG level 0
    double x0, y0, z0;
    __device__ double x(int b, int a, int i) const
    {
        double o = a == 0 ? x0 : (a == 1 ? y0 : z0);
        return o + ((double)(ijk[3 * b + a] * 8 + i) + 0.5) * h0_lev[level[b]];
    }
};

template <int N> __global__ void k_sphere_mask(double *A, int nblk, Geo G, double rad2)
{
    using B = Blk<N>;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nblk * B::LEN) return;
    int b = (int)(t / B::LEN), q = (int)(t % B::LEN);
    int i = q % B::W - 1, j = (q / B::W) % B::W - 1, k = q / (B::W * B::W) - 1;
    double x = G.x(b, 0, i), y = G.x(b, 1, j), z = G.x(b, 2, k);
    if (x * x + y * y + z * z >= rad2) A[t] = 0.0;
}

// multipole moments (l <= 3, real form) of one base block, cells in (k, j, i) order
template <int N> __global__ void k_sphere_moments(const double *A, int nbase, int P, Geo G, double dv, double *part)
{
    using B = Blk<N>;
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbase) return;
    const double *rho = A + (long)b * B::LEN;
    double a[16];
    for (int q = 0; q < 16; ++q) a[q] = 0.0;
    for (int k = 0; k < N; ++k)
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i) {
                double x = G.x(b, 0, i), y = G.x(b, 1, j), z = G.x(b, 2, k);
                double r2 = x * x + y * y + z * z;
                double r = sqrt(r2);
                double r3 = r2 * r;
                double rc = sqrt(x * x + y * y);
                double ri = 1.0 / r, rci = 1.0 / rc;
                double cost = z * ri, sint = rc * ri, cosp = x * rci, sinp = y * rci;
                double c2 = cost * cost, s2 = sint * sint, s3 = s2 * sint;
                double cp2 = cosp * cosp, sp2 = sinp * sinp;
                double rdv = rho[B::gi(i, j, k)] * dv;
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
    for (int q = 0; q < 16; ++q) part[(long)q * P + b] = a[q];
}

__global__ void k_sphere_coef(double *part, int nbase, int P, double *coef)
{
    for (int q = 0; q < 16; ++q) wg_tree_sum(part + (long)q * P, nbase, P);
    if (threadIdx.x != 0) return;
    const double pi = 3.14159265358979323846;
    double A[16];
    for (int q = 0; q < 16; ++q) A[q] = part[(long)q * P];
    coef[0] = -A[0] / (4.0 * pi);
    coef[1] = -A[1] / (4.0 * pi);
    coef[2] = -A[2] / (8.0 * pi);
    coef[3] = -A[3] / (8.0 * pi);
    coef[4] = -A[4] / (16.0 * pi);
    coef[5] = -A[5] / (8.0 * pi) * 3.0;
    coef[6] = -A[6] / (8.0 * pi) * 3.0;
    coef[7] = -A[7] / (32.0 * pi) * 3.0;
    coef[8] = -A[8] / (32.0 * pi) * 3.0;
    coef[9] = -A[9] / (16.0 * pi);
    coef[10] = -A[10] / (64.0 * pi) * 3.0;
    coef[11] = -A[11] / (64.0 * pi) * 3.0;
    coef[12] = -A[12] / (32.0 * pi) * 15.0;
    coef[13] = -A[13] / (32.0 * pi) * 15.0;
    coef[14] = -A[14] / (64.0 * pi) * 5.0;
    coef[15] = -A[15] / (64.0 * pi) * 5.0;
}

template <int N> __global__ void k_sphere_ghost(double *A, const int *faces, int nf, Geo G, const double *coef)
{
    using B = Blk<N>;
    constexpr int W = B::W;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nf * W * W) return;
    int t1 = (int)(t % W) - 1, t2 = (int)((t / W) % W) - 1, e = faces[t / (W * W)];
    int b = e / 6, f = e % 6, d = f / 2, ag = (f & 1) ? N : -1;
    int ii = d == 0 ? ag : t1, jj = d == 1 ? ag : (d == 0 ? t1 : t2), kk = d == 2 ? ag : t2;
    double x = G.x(b, 0, ii), y = G.x(b, 1, jj), z = G.x(b, 2, kk);
    const double *c = coef;
    double r2i = 1.0 / (x * x + y * y + z * z);
    double ri = sqrt(r2i);
    double r3i = r2i * ri, r4i = r2i * r2i;
    double rc = sqrt(x * x + y * y);
    double rci = 1.0 / rc;
    double cost = z * ri, sint = rc * ri, cosp = x * rci, sinp = y * rci;
    double c2 = cost * cost, s2 = sint * sint, s3 = s2 * sint;
    double cp2 = cosp * cosp, sp2 = sinp * sinp;
    double ub = ri * c[0];
    ub = ub + r2i * cost * c[1];
    ub = ub + r2i * sint * (cosp * c[2] + sinp * c[3]) * 2.0;
    ub = ub + r3i * (3.0 * c2 - 1.0) * c[4];
    ub = ub + r3i * sint * cost * (cosp * c[5] + sinp * c[6]) * 2.0;
    ub = ub + r3i * s2 * ((cp2 - sp2) * c[7] + 2.0 * sinp * cosp * c[8]) * 2.0;
    ub = ub + r4i * cost * (5.0 * c2 - 3.0) * c[9];
    ub = ub + r4i * sint * (5.0 * c2 - 1.0) * (cosp * c[10] + sinp * c[11]) * 2.0;
    ub = ub + r4i * cost * s2 * ((cp2 - sp2) * c[12] + 2.0 * sinp * cosp * c[13]) * 2.0;
    ub = ub + r4i * s3 * (cosp * (cp2 - 3.0 * sp2) * c[14] + sinp * (3.0 * cp2 - sp2) * c[15]) * 2.0;
    A[(long)b * B::LEN + B::rot(d, ag, t1, t2)] = ub;
}

// ---- base grid: gather of the base blocks' interiors (mg_dataPrepare) and scatter back ----
template <int N, typename T> __global__ void k_base_gather(T *psi, T *src, const T *Au, const T *Arhs,
                                              int nbase, const int *ijk, int NG)
{
    using B = Blk<N>;
    constexpr int PER = N * N * N;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nbase * PER) return;
    int i = (int)(t % N), j = (int)((t / N) % N), k = (int)((t / (N * N)) % N), b = (int)(t / PER);
    long W = NG + 2;
    long q = ((long)(ijk[3 * b + 2] * N + k + 1) * W + (ijk[3 * b + 1] * N + j + 1)) * W + (ijk[3 * b] * N + i + 1);
    psi[q] = Au[(long)b * B::LEN + B::gi(i, j, k)];
    src[q] = Arhs[(long)b * B::LEN + B::gi(i, j, k)];
}

template <int N, typename T> __global__ void k_base_scatter(T *Au, const T *psi, int nbase, const int *ijk, int NG)
{
    using B = Blk<N>;
    constexpr int PER = N * N * N;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nbase * PER) return;
    int i = (int)(t % N), j = (int)((t / N) % N), k = (int)((t / (N * N)) % N), b = (int)(t / PER);
    long W = NG + 2;
    long q = ((long)(ijk[3 * b + 2] * N + k + 1) * W + (ijk[3 * b + 1] * N + j + 1)) * W + (ijk[3 * b] * N + i + 1);
    Au[(long)b * B::LEN + B::gi(i, j, k)] = psi[q];
}

// ---- mixed precision: FP64 residual -> FP32 inner source, FP64 potential += FP32 correction ----
__global__ void k_cast_blocks(float *O, const double *I, const int *list, int nb, int len)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nb * len) return;
    long q = (long)pick_block(list, 0, (int)(t / len)) * len + t % len;
    O[q] = (float)I[q];
}

__global__ void k_add_cast_blocks(double *A, const float *D, const int *list, int nb, int len)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nb * len) return;
    long q = (long)pick_block(list, 0, (int)(t / len)) * len + t % len;
    A[q] = A[q] + (double)D[q];
}

} // namespace amr
} // namespace sg
