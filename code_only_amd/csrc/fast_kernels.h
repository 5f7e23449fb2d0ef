// Fused FP32 kernels of the mixed-precision inner solve (gfx950).
//
//  * k_sweep_tile: one full red-black sweep (mg_poisson.F90:55-74, both colours) of a TX*TY*TZ
//    tile held in LDS with an H-cell halo, optionally
//      - starting from u = 0 (the zero fill of mg.F90:189 folded into the load),
//      - adding the trilinear coarse-grid correction on load (mg_addint, mg.F90:727-731), the
//        coarse tile staged in LDS first,
//      - finishing with the residual (mg_poisson.F90:17-27) restricted to the coarse grid
//        (restriction.F90:132-147), which needs H = 3.
//    Reads u from one buffer and writes the tile core to another: a workgroup's halo is then
//    never a value another workgroup has already updated.  The redundant halo updates repeat the
//    owner's arithmetic exactly, so the result equals the two-launch colour sweep bit for bit.
//    Tile shape and workgroup size are template parameters; the host picks one per level size by
//    timing (all shapes give identical values).
//  * k_interp_cubic_tile: the FMG tricubic prolongation (interpolation.F90:98-217) as the
//    reference's three separable passes, on an LDS-staged coarse tile.
//  * k_coarse: every level of <= 16^3 cells in one workgroup's LDS: the FMG part of mg_lin
//    (mg.F90:142-215) from the coarsest level up to the first fine level, or one V-cycle from that
//    level down and back (the bottom of each fine-level V-cycle); or, in both places, the exact
//    solve of the first level of <= 16^3 cells by Hartley transforms (coarse_exact).
//
// Every cell value is produced by the same helpers (relax_value, resid_value, restrict_value,
// trilinear_value, cubic_stencil) as the one-kernel-per-loop path, in the same order, so the fused
// solve reproduces the unfused FP32 solve exactly.
#pragma once
#include <cmath>

#include "uniform_kernels.h"

namespace sg {

constexpr int kCoarseMaxN = 16;
constexpr int kCoarseThreads = 1024;

__device__ __forceinline__ int wrapn(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// Trilinear prolongation value at fine cell (i, j, k) of a periodic coarse grid of nc cells;
// C(a, b, d) returns coarse cell (a, b, d).
template <typename T, typename Acc>
__device__ __forceinline__ T trilinear_point(const Acc &C, int i, int j, int k, int nc)
{
    const int ic = i >> 1, jc = j >> 1, kc = k >> 1;
    const int ic1 = wrap_g(ic + 2 * (i & 1) - 1, nc);
    const int jc1 = wrap_g(jc + 2 * (j & 1) - 1, nc);
    const int kc1 = wrap_g(kc + 2 * (k & 1) - 1, nc);
    return trilinear_value<T>(C(ic, jc, kc), C(ic1, jc, kc), C(ic, jc1, kc), C(ic, jc, kc1), C(ic, jc1, kc1),
                              C(ic1, jc, kc1), C(ic1, jc1, kc), C(ic1, jc1, kc1));
}

// 4-point cubic combination in the reference's order: (a0 v0 + a1 v1) + a2 v2 + a3 v3.
template <typename T> __device__ __forceinline__ T cubic4(const T w[4], T v0, T v1, T v2, T v3)
{
    T s = w[0] * v0 + w[1] * v1;
    s = s + w[2] * v2;
    s = s + w[3] * v3;
    return s;
}

// Tricubic value at fine cell (i, j, k) (interpolation.F90:98-217), nc >= 2.
template <typename T, typename Acc>
__device__ __forceinline__ T tricubic_point(const Acc &C, int i, int j, int k, int nc)
{
    int ix[4], iy[4], iz[4];
    double ax[4], ay[4], az[4];
    cubic_stencil(i, -1, nc, ix, ax);
    cubic_stencil(j, -1, nc, iy, ay);
    cubic_stencil(k, -1, nc, iz, az);
    T wx[4], wy[4], wz[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        wx[q] = (T)ax[q]; wy[q] = (T)ay[q]; wz[q] = (T)az[q];
        ix[q] = wrap_g(ix[q], nc); iy[q] = wrap_g(iy[q], nc); iz[q] = wrap_g(iz[q], nc);
    }
    T bz[4];
#pragma unroll
    for (int kq = 0; kq < 4; ++kq) {
        T by[4];
#pragma unroll
        for (int jq = 0; jq < 4; ++jq)
            by[jq] = cubic4<T>(wx, C(ix[0], iy[jq], iz[kq]), C(ix[1], iy[jq], iz[kq]), C(ix[2], iy[jq], iz[kq]),
                               C(ix[3], iy[jq], iz[kq]));
        bz[kq] = cubic4<T>(wy, by[0], by[1], by[2], by[3]);
    }
    return cubic4<T>(wz, bz[0], bz[1], bz[2], bz[3]);
}

// Tile index with XCD-aware placement: consecutive workgroups are dealt round-robin to the 8
// XCDs, so workgroup b is given tile (b % 8) * per + b / 8 -- each XCD works through a contiguous
// range of tiles (z-slabs), and the halos its tiles share stay in that XCD's L2.
__device__ __forceinline__ int xcd_tile(int ntiles)
{
    const int per = (ntiles + 7) / 8;
    const int b = (int)blockIdx.x;
    return (b % 8) * per + b / 8;
}

template <int TX_, int TY_, int TZ_, int NT_> struct Tile {
    static constexpr int TX = TX_, TY = TY_, TZ = TZ_, NT = NT_;
};

// Tile shapes the autotuners choose from (all give identical values), each with 1 or 2 sweeps
// per launch.
using Tile0 = Tile<32, 16, 16, 1024>;
using Tile1 = Tile<16, 8, 8, 1024>;
using Tile2 = Tile<16, 8, 8, 256>;
using Tile3 = Tile<64, 8, 8, 512>;
using Tile4 = Tile<8, 8, 8, 512>;
constexpr int kNumTiles = 5;
constexpr int kMaxSweepsPerLaunch = 2;
static const int kTileDims[kNumTiles][4] = {
    {32, 16, 16, 1024}, {16, 8, 8, 1024}, {16, 8, 8, 256}, {64, 8, 8, 512}, {8, 8, 8, 512}};

inline bool tile_fits(int cfg, int n)
{
    const int *d = kTileDims[cfg];
    return n >= d[0] && n >= d[1] && n >= d[2] && n % d[0] == 0 && n % d[1] == 0 && n % d[2] == 0;
}

// One colour of the sweep over the LDS region shrunk by margin M: cells with
// (gx + gy + gz) % 2 == colour, where colour 1 is the reference's first pass (ipass = 1).  The
// source values are fetched for all of a thread's cells before any update.
template <class TL, int LX, int LY, int LZ, int M, int H, bool FS>
__device__ __forceinline__ void tile_colour(float *su, const float *sf, const float *__restrict__ f, int n, int x0,
                                            int y0, int z0, int colour, float h2, double &amax)
{
    constexpr int RX = (LX - 2 * M) / 2, RY = LY - 2 * M, RZ = LZ - 2 * M;
    constexpr int CNT = RX * RY * RZ;
    constexpr int IT = (CNT + TL::NT - 1) / TL::NT;
    static_assert((LX - 2 * M) % 2 == 0, "region x extent must be even");
    static_assert(IT <= 32, "too many cells per thread");
    float fv[IT];
    int cv[IT];
    unsigned core = 0u;
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int q = threadIdx.x + it * TL::NT;
        cv[it] = -1;
        fv[it] = 0.0f;
        if (CNT % TL::NT == 0 || q < CNT) {
            const int xh = q % RX;
            const int y = M + (q / RX) % RY;
            const int z = M + q / (RX * RY);
            const int d = (colour + x0 + M + y0 + y + z0 + z) & 1;
            const int x = M + 2 * xh + d;
            cv[it] = (z * LY + y) * LX + x;
            if (x >= H && x < H + TL::TX && y >= H && y < H + TL::TY && z >= H && z < H + TL::TZ) core |= 1u << it;
            if (FS) {
                fv[it] = sf[cv[it]];
            } else {
                const int gi = wrapn(x0 + x, n), gj = wrapn(y0 + y, n), gk = wrapn(z0 + z, n);
                fv[it] = f[(gk * n + gj) * n + gi];
            }
        }
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int c = cv[it];
        if (c >= 0) {
            float s;
            su[c] = relax_value<float>(su[c], su[c + 1], su[c - 1], su[c + LX], su[c - LX], su[c + LX * LY],
                                       su[c - LX * LY], fv[it], h2, &s);
            if (core & (1u << it)) amax = fmax(amax, fabs((double)s));
        }
    }
}

// Sweeps S .. NS-1 of a multi-sweep tile: sweep s updates the first colour on margin 2s+1 and the
// second on margin 2s+2.  Only the last sweep feeds the max|resh2| (one mg_relax call's Resmaxg).
template <class TL, int LX, int LY, int LZ, int H, int S, int NS, bool FS> struct TileSweeps {
    static __device__ __forceinline__ void run(float *su, const float *sf, const float *__restrict__ f, int n, int x0,
                                               int y0, int z0, float h2, double &amax)
    {
        double dummy = 0.0;
        double &acc = (S == NS - 1) ? amax : dummy;
        tile_colour<TL, LX, LY, LZ, 2 * S + 1, H, FS>(su, sf, f, n, x0, y0, z0, 1, h2, acc);
        __syncthreads();
        tile_colour<TL, LX, LY, LZ, 2 * S + 2, H, FS>(su, sf, f, n, x0, y0, z0, 0, h2, acc);
        __syncthreads();
        TileSweeps<TL, LX, LY, LZ, H, S + 1, NS, FS>::run(su, sf, f, n, x0, y0, z0, h2, amax);
    }
};
template <class TL, int LX, int LY, int LZ, int H, int NS, bool FS> struct TileSweeps<TL, LX, LY, LZ, H, NS, NS, FS> {
    static __device__ __forceinline__ void run(float *, const float *, const float *__restrict__, int, int, int, int,
                                               float, double &)
    {
    }
};

// LDS bytes of the largest variant of a (tile, sweeps, staging) choice: the residual variant has
// the widest halo, the prolongation variant adds the staged coarse tile.
template <class TL, int NS, bool FS> constexpr int sweep_lds_bytes()
{
    constexpr int H1 = 2 * NS + 1, H2 = 2 * NS;
    constexpr int r1 = (TL::TX + 2 * H1) * (TL::TY + 2 * H1) * (TL::TZ + 2 * H1) * (FS ? 2 : 1);
    constexpr int r2 = (TL::TX + 2 * H2) * (TL::TY + 2 * H2) * (TL::TZ + 2 * H2) * (FS ? 2 : 1) +
                       ((TL::TX + 2 * H2) / 2 + 2) * ((TL::TY + 2 * H2) / 2 + 2) * ((TL::TZ + 2 * H2) / 2 + 2);
    return 4 * (r1 > r2 ? r1 : r2) + 1024;
}
template <class TL, int NS, bool FS> constexpr bool sweep_fits_lds() { return sweep_lds_bytes<TL, NS, FS>() <= 163840; }

template <class TL> inline bool fits_lds(int ns, bool fs)
{
    if (ns == 2) return fs ? sweep_fits_lds<TL, 2, true>() : sweep_fits_lds<TL, 2, false>();
    return fs ? sweep_fits_lds<TL, 1, true>() : sweep_fits_lds<TL, 1, false>();
}

// whether (tile shape, sweeps per launch, staging) is instantiated (fits the 160 KB of LDS)
inline bool cand_ok(int cfg, int ns, bool fs)
{
    switch (cfg) {
    case 0: return fits_lds<Tile0>(ns, fs);
    case 1: return fits_lds<Tile1>(ns, fs);
    case 2: return fits_lds<Tile2>(ns, fs);
    case 3: return fits_lds<Tile3>(ns, fs);
    default: return fits_lds<Tile4>(ns, fs);
    }
}

// NS full sweeps (temporal blocking: halo 2 per sweep, +1 for the residual).  FS stages the source
// on the same LDS region as u, loaded together with u, so the colour passes touch LDS only.
// Body of the fused sweep for tile t of an n^3 level (n even, divisible by the tile extents);
// called by every thread of a workgroup with the same t.
template <class TL, int NS, bool ZERO, bool PROLONG, bool RESID, bool FS>
__device__ __forceinline__ void sweep_tile(const float *__restrict__ uin, float *__restrict__ uout,
                                           const float *__restrict__ f, int n, float h2, float h2i,
                                           const float *__restrict__ uc, float *__restrict__ rc, double *resmax, int t)
{
    constexpr int H = 2 * NS + (RESID ? 1 : 0);
    constexpr int TX = TL::TX, TY = TL::TY, TZ = TL::TZ, NT = TL::NT;
    constexpr int LX = TX + 2 * H, LY = TY + 2 * H, LZ = TZ + 2 * H;
    constexpr int CX = LX / 2 + 2, CY = LY / 2 + 2, CZ = LZ / 2 + 2;  // staged coarse region
    static_assert(!RESID || H >= 3, "residual needs a 3-cell halo");
    static_assert(H % 2 == 0 || !PROLONG, "prolongation staging assumes an even tile origin");
    __shared__ float su[LX * LY * LZ];
    __shared__ float sf[FS ? LX * LY * LZ : 1];
    __shared__ float sc[PROLONG ? CX * CY * CZ : 1];
    const int ntx = n / TX, nty = n / TY;
    const int tx = t % ntx, ty = (t / ntx) % nty, tz = t / (ntx * nty);
    const int x0 = tx * TX - H, y0 = ty * TY - H, z0 = tz * TZ - H;
    const int nc = n >> 1;

    {
        // one batch of global loads: u and (FS) the source on tile + halo, (PROLONG) the coarse
        // cells x0/2 - 1 .. x0/2 + LX/2 (x0 even) that the trilinear correction reads
        constexpr int CNT = LX * LY * LZ, IT = (CNT + NT - 1) / NT;
        constexpr int CNTC = CX * CY * CZ, ITC = PROLONG ? (CNTC + NT - 1) / NT : 1;
        float v[IT];
        float fvl[FS ? IT : 1];
        float cvl[ITC];
        if (!ZERO || FS) {
#pragma unroll
            for (int it = 0; it < IT; ++it) {
                const int q = threadIdx.x + it * NT;
                if (CNT % NT == 0 || q < CNT) {
                    const int x = q % LX, y = (q / LX) % LY, z = q / (LX * LY);
                    const int g = (wrapn(z0 + z, n) * n + wrapn(y0 + y, n)) * n + wrapn(x0 + x, n);
                    if (!ZERO) v[it] = uin[g];
                    if (FS) fvl[FS ? it : 0] = f[g];
                }
            }
        }
        if (PROLONG) {
            const int cx0 = x0 / 2 - 1, cy0 = y0 / 2 - 1, cz0 = z0 / 2 - 1;
#pragma unroll
            for (int it = 0; it < ITC; ++it) {
                const int q = threadIdx.x + it * NT;
                if (CNTC % NT == 0 || q < CNTC) {
                    const int x = q % CX, y = (q / CX) % CY, z = q / (CX * CY);
                    cvl[it] = uc[(wrapn(cz0 + z, nc) * nc + wrapn(cy0 + y, nc)) * nc + wrapn(cx0 + x, nc)];
                }
            }
#pragma unroll
            for (int it = 0; it < ITC; ++it) {
                const int q = threadIdx.x + it * NT;
                if (CNTC % NT == 0 || q < CNTC) sc[q] = cvl[it];
            }
        }
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int q = threadIdx.x + it * NT;
            if (CNT % NT == 0 || q < CNT) {
                su[q] = ZERO ? 0.0f : v[it];
                if (FS) sf[q] = fvl[FS ? it : 0];
            }
        }
    }
    __syncthreads();
    if (PROLONG) {
        // u += trilinear(coarse) on tile + halo (mg_addint, mg.F90:727-731)
        constexpr int CNT = LX * LY * LZ, IT = (CNT + NT - 1) / NT;
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int q = threadIdx.x + it * NT;
            if (CNT % NT == 0 || q < CNT) {
                const int x = q % LX, y = (q / LX) % LY, z = q / (LX * LY);
                // local coarse index of the parent: floor(x/2) + 1; neighbour toward the cell
                const int ic = (x >> 1) + 1, jc = (y >> 1) + 1, kc = (z >> 1) + 1;
                const int ic1 = ic + 2 * (x & 1) - 1, jc1 = jc + 2 * (y & 1) - 1, kc1 = kc + 2 * (z & 1) - 1;
                auto C = [&](int a, int b, int d) { return sc[(d * CY + b) * CX + a]; };
                su[q] = su[q] + trilinear_value<float>(C(ic, jc, kc), C(ic1, jc, kc), C(ic, jc1, kc), C(ic, jc, kc1),
                                                       C(ic, jc1, kc1), C(ic1, jc, kc1), C(ic1, jc1, kc),
                                                       C(ic1, jc1, kc1));
            }
        }
        __syncthreads();
    }
    double amax = 0.0;
    TileSweeps<TL, LX, LY, LZ, H, 0, NS, FS>::run(su, sf, f, n, x0, y0, z0, h2, amax);
    {
        constexpr int CNT = TX * TY * TZ, IT = (CNT + NT - 1) / NT;
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int q = threadIdx.x + it * NT;
            if (CNT % NT == 0 || q < CNT) {
                const int x = q % TX, y = (q / TX) % TY, z = q / (TX * TY);
                const int gi = x0 + H + x, gj = y0 + H + y, gk = z0 + H + z;
                uout[(gk * n + gj) * n + gi] = su[((z + H) * LY + (y + H)) * LX + (x + H)];
            }
        }
    }
    if (RESID) {
        // residual of the 8 fine cells of each coarse cell of the tile, restricted
        constexpr int QX = TX / 2, QY = TY / 2, QZ = TZ / 2;
        constexpr int CNT = QX * QY * QZ, IT = (CNT + NT - 1) / NT;
        float fv[IT][8];
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int q = threadIdx.x + it * NT;
            if (CNT % NT == 0 || q < CNT) {
                const int cx = q % QX, cy = (q / QX) % QY, cz = q / (QX * QY);
#pragma unroll
                for (int o = 0; o < 8; ++o) {
                    const int a = o & 1, b = (o >> 1) & 1, d = (o >> 2) & 1;
                    if (FS) {
                        fv[it][o] = sf[((H + 2 * cz + d) * LY + (H + 2 * cy + b)) * LX + (H + 2 * cx + a)];
                    } else {
                        const int gi = x0 + H + 2 * cx + a, gj = y0 + H + 2 * cy + b, gk = z0 + H + 2 * cz + d;
                        fv[it][o] = f[(gk * n + gj) * n + gi];
                    }
                }
            }
        }
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int q = threadIdx.x + it * NT;
            if (CNT % NT == 0 || q < CNT) {
                const int cx = q % QX, cy = (q / QX) % QY, cz = q / (QX * QY);
                float r[8];
#pragma unroll
                for (int o = 0; o < 8; ++o) {
                    const int a = o & 1, b = (o >> 1) & 1, d = (o >> 2) & 1;
                    const int c = ((H + 2 * cz + d) * LY + (H + 2 * cy + b)) * LX + (H + 2 * cx + a);
                    r[o] = resid_value<float>(su[c], su[c + 1], su[c - 1], su[c + LX], su[c - LX], su[c + LX * LY],
                                              su[c - LX * LY], fv[it][o], h2i);
                }
                // restriction.F90:140-143 order: 000 100 010 001 110 101 011 111
                const int gci = tx * QX + cx, gcj = ty * QY + cy, gck = tz * QZ + cz;
                rc[(gck * nc + gcj) * nc + gci] =
                    restrict_value<float>(r[0], r[1], r[2], r[4], r[3], r[5], r[6], r[7]);
            }
        }
    }
    if (resmax) block_atomic_max(resmax, amax);
}

template <class TL, int NS, bool ZERO, bool PROLONG, bool RESID, bool FS>
__global__ void __launch_bounds__(TL::NT)
k_sweep_tile(const float *__restrict__ uin, float *__restrict__ uout, const float *__restrict__ f, int n, float h2,
             float h2i, const float *__restrict__ uc, float *__restrict__ rc, double *resmax)
{
    const int ntiles = (n / TL::TX) * (n / TL::TY) * (n / TL::TZ);
    const int t = xcd_tile(ntiles);
    if (t >= ntiles) return;  // uniform per workgroup, before any barrier
    sweep_tile<TL, NS, ZERO, PROLONG, RESID, FS>(uin, uout, f, n, h2, h2i, uc, rc, resmax, t);
}

// ------------------------------------------------------------------------------------------------
// Streaming form of the fused sweep: a workgroup owns an SX x SY column of the grid and a chunk of
// ZC planes, and marches up the chunk with the planes it needs in an LDS ring of R planes (u and
// f).  Step p:
//   1. first colour (i+j+k odd, the reference's ipass 1) of plane p, on the column with margin 1;
//   -- barrier --
//   2. plane p + 2 (fetched into registers one step earlier, with the trilinear correction added
//      for PROLONG) into the slot of plane p - 3 (p - 4 with RESID), plane p + 3 fetched into the
//      registers; second colour of plane p - 1, margin 2, whose first-colour neighbours on planes
//      p - 2, p - 1 and p are all new: plane p - 1 is final;
//   -- barrier --
//   3. plane p - 1 stored; with RESID, the residual of plane p - 2 (final, as are p - 3 and
//      p - 1), restricted pairwise (restriction.F90:132-147 operand order).
// Phase 3 overlaps the next step's first colour, which touches none of its planes.  Every
// thread's cells and global offsets are fixed for the whole march and computed once; the march is
// unrolled by R, so the ring slots are immediate LDS offsets.  Halo: H = 2 cells in x and y, H
// planes below and above the chunk (H = 3 with RESID).  max|resh2| is kept in FP32 only when
// asked for (the max of the FP32 values, widened once, equals the max of their widened values).
// Every cell update sees the same operands as the two-launch colour passes, so the values are
// those of the tile kernel, bit for bit.
// ------------------------------------------------------------------------------------------------
template <int SX_, int SY_, int ZC_, int NT_> struct StreamShape {
    static constexpr int SX = SX_, SY = SY_, ZC = ZC_, NT = NT_;
};

template <class SS, bool PROLONG, bool RESID> struct StreamGeom {
    static constexpr int H = RESID ? 3 : 2;
    static constexpr int PX = SS::SX + 2 * H, PY = SS::SY + 2 * H, PP = PX * PY;
    static constexpr int R = RESID ? 6 : 5;                    // ring slots
    static constexpr int CX = PX / 2 + 2, CY = PY / 2 + 2, CP = CX * CY;  // staged coarse plane
    static constexpr int CR = 8;                               // coarse ring slots
    static constexpr int IL = (PP + SS::NT - 1) / SS::NT;      // ring cells per thread
    static constexpr int ICL = (CP + SS::NT - 1) / SS::NT;     // coarse cells per thread
    static constexpr int IS = (SS::SX * SS::SY + SS::NT - 1) / SS::NT;  // core cells per thread
    static constexpr int RX1 = (PX - 2) / 2, RY1 = PY - 2, C1 = RX1 * RY1, I1 = (C1 + SS::NT - 1) / SS::NT;
    static constexpr int RX2 = (PX - 4) / 2, RY2 = PY - 4, C2 = RX2 * RY2, I2 = (C2 + SS::NT - 1) / SS::NT;
    static constexpr int lds_bytes = 4 * (2 * R * PP + (PROLONG ? CR * CP : 0) + (RESID ? SS::SX * SS::SY : 0));
    static_assert(IL <= 32 && I1 <= 32 && I2 <= 32 && IS <= 32, "per-thread bit masks hold 32 cells");
};

template <class SS, bool PROLONG, bool RESID> constexpr bool stream_fits_lds()
{
    return StreamGeom<SS, PROLONG, RESID>::lds_bytes + 1024 <= 163840;
}

__device__ __forceinline__ int pmod(int a, int m) { const int r = a % m; return r < 0 ? r + m : r; }
__device__ __forceinline__ int fdiv2(int a) { return a >> 1; }  // floor(a / 2), arithmetic shift

// Cells of one colour pass (margin M) owned by this thread: base offset y PX + M + 2 xh (the cell
// is base + (y + k) % 2 on plane k, since x0, y0 and colour + M are even), y parity, and for
// either x parity whether the cell is in the column's core.
template <int N> struct PassCells {
    int cb[N];
    unsigned yb = 0u, valid = 0u, core0 = 0u, core1 = 0u;
};

template <int N, int RX, int RY, int CNT, int NT, int PX, int M, int H, int SX, int SY>
__device__ __forceinline__ void pass_cells(PassCells<N> &P, int tid)
{
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const int q = tid + i * NT;
        const int xh = q % RX, y = M + q / RX;
        P.cb[i] = y * PX + M + 2 * xh;
        if (q < CNT) P.valid |= 1u << i;
        if (y & 1) P.yb |= 1u << i;
        const bool yc = y >= H && y < H + SY;
        if (yc && M + 2 * xh >= H && M + 2 * xh < H + SX) P.core0 |= 1u << i;
        if (yc && M + 2 * xh + 1 >= H && M + 2 * xh + 1 < H + SX) P.core1 |= 1u << i;
    }
}

template <class SS, bool ZERO, bool PROLONG, bool RESID>
__device__ __forceinline__ void sweep_stream(const float *__restrict__ uin, float *__restrict__ uout,
                                             const float *__restrict__ f, int n, float h2, float h2i,
                                             const float *__restrict__ uc, float *__restrict__ rc, double *resmax, int t)
{
    using G = StreamGeom<SS, PROLONG, RESID>;
    constexpr int SX = SS::SX, SY = SS::SY, ZC = SS::ZC, NT = SS::NT;
    constexpr int H = G::H, PX = G::PX, PY = G::PY, PP = G::PP, R = G::R, CX = G::CX, CP = G::CP, CR = G::CR;
    static_assert(SX % 2 == 0 && SY % 2 == 0 && ZC % 2 == 0, "stream shape");
    static_assert(!PROLONG || H % 2 == 0, "the coarse-value window starts on an even fine plane");
    __shared__ float su[R * PP];
    __shared__ float sf[R * PP];
    __shared__ float scz[PROLONG ? CR * CP : 1];
    __shared__ float srp[RESID ? SX * SY : 1];
    const int ntx = n / SX, nty = n / SY;
    const int tx = t % ntx, ty = (t / ntx) % nty, tz = t / (ntx * nty);
    const int x0 = tx * SX, y0 = ty * SY, z0 = tz * ZC;
    const int nc = n >> 1;
    const int tid = threadIdx.x;
    const bool want = resmax != nullptr;
    float amax = 0.f;

    // ---- per-thread geometry, fixed for the whole march ----
    int goff[G::IL];         // global (y, x) offset of ring cell tid + it NT
    int coff[PROLONG ? G::IL : 1][4];  // its trilinear stencil in the staged coarse plane
#pragma unroll
    for (int it = 0; it < G::IL; ++it) {
        const int q = tid + it * NT;
        const int x = q % PX, y = q / PX;
        goff[it] = wrapn(y0 - H + y, n) * n + wrapn(x0 - H + x, n);
        if (PROLONG) {
            const int ic = (x >> 1) + 1, jc = (y >> 1) + 1;
            const int ic1 = ic + 2 * (x & 1) - 1, jc1 = jc + 2 * (y & 1) - 1;
            coff[PROLONG ? it : 0][0] = jc * CX + ic;
            coff[PROLONG ? it : 0][1] = jc * CX + ic1;
            coff[PROLONG ? it : 0][2] = jc1 * CX + ic;
            coff[PROLONG ? it : 0][3] = jc1 * CX + ic1;
        }
    }
    int cgoff[PROLONG ? G::ICL : 1];  // global (y, x) offset of coarse staged cell tid + it NT
    if (PROLONG) {
#pragma unroll
        for (int it = 0; it < G::ICL; ++it) {
            const int q = tid + it * NT;
            const int x = q % CX, y = q / CX;
            cgoff[PROLONG ? it : 0] = wrapn(fdiv2(y0 - H) - 1 + y, nc) * nc + wrapn(fdiv2(x0 - H) - 1 + x, nc);
        }
    }
    PassCells<G::I1> P1;
    PassCells<G::I2> P2;
    pass_cells<G::I1, G::RX1, G::RY1, G::C1, NT, PX, 1, H, SX, SY>(P1, tid);
    pass_cells<G::I2, G::RX2, G::RY2, G::C2, NT, PX, 2, H, SX, SY>(P2, tid);
    int soff[G::IS], sl[G::IS];  // core cells: global (y, x) offset, ring offset
#pragma unroll
    for (int it = 0; it < G::IS; ++it) {
        const int q = tid + it * NT;
        const int x = q % SX, y = q / SX;
        soff[it] = (y0 + y) * n + x0 + x;
        sl[it] = (y + H) * PX + x + H;
    }

    float ru[G::IL], rf[G::IL], rcz[PROLONG ? G::ICL : 1];
    auto fetch = [&](int k) {
        const int base = wrapn(k, n) * n * n;
#pragma unroll
        for (int it = 0; it < G::IL; ++it) {
            if (PP % NT == 0 || tid + it * NT < PP) {
                ru[it] = ZERO ? 0.0f : uin[base + goff[it]];
                rf[it] = f[base + goff[it]];
            }
        }
    };
    auto fetch_coarse = [&](int c) {
        const int base = wrapn(c, nc) * nc * nc;
#pragma unroll
        for (int it = 0; it < G::ICL; ++it)
            if (CP % NT == 0 || tid + it * NT < CP) rcz[PROLONG ? it : 0] = uc[base + cgoff[PROLONG ? it : 0]];
    };
    auto put_coarse = [&](int c) {
        float *dst = scz + (c & (CR - 1)) * CP;
#pragma unroll
        for (int it = 0; it < G::ICL; ++it)
            if (CP % NT == 0 || tid + it * NT < CP) dst[tid + it * NT] = rcz[PROLONG ? it : 0];
    };
    // PROLONG: the four coarse values of each ring cell's trilinear stencil on the two coarse
    // planes a put reads, held in registers: win[0] plane m, win[1] plane m + 1 for the fine
    // planes 2m + 1 and 2m + 2, so each coarse plane is read from LDS once per thread, not 4 times
    constexpr int WI = PROLONG ? G::IL : 1;
    float win[2][WI][4];
    auto win_load = [&](int w, int c) {
        const float *src = scz + (c & (CR - 1)) * CP;
#pragma unroll
        for (int it = 0; it < WI; ++it)
            if (PP % NT == 0 || tid + it * NT < PP)
#pragma unroll
                for (int e = 0; e < 4; ++e) win[w][it][e] = src[coff[PROLONG ? it : 0][e]];
    };
    // the fetched plane k into ring slot `slot` (adding the trilinear correction with PROLONG:
    // coarse planes kc = floor(k / 2) and kc1 = kc - 1 (k even) or kc + 1 (k odd))
    auto put = [&](int k, int slot) {
        float *du = su + slot * PP, *df = sf + slot * PP;
        const bool odd = (k & 1) != 0;
        if (PROLONG && odd) {
#pragma unroll
            for (int it = 0; it < WI; ++it)
#pragma unroll
                for (int e = 0; e < 4; ++e) win[0][it][e] = win[1][it][e];
            win_load(1, fdiv2(k) + 1);
        }
#pragma unroll
        for (int it = 0; it < G::IL; ++it) {
            const int q = tid + it * NT;
            if (PP % NT == 0 || q < PP) {
                float v = ru[it];
                if (PROLONG) {
                    const int w = PROLONG ? it : 0;
                    float a[4], b[4];  // plane kc, plane kc1
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        a[e] = odd ? win[0][w][e] : win[1][w][e];
                        b[e] = odd ? win[1][w][e] : win[0][w][e];
                    }
                    v = v + trilinear_value<float>(a[0], a[1], a[2], b[0], b[2], b[1], a[3], b[3]);
                }
                du[q] = v;
                df[q] = rf[it];
            }
        }
    };
    // one colour pass over the cells P on plane k: u (slot of k), um / up (k - 1 / k + 1), fk
    auto colour = [&](auto &P, int k, float *u, const float *um, const float *up, const float *fk, bool out) {
        constexpr int N = sizeof(P.cb) / sizeof(int);
        const unsigned kb = (unsigned)(k & 1);
#pragma unroll
        for (int i = 0; i < N; ++i) {
            if (((P.valid >> i) & 1u) == 0u) continue;
            const unsigned d = ((P.yb >> i) & 1u) ^ kb;
            const int c = P.cb[i] + (int)d;
            float s;
            u[c] = relax_value<float>(u[c], u[c + 1], u[c - 1], u[c + PX], u[c - PX], up[c], um[c], fk[c], h2, &s);
            if (want && out && ((((d ? P.core1 : P.core0) >> i) & 1u) != 0u)) amax = fmaxf(amax, fabsf(s));
        }
    };
    auto resid_at = [&](int slot, int slotm, int slotp, int x, int y) {
        const float *u = su + slot * PP, *um = su + slotm * PP, *up = su + slotp * PP;
        const float *fk = sf + slot * PP;
        const int c = (y + H) * PX + x + H;
        return resid_value<float>(u[c], u[c + 1], u[c - 1], u[c + PX], u[c - PX], up[c], um[c], fk[c], h2i);
    };

    const int zlo = z0 - H, zhi = z0 + ZC + H - 1;  // planes held at some point
    // ---- prologue: coarse planes for fine planes zlo .. zlo + 3, fine planes zlo .. zlo + 2 into
    // slots 0 .. 2, plane zlo + 3 (and the next coarse plane) into registers ----
    auto cneed = [&](int k) { return fdiv2(k) + 1; };  // highest coarse plane put(k) reads
    int cmax = 0;
    if (PROLONG) {
        for (int c = fdiv2(zlo) - 1; c <= cneed(zlo + 3); ++c) {
            fetch_coarse(c);
            put_coarse(c);
        }
        cmax = cneed(zlo + 3);
        __syncthreads();
    }
    if (PROLONG) {
        // zlo is even (H = 2, z0 a multiple of the even ZC): put(zlo) reads planes zlo/2 - 1, zlo/2
        win_load(0, fdiv2(zlo) - 1);
        win_load(1, fdiv2(zlo));
    }
    for (int k = zlo; k <= zlo + 2; ++k) {
        fetch(k);
        put(k, k - zlo);
    }
    fetch(zlo + 3);
    bool cpend = false;  // a coarse plane (cmax + 1) waits in the registers
    if (PROLONG && cneed(zlo + 4) > cmax) {
        fetch_coarse(cmax + 1);
        cpend = true;
    }
    __syncthreads();
    for (int pb = zlo + 1; pb <= zhi - 1; pb += R) {
#pragma unroll
        for (int uu = 0; uu < R; ++uu) {
            const int pa = pb + uu;
            if (pa > zhi - 1) break;  // uniform per workgroup
            // ring slot of plane pa + delta: (pa + delta - zlo) mod R, pa - zlo = 1 + uu (mod R)
            auto S = [&](int delta) { return ((1 + uu + delta) % R + R) % R; };
            // 1. first colour of plane pa (margin 1)
            colour(P1, pa, su + S(0) * PP, su + S(-1) * PP, su + S(1) * PP, sf + S(0) * PP, pa >= z0 && pa < z0 + ZC);
            __syncthreads();
            // 2. coarse plane for the next step's put, plane pa + 2 into its slot, fetches
            if (PROLONG && cpend) {
                put_coarse(cmax + 1);
                ++cmax;
                cpend = false;
            }
            if (pa + 2 <= zhi) put(pa + 2, S(2));
            if (pa + 3 <= zhi) fetch(pa + 3);
            if (PROLONG && pa + 3 <= zhi && cneed(pa + 4 <= zhi ? pa + 4 : zhi) > cmax) {
                fetch_coarse(cmax + 1);
                cpend = true;
            }
            //    second colour of plane pa - 1 (margin 2): final
            if (pa - 1 >= zlo + 2 && pa - 1 <= zhi - 2)
                colour(P2, pa - 1, su + S(-1) * PP, su + S(-2) * PP, su + S(0) * PP, sf + S(-1) * PP,
                       pa - 1 >= z0 && pa - 1 < z0 + ZC);
            __syncthreads();
            // 3. store plane pa - 1; residual of plane pa - 2
            if (pa - 1 >= z0 && pa - 1 < z0 + ZC) {
                const float *u = su + S(-1) * PP;
                const int gz = wrapn(pa - 1, n) * n * n;
#pragma unroll
                for (int it = 0; it < G::IS; ++it)
                    if ((SX * SY) % NT == 0 || tid + it * NT < SX * SY) uout[gz + soff[it]] = u[sl[it]];
            }
            if (RESID) {
                const int kq = pa - 2;
                if (kq >= z0 && kq < z0 + ZC) {
                    if ((kq & 1) == 0) {
                        for (int q = tid; q < SX * SY; q += NT) srp[q] = resid_at(S(-2), S(-3), S(-1), q % SX, q / SX);
                    } else {
                        // coarse cell (cx, cy) of the pair (kq - 1, kq), shared by two adjacent lanes:
                        // lane b forms the two residuals of row 2 cy + b on the odd plane and takes
                        // that row's two of the even plane from srp; the pair swaps its four values
                        // and lane 0 restricts the eight (NT is a multiple of the wave size, so both
                        // lanes of a pair are in one wave and in the same loop trip)
                        for (int q = tid; q < SX * SY / 2; q += NT) {
                            const int b = q & 1, cx = (q >> 1) % (SX / 2), cy = (q >> 1) / (SX / 2);
                            const int y = 2 * cy + b;
                            const float e0 = srp[y * SX + 2 * cx], e1 = srp[y * SX + 2 * cx + 1];
                            const float o0 = resid_at(S(-2), S(-3), S(-1), 2 * cx, y);
                            const float o1 = resid_at(S(-2), S(-3), S(-1), 2 * cx + 1, y);
                            const float pe0 = __shfl_xor(e0, 1), pe1 = __shfl_xor(e1, 1);
                            const float po0 = __shfl_xor(o0, 1), po1 = __shfl_xor(o1, 1);
                            if (b == 0) {
                                // r[a + 2 b + 4 dz]: own row is b = 0, the partner's b = 1
                                const float r[8] = {e0, e1, pe0, pe1, o0, o1, po0, po1};
                                const int gcx = x0 / 2 + cx, gcy = y0 / 2 + cy, gcz = wrapn(kq, n) / 2;
                                rc[(gcz * nc + gcy) * nc + gcx] =
                                    restrict_value<float>(r[0], r[1], r[2], r[4], r[3], r[5], r[6], r[7]);
                            }
                        }
                    }
                }
            }
        }
    }
    if (want) block_atomic_max(resmax, (double)amax);
}

template <class SS, bool ZERO, bool PROLONG, bool RESID>
__global__ void __launch_bounds__(SS::NT)
k_sweep_stream(const float *__restrict__ uin, float *__restrict__ uout, const float *__restrict__ f, int n, float h2,
               float h2i, const float *__restrict__ uc, float *__restrict__ rc, double *resmax)
{
    const int ntiles = (n / SS::SX) * (n / SS::SY) * (n / SS::ZC);
    const int t = xcd_tile(ntiles);
    if (t >= ntiles) return;
    sweep_stream<SS, ZERO, PROLONG, RESID>(uin, uout, f, n, h2, h2i, uc, rc, resmax, t);
}

// Streaming shapes the autotuners choose from beside the tiles (one sweep per launch, source always
// staged): column SX x SY, chunk of ZC planes, NT threads.
using Stream0 = StreamShape<64, 16, 32, 512>;
using Stream1 = StreamShape<64, 16, 16, 512>;
using Stream2 = StreamShape<64, 32, 32, 512>;
using Stream3 = StreamShape<32, 16, 32, 256>;
using Stream4 = StreamShape<128, 16, 32, 512>;
using Stream5 = StreamShape<32, 16, 16, 256>;
using Stream6 = StreamShape<32, 8, 16, 256>;
using Stream7 = StreamShape<32, 32, 32, 512>;
using Stream8 = StreamShape<16, 16, 16, 256>;
constexpr int kNumStream = 9;
static const int kStreamDims[kNumStream][4] = {
    {64, 16, 32, 512}, {64, 16, 16, 512}, {64, 32, 32, 512}, {32, 16, 32, 256}, {128, 16, 32, 512},
    {32, 16, 16, 256}, {32, 8, 16, 256},  {32, 32, 32, 512}, {16, 16, 16, 256}};
static_assert(stream_fits_lds<Stream2, true, false>() && stream_fits_lds<Stream2, false, true>(), "stream LDS");
static_assert(stream_fits_lds<Stream4, true, false>() && stream_fits_lds<Stream4, false, true>(), "stream LDS");

// sweep configurations: 0 .. kNumTiles-1 tiles, kNumTiles .. kNumTiles+kNumStream-1 streams
constexpr int kNumCfg = kNumTiles + kNumStream;
constexpr int kNumCand = kNumTiles * kMaxSweepsPerLaunch * 2 + kNumStream;
static_assert(kNumCand <= 32, "autotuner timing tables hold 32 candidates");

inline bool cfg_fits(int cfg, int n)
{
    if (cfg < kNumTiles) return tile_fits(cfg, n);
    const int *d = kStreamDims[cfg - kNumTiles];
    return n >= d[0] && n >= d[1] && n >= d[2] && n % d[0] == 0 && n % d[1] == 0 && n % d[2] == 0;
}

// whether configuration cfg with (ns, fs) exists: streams take one sweep per launch and always
// stage the source (reported as fs = 0)
inline bool cfg_ok(int cfg, int ns, bool fs)
{
    if (cfg < kNumTiles) return cand_ok(cfg, ns, fs);
    return cfg < kNumCfg && ns == 1 && !fs;
}

// autotuner candidate index -> (cfg, ns, fs)
inline void cand_decode(int cand, int &cfg, int &ns, bool &fs)
{
    constexpr int nt = kNumTiles * kMaxSweepsPerLaunch * 2;
    if (cand < nt) {
        cfg = cand / (2 * kMaxSweepsPerLaunch);
        ns = 1 + (cand / 2) % kMaxSweepsPerLaunch;
        fs = cand % 2;
    } else {
        cfg = kNumTiles + (cand - nt);
        ns = 1;
        fs = false;
    }
}

// FMG tricubic prolongation uf = P3 uc (interpolation.F90:98-217) on an LDS tile: the coarse cells
// any stencil of the fine tile can touch are staged, then the x, y and z passes run on LDS exactly
// as the reference's bufx / bufxy sweeps.  Fine tile FX*FY*FZ, nf >= FX, nf % FY == nf % FZ == 0.
// T = float (mixed path) or double (plain path): the weights are rounded to T as k_interp_cubic
// rounds them, and every value is formed in the same operand order, so both give its results.
template <int FX, int FY, int FZ, int NT, typename T = float>
__device__ __forceinline__ void interp_cubic_tile(T *__restrict__ uf, const T *__restrict__ uc, int nf, int t)
{
    constexpr int CX = FX / 2 + 4, CY = FY / 2 + 4, CZ = FZ / 2 + 4;
    __shared__ T sc[CX * CY * CZ];
    __shared__ T bx[CZ * CY * FX];
    __shared__ T bxy[CZ * FY * FX];
    const int nc = nf >> 1;
    const int ntx = nf / FX, nty = nf / FY;
    const int tx = t % ntx, ty = (t / ntx) % nty, tz = t / (ntx * nty);
    const int fx0 = tx * FX, fy0 = ty * FY, fz0 = tz * FZ;
    const int bxo = fx0 / 2 - 2, byo = fy0 / 2 - 2, bzo = fz0 / 2 - 2;  // ghosted coarse index of local 0
    {
        constexpr int CNT = CX * CY * CZ, IT = (CNT + NT - 1) / NT;
        T v[IT];
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int q = threadIdx.x + it * NT;
            if (CNT % NT == 0 || q < CNT) {
                const int x = q % CX, y = (q / CX) % CY, z = q / (CX * CY);
                v[it] = uc[(wrapn(bzo + z, nc) * nc + wrapn(byo + y, nc)) * nc + wrapn(bxo + x, nc)];
            }
        }
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int q = threadIdx.x + it * NT;
            if (CNT % NT == 0 || q < CNT) sc[q] = v[it];
        }
    }
    // per-axis stencil tables (interpolation.F90:106-127), computed once per tile
    __shared__ int tix[FX][4], tiy[FY][4], tiz[FZ][4];
    __shared__ T twx[FX][4], twy[FY][4], twz[FZ][4];
    if (threadIdx.x < FX + FY + FZ) {
        const int q = threadIdx.x;
        int idx[4];
        double a[4];
        if (q < FX) {
            cubic_stencil(fx0 + q, -1, nc, idx, a);
            for (int e = 0; e < 4; ++e) { tix[q][e] = idx[e] - bxo; twx[q][e] = (T)a[e]; }
        } else if (q < FX + FY) {
            cubic_stencil(fy0 + q - FX, -1, nc, idx, a);
            for (int e = 0; e < 4; ++e) { tiy[q - FX][e] = idx[e] - byo; twy[q - FX][e] = (T)a[e]; }
        } else {
            cubic_stencil(fz0 + q - FX - FY, -1, nc, idx, a);
            for (int e = 0; e < 4; ++e) { tiz[q - FX - FY][e] = idx[e] - bzo; twz[q - FX - FY][e] = (T)a[e]; }
        }
    }
    __syncthreads();
    // x pass: bx(fx, jc, kc) over the staged coarse y, z range
    for (int q = threadIdx.x; q < CZ * CY * FX; q += NT) {
        const int x = q % FX, r = q / FX;
        const T *row = sc + r * CX;
        bx[q] = cubic4<T>(twx[x], row[tix[x][0]], row[tix[x][1]], row[tix[x][2]], row[tix[x][3]]);
    }
    __syncthreads();
    // y pass: bxy(fx, fy, kc)
    for (int q = threadIdx.x; q < CZ * FY * FX; q += NT) {
        const int x = q % FX, y = (q / FX) % FY, z = q / (FX * FY);
        const T *col = bx + z * CY * FX + x;
        bxy[q] = cubic4<T>(twy[y], col[tiy[y][0] * FX], col[tiy[y][1] * FX], col[tiy[y][2] * FX],
                               col[tiy[y][3] * FX]);
    }
    __syncthreads();
    // z pass: uf(fx, fy, fz)
    for (int q = threadIdx.x; q < FZ * FY * FX; q += NT) {
        const int x = q % FX, y = (q / FX) % FY, z = q / (FX * FY);
        const T *p = bxy + y * FX + x;
        constexpr int PL = FY * FX;
        uf[((fz0 + z) * nf + (fy0 + y)) * nf + (fx0 + x)] =
            cubic4<T>(twz[z], p[tiz[z][0] * PL], p[tiz[z][1] * PL], p[tiz[z][2] * PL], p[tiz[z][3] * PL]);
    }
}

template <int FX, int FY, int FZ, int NT, typename T = float>
__global__ void __launch_bounds__(NT) k_interp_cubic_tile(T *__restrict__ uf, const T *__restrict__ uc, int nf)
{
    const int ntiles = (nf / FX) * (nf / FY) * (nf / FZ);
    const int t = xcd_tile(ntiles);
    if (t >= ntiles) return;
    interp_cubic_tile<FX, FY, FZ, NT, T>(uf, uc, nf, t);
}

// ------------------------------------------------------------------------------------------------
// Single-workgroup solver for the levels of <= 16^3 cells, all resident in LDS.
// ------------------------------------------------------------------------------------------------
struct CoarseLevels {
    int nlev;               // levels held: 0 .. nlev-1 (local numbering), n_0 <= 16
    int n[8], lg[8];        // cells per side and its log2
    float h2[8], h2i[8];
    int off[8];             // offset of level l inside each LDS array
    int exact;              // 1: held level 0 solved exactly instead (coarse_exact)
    float Q[kCoarseMaxN * kCoarseMaxN], lam[kCoarseMaxN];  // its Hartley basis and eigenvalues
};

constexpr int kCoarseCells = 16 * 16 * 16 + 8 * 8 * 8 + 4 * 4 * 4 + 2 * 2 * 2 + 1;

// Hartley basis of the periodic 3-point second difference on n cells, Q[m n + k] =
// (cos a + sin a) / sqrt(n) with a = 2 pi ((m k) mod n) / n, and its eigenvalues
// lam[m] = (2 cos(2 pi m / n) - 2) / h^2 (host; the CPU restatement forms them the same way).
inline void hartley_basis(int n, double h, double *Q, double *lam)
{
    const double pi = 3.14159265358979323846;
    for (int m = 0; m < n; ++m) {
        for (int k = 0; k < n; ++k) {
            const double a = 2.0 * pi * (double)((m * k) % n) / (double)n;
            Q[m * n + k] = (std::cos(a) + std::sin(a)) / std::sqrt((double)n);
        }
        lam[m] = (2.0 * std::cos(2.0 * pi * (double)m / (double)n) - 2.0) / (h * h);
    }
}

// ---- exact solve of a periodic level of n^3 cells (n <= 16) in one workgroup ---------------------
// The coarse-grid problem L u = f that the V-cycle bottom and the FMG start approximate (mg.F90:
// 142-210 below the first fine level) is solved exactly instead: the 7-point operator separates,
// the symmetric orthonormal Hartley basis Q diagonalises the periodic second difference along
// each axis, so u = Q (Lambda^-1 (Q f)) with Lambda = (lam_i + lam_j) + lam_k and the zero mode set
// to zero (the zero-mean solution).  Three axis passes serve as forward and inverse transform.
// The grid sits in LDS with row pitch n + 1.  Thread t works on grid line t / 4 and owns the
// outputs m = t % 4 + 4 r (r < 4) of every line it is given: its four rows of Q stay in registers
// for all six passes, and the four threads of a line sit in adjacent lanes, so the line's reads
// are broadcasts.  Each output is accumulated k = 0 .. n-1 in order with fused multiply-adds.
constexpr int kCeParts = 4;
static_assert(kCoarseMaxN * kCoarseMaxN * kCeParts <= kCoarseThreads, "one thread per (line, part)");
static_assert(kCoarseMaxN <= 4 * kCeParts, "four outputs per thread");

__device__ inline void ce_basis(float (&q)[4][kCoarseMaxN], const float *Q, int n)
{
    const int part = threadIdx.x % kCeParts;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int m = part + kCeParts * r;
#pragma unroll
        for (int k = 0; k < kCoarseMaxN; ++k) q[r][k] = (m < n && k < n) ? Q[m * n + k] : 0.f;
    }
}

__device__ inline void ce_pass(float *g, const float (&q)[4][kCoarseMaxN], int n, int axis)
{
    const int P = n + 1;
    const int line = threadIdx.x / kCeParts, part = threadIdx.x % kCeParts;
    const int a = line % n, b = line / n;
    int base, stride;
    if (axis == 0) { base = (b * n + a) * P; stride = 1; }        // line (j, k) = (a, b)
    else if (axis == 1) { base = b * n * P + a; stride = P; }     // line (i, k)
    else { base = b * P + a; stride = n * P; }                    // line (i, j)
    const bool act = line < n * n;
    float x[kCoarseMaxN];
#pragma unroll
    for (int k = 0; k < kCoarseMaxN; ++k) x[k] = (act && k < n) ? g[base + k * stride] : 0.f;
    __syncthreads();
    if (act) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int m = part + kCeParts * r;
            if (m < n) {
                float acc = 0.f;
#pragma unroll
                for (int k = 0; k < kCoarseMaxN; ++k)
                    if (k < n) acc = __builtin_fmaf(q[r][k], x[k], acc);
                g[base + m * stride] = acc;
            }
        }
    }
    __syncthreads();
}

// u_out := exact solution for rhs_in (dense n^3 arrays, i fastest).  g: n^2 (n + 1) floats of LDS,
// s: n^2 + n floats of LDS for the basis (staged there once, then read into registers as
// broadcasts) and the eigenvalues; Qg, lamg: basis and eigenvalues (any address space).
__device__ inline void coarse_exact(float *g, float *s, const float *Qg, const float *lamg, const float *rhs_in,
                                    float *u_out, int n)
{
    const int P = n + 1, n3 = n * n * n;
    float *Qs = s, *lam = s + n * n;
    for (int t = threadIdx.x; t < n * n; t += kCoarseThreads) Qs[t] = Qg[t];
    for (int t = threadIdx.x; t < n; t += kCoarseThreads) lam[t] = lamg[t];
    for (int c = threadIdx.x; c < n3; c += kCoarseThreads) {
        const int i = c % n, j = (c / n) % n, k = c / (n * n);
        g[(k * n + j) * P + i] = rhs_in[c];
    }
    __syncthreads();
    float q[4][kCoarseMaxN];
    ce_basis(q, Qs, n);
    for (int a = 0; a < 3; ++a) ce_pass(g, q, n, a);
    for (int c = threadIdx.x; c < n3; c += kCoarseThreads) {
        const int i = c % n, j = (c / n) % n, k = c / (n * n);
        const float den = (lam[i] + lam[j]) + lam[k];
        float &v = g[(k * n + j) * P + i];
        v = den != 0.f ? v / den : 0.f;
    }
    __syncthreads();
    for (int a = 0; a < 3; ++a) ce_pass(g, q, n, a);
    for (int c = threadIdx.x; c < n3; c += kCoarseThreads) {
        const int i = c % n, j = (c / n) % n, k = c / (n * n);
        u_out[c] = g[(k * n + j) * P + i];
    }
}

__device__ __forceinline__ void c_relax(float *U, const float *R, int n, int lg, float h2, double *resmax)
{
    double amax = 0.0;
    const int lgh = lg > 0 ? lg - 1 : 0;  // log2 of cells per colour row (n >= 2), 1 for n == 1
    const int m = n - 1;
    for (int ipass = 1; ipass <= 2; ++ipass) {
        for (int t = threadIdx.x; t < (n * n * n + 1) / 2; t += kCoarseThreads) {
            const int ih = t & ((1 << lgh) - 1), j = (t >> lgh) & m, k = t >> (lgh + lg);
            const int i = (n == 1) ? 0 : 2 * ih + ((j + k + ipass) & 1);
            if (n == 1 && ipass == 1) continue;   // single cell has colour "even" (ipass 2)
            const int c = (k * n + j) * n + i, row = (k * n + j) * n;
            float s;
            U[c] = relax_value<float>(U[c], U[row + ((i + 1) & m)], U[row + ((i - 1) & m)],
                                      U[(k * n + ((j + 1) & m)) * n + i], U[(k * n + ((j - 1) & m)) * n + i],
                                      U[(((k + 1) & m) * n + j) * n + i], U[(((k - 1) & m) * n + j) * n + i], R[c],
                                      h2, &s);
            amax = fmax(amax, fabs((double)s));
        }
        __syncthreads();
    }
    if (resmax) block_atomic_max(resmax, amax);
}

// residual of level (n, U, R) restricted into Rc: each coarse cell evaluates its 8 fine residuals
// (mg_poisson.F90:17-27 then restriction.F90:132-147, same values as the two separate passes).
__device__ __forceinline__ void c_resid_restrict(const float *U, const float *R, int n, int lg, float h2i, float *Rc)
{
    const int nc = n >> 1, lgc = lg - 1, m = n - 1;
    for (int q = threadIdx.x; q < nc * nc * nc; q += kCoarseThreads) {
        const int cx = q & (nc - 1), cy = (q >> lgc) & (nc - 1), cz = q >> (2 * lgc);
        float r[8];
#pragma unroll
        for (int o = 0; o < 8; ++o) {
            const int i = 2 * cx + (o & 1), j = 2 * cy + ((o >> 1) & 1), k = 2 * cz + ((o >> 2) & 1);
            const int c = (k * n + j) * n + i, row = (k * n + j) * n;
            r[o] = resid_value<float>(U[c], U[row + ((i + 1) & m)], U[row + ((i - 1) & m)],
                                      U[(k * n + ((j + 1) & m)) * n + i], U[(k * n + ((j - 1) & m)) * n + i],
                                      U[(((k + 1) & m) * n + j) * n + i], U[(((k - 1) & m) * n + j) * n + i], R[c],
                                      h2i);
        }
        Rc[q] = restrict_value<float>(r[0], r[1], r[2], r[4], r[3], r[5], r[6], r[7]);
    }
    __syncthreads();
}

__device__ __forceinline__ void c_restrict(float *Rc, const float *Rf, int nc, int lgc)
{
    const int nf = 2 * nc;
    for (int q = threadIdx.x; q < nc * nc * nc; q += kCoarseThreads) {
        const int i = 2 * (q & (nc - 1)), j = 2 * ((q >> lgc) & (nc - 1)), k = 2 * (q >> (2 * lgc));
        auto F = [&](int a, int b, int d) { return Rf[((k + d) * nf + (j + b)) * nf + (i + a)]; };
        Rc[q] = restrict_value<float>(F(0, 0, 0), F(1, 0, 0), F(0, 1, 0), F(0, 0, 1), F(1, 1, 0), F(1, 0, 1),
                                      F(0, 1, 1), F(1, 1, 1));
    }
    __syncthreads();
}

__device__ __forceinline__ void c_fill0(float *U, int n)
{
    for (int c = threadIdx.x; c < n * n * n; c += kCoarseThreads) U[c] = 0.0f;
    __syncthreads();
}

// mg_interp (mg.F90:608-633): cubic (FMG) unless the coarse grid plus ghosts is narrower than 4.
__device__ __forceinline__ void c_interp(float *Uf, const float *Uc, int nf, int lgf, bool cubic, bool add)
{
    const int nc = nf >> 1;
    auto C = [&](int a, int b, int d) { return Uc[(d * nc + b) * nc + a]; };
    for (int c = threadIdx.x; c < nf * nf * nf; c += kCoarseThreads) {
        const int i = c & (nf - 1), j = (c >> lgf) & (nf - 1), k = c >> (2 * lgf);
        float v;
        if (cubic && nc + 2 >= 4) v = tricubic_point<float>(C, i, j, k, nc);
        else v = trilinear_point<float>(C, i, j, k, nc);
        Uf[c] = add ? Uf[c] + v : v;
    }
    __syncthreads();
}

// mode 0: FMG over all held levels from RHO of local level 0 (mg.F90:142-215 restricted to them);
// mode 1: one V-cycle from local level 0 with zero initial guess (mg.F90:179-210).
// resmax receives max|resh2| of the final smoothing call when non-null (only meaningful when the
// held level 0 is the finest level of the problem).  With L.exact both modes solve held level 0
// exactly (coarse_exact).
static __global__ void __launch_bounds__(kCoarseThreads)
k_coarse(const float *__restrict__ rhs_in, float *__restrict__ u_out, CoarseLevels L, int mode, int ncycle,
         int npre, int npost, double *resmax)
{
    __shared__ float sU[kCoarseCells], sR[kCoarseCells];
    static_assert(kCoarseMaxN * kCoarseMaxN * (kCoarseMaxN + 1) <= kCoarseCells &&
                      kCoarseMaxN * kCoarseMaxN + kCoarseMaxN <= kCoarseCells,
                  "exact coarse solve exceeds the coarse kernel's LDS");
    const int lmax = L.nlev - 1;
    const int n0 = L.n[0];
    if (L.exact) {
        coarse_exact(sU, sR, L.Q, L.lam, rhs_in, u_out, n0);
        return;
    }
    for (int c = threadIdx.x; c < n0 * n0 * n0; c += kCoarseThreads) sR[c] = rhs_in[c];
    __syncthreads();
    auto U = [&](int l) { return sU + L.off[l]; };
    auto R = [&](int l) { return sR + L.off[l]; };
    auto vcycle = [&](int ml, bool last_cycle) {
        for (int v = ml; v <= lmax - 1; ++v) {
            for (int p = 0; p < npre; ++p) c_relax(U(v), R(v), L.n[v], L.lg[v], L.h2[v], nullptr);
            c_resid_restrict(U(v), R(v), L.n[v], L.lg[v], L.h2i[v], R(v + 1));
            c_fill0(U(v + 1), L.n[v + 1]);
        }
        for (int it = 0; it < 2; ++it) c_relax(U(lmax), R(lmax), L.n[lmax], L.lg[lmax], L.h2[lmax], nullptr);
        for (int v = lmax - 1; v >= ml; --v) {
            c_interp(U(v), U(v + 1), L.n[v], L.lg[v], false, true);
            for (int p = 0; p < npost; ++p) {
                const bool last = last_cycle && v == 0 && p == npost - 1;
                c_relax(U(v), R(v), L.n[v], L.lg[v], L.h2[v], last ? resmax : nullptr);
            }
        }
    };
    if (mode == 0) {
        for (int l = 1; l <= lmax; ++l) c_restrict(R(l), R(l - 1), L.n[l], L.lg[l]);
        c_fill0(U(lmax), L.n[lmax]);
        for (int it = 0; it < 2; ++it) c_relax(U(lmax), R(lmax), L.n[lmax], L.lg[lmax], L.h2[lmax], nullptr);
        for (int ml = lmax - 1; ml >= 0; --ml) {
            c_interp(U(ml), U(ml + 1), L.n[ml], L.lg[ml], true, false);
            for (int jc = 0; jc < ncycle; ++jc) vcycle(ml, ml == 0 && jc == ncycle - 1);
        }
    } else {
        c_fill0(U(0), n0);
        if (lmax == 0) {
            for (int it = 0; it < 2; ++it) c_relax(U(0), R(0), n0, L.lg[0], L.h2[0], nullptr);
        } else {
            vcycle(0, false);
        }
    }
    for (int c = threadIdx.x; c < n0 * n0 * n0; c += kCoarseThreads) u_out[c] = sU[c];
}

} // namespace sg
