// Uniform periodic-grid multigrid on one MI355X: host orchestration of mg_cycle / mg_lin
// (mg.F90:42-218) over device-resident level hierarchies, and the uniform-grid part of the C ABI
// declared in sg.h.
//
// Two inner solvers share the same kernels:
//  * SG_PREC_FP64 ("plain"): every level array FP64, one launch per reference loop nest,
//    launched eagerly.  Bitwise equal to the CPU FP64 restatement (same operand order, no FMA
//    contraction).
//  * SG_PREC_MIXED: the inner mg_lin runs on FP32 level arrays; the outer residual
//    r = f - L phi and the update phi += e stay FP64.  One outer iteration (inner solve, update,
//    next outer residual and its norms) is captured once into a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include "exact_planes.h"
#include "fast_kernels.h"
#include "sg_common.h"
#include "uniform_kernels.h"

namespace sg {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

const char *last_error() { return g_err; }

constexpr int kMaxLev = 16;
constexpr int kBlock = 256;
constexpr int kMaxN = 1024;
constexpr int kNormBlocks = 2048;  // grid of the norm-reducing kernels (fixed: deterministic sums)

template <typename T> struct Levels {
    T *U[kMaxLev] = {};
    T *RES[kMaxLev] = {};
    T *RHO[kMaxLev] = {};  // RHO[l] also serves as RHS[l]: see mg_lin below
};

static inline unsigned nblocks(long work) { return (unsigned)((work + kBlock - 1) / kBlock); }
static long n3(int n) { return (long)n * n * n; }
static int ilog2(int n)
{
    int lg = 0;
    while ((1 << lg) < n) ++lg;
    return lg;
}

} // namespace sg

using namespace sg;

struct sg_uniform {
    int n = 0, nlev = 0, precisions = 0;
    double h = 0.0;
    int nl[kMaxLev] = {};
    double hl[kMaxLev] = {};
    Levels<double> d;
    Levels<float> f;
    float *U1[kMaxLev] = {};   // second FP32 U buffer per level: the fused sweeps ping-pong
    int tile_cfg[kMaxLev] = {}; // autotuned sweep tile shape per fine level
    int tile_ns[kMaxLev] = {};  // autotuned sweeps per launch per fine level
    int tile_fs[kMaxLev] = {};  // autotuned: source staged in LDS
    float tune_ms[kMaxLev][32] = {};
    CoarseLevels coarse;       // levels held by k_coarse, with the exact-solve basis of the first
    int lc = 0;                // first level of <= 16^3 cells
    int lb = 0;                // bottom level of the exact-coarse cycle (lc, or a level of <= 64^3)
    float *d_qx = nullptr, *d_invx = nullptr;  // Hartley basis of level lb, 1 / Lambda (planes)
    double *psi_dev = nullptr, *src_dev = nullptr;   // staging for the host-pointer entry point
    double *adj_src = nullptr, *adj_part = nullptr;  // adjoint: projected source, mean partials
    double *psi_alt = nullptr;                         // second phi buffer of the fused outer step
    double *pmax = nullptr, *psum = nullptr;          // norm partials
    int npart = 0;
    int last_odd = 0;          // the previous fused mixed solve took an odd number of iterations
    double *dnorm = nullptr;   // [0] max|r| [1] sum r^2 [2] resmaxg*h^2 [3] max|f| [4] sum f^2
    double *hnorm = nullptr;   // pinned host mirror of dnorm
    double *hnorm_dev = nullptr;  // its device-side address (hipHostGetDevicePointer)
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evn = nullptr;
    // graphs of one mixed-precision outer iteration (one per phi buffer parity), keyed on the
    // arguments they were captured with
    struct Graph {
        hipGraphExec_t exec = nullptr;
        const double *pin = nullptr;
        double *pout = nullptr;
        const double *src = nullptr;
        int ncycle = -1, npre = -1, npost = -1, fused = -1, stop = -1, coarse = -1;
    } graph[2];
};

namespace sg {

template <typename T> static Levels<T> &levels(sg_uniform *c);
template <> Levels<double> &levels<double>(sg_uniform *c) { return c->d; }
template <> Levels<float> &levels<float>(sg_uniform *c) { return c->f; }

template <typename T> static int launch_fill(hipStream_t s, T *p, long len)
{
    hipLaunchKernelGGL(k_fill<T>, dim3(nblocks(len)), dim3(kBlock), 0, s, p, len, T(0));
    SG_CHECK_LAUNCH();
    return SG_OK;
}

template <typename T>
static int launch_relax(hipStream_t s, T *u, const T *rhs, int n, double h, double *resmax)
{
    const long work = (long)((n + 1) / 2) * n * n;
    const T h2 = (T)(h * h);
    for (int ipass = 1; ipass <= 2; ++ipass) {
        hipLaunchKernelGGL(k_relax_color<T>, dim3(nblocks(work)), dim3(kBlock), 0, s, u, rhs, ilog2(n), h2, ipass,
                           resmax);
        SG_CHECK_LAUNCH();
    }
    return SG_OK;
}

// norms != nullptr: also reduce max|res| and sum res^2 into norms[0..1] (deterministic order).
template <typename Tin, typename Tout>
static int launch_resid(sg_uniform *c, Tout *res, const Tin *u, const Tin *rhs, int n, double h, double *norms)
{
    const long work = n3(n);
    const Tin h2i = (Tin)(1.0 / (h * h));
    const unsigned nb = std::min<unsigned>(nblocks(work), (unsigned)kNormBlocks);
    if (norms && (int)nb > c->npart) {
        set_error("norm partial buffer too small (%u > %d)", nb, c->npart);
        return SG_ERR_ARG;
    }
    hipLaunchKernelGGL((k_resid<Tin, Tout>), dim3(nb), dim3(kBlock), 0, c->stream, res, u, rhs, ilog2(n), h2i,
                       norms ? c->pmax : nullptr, norms ? c->psum : nullptr);
    SG_CHECK_LAUNCH();
    if (norms) {
        hipLaunchKernelGGL(k_norm_finalize, dim3(1), dim3(1024), 0, c->stream, c->pmax, c->psum, (int)nb, norms);
        SG_CHECK_LAUNCH();
    }
    return SG_OK;
}

// max|a| and sum a^2 of a dense FP64 array (the source norms of the relative residual).
__global__ void __launch_bounds__(256) k_norm_partial(const double *__restrict__ a, long len, double *pmax,
                                                      double *psum)
{
    double mx = 0.0, sm = 0.0;
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < len; c += (long)gridDim.x * blockDim.x) {
        mx = fmax(mx, fabs(a[c]));
        sm += a[c] * a[c];
    }
    block_max_sum(mx, sm);
    if (threadIdx.x == 0) { pmax[blockIdx.x] = mx; psum[blockIdx.x] = sm; }
}

static int launch_norm(sg_uniform *c, const double *a, long len, double *out)
{
    const unsigned nb = std::min<unsigned>(nblocks(len), (unsigned)kNormBlocks);
    SG_REQUIRE((int)nb <= c->npart, "norm partial buffer too small");
    hipLaunchKernelGGL(k_norm_partial, dim3(nb), dim3(kBlock), 0, c->stream, a, len, c->pmax, c->psum);
    SG_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_norm_finalize, dim3(1), dim3(1024), 0, c->stream, c->pmax, c->psum, (int)nb, out);
    SG_CHECK_LAUNCH();
    return SG_OK;
}

template <typename T> static int launch_restrict(hipStream_t s, T *uc, const T *uf, int nc)
{
    hipLaunchKernelGGL(k_restrict<T>, dim3(nblocks(n3(nc))), dim3(kBlock), 0, s, uc, uf, ilog2(nc));
    SG_CHECK_LAUNCH();
    return SG_OK;
}

constexpr int kCubicFX = 32;

// FMG tricubic prolongation of a fine level n >= 32 through LDS tiles: 32 x 16 x 16 fine cells per
// 512-thread workgroup in FP32, 32 x 8 x 8 per 256 threads in FP64 (its LDS bound).
template <typename T> static int launch_interp_cubic_tile(hipStream_t s, T *uf, const T *uc, int nf)
{
    constexpr bool F = std::is_same<T, float>::value;
    constexpr int FY = F ? 16 : 8, FZ = F ? 16 : 8, NT = F ? 512 : 256;
    SG_REQUIRE(nf >= kCubicFX && nf % kCubicFX == 0 && nf % FY == 0 && nf % FZ == 0,
               "tiled tricubic needs nf >= 32, got %d", nf);
    const int ntiles = (nf / kCubicFX) * (nf / FY) * (nf / FZ);
    hipLaunchKernelGGL((k_interp_cubic_tile<kCubicFX, FY, FZ, NT, T>), dim3(8 * ((ntiles + 7) / 8)), dim3(NT), 0, s,
                       uf, uc, nf);
    SG_CHECK_LAUNCH();
    return SG_OK;
}

// mg_interp (mg.F90:608-633) on a periodic grid: the ghost fills become index wrap, so only the
// prolongation itself remains.  Tricubic needs a coarse array of >= 4 cells with ghosts
// (interpolation.F90:87-96); narrower ones fall back to trilinear.  Fine levels of >= 32 cells
// take the LDS-tiled tricubic (same values as the per-cell form).
template <typename T> static int launch_interp(hipStream_t s, T *uf, const T *uc, int nf, bool cubic, bool add)
{
    const int nc = nf / 2;
    const unsigned nb = nblocks(n3(nf));
    if (cubic && !add && nf >= kCubicFX) return launch_interp_cubic_tile<T>(s, uf, uc, nf);
    if (cubic && nc + 2 >= 4 && !add) {
        hipLaunchKernelGGL(k_interp_cubic<T>, dim3(nb), dim3(kBlock), 0, s, uf, uc, ilog2(nf));
    } else if (add) {
        hipLaunchKernelGGL((k_interp_linear<T, true>), dim3(nb), dim3(kBlock), 0, s, uf, uc, ilog2(nf));
    } else {
        hipLaunchKernelGGL((k_interp_linear<T, false>), dim3(nb), dim3(kBlock), 0, s, uf, uc, ilog2(nf));
    }
    SG_CHECK_LAUNCH();
    return SG_OK;
}

// mg.F90:106-218 with Ncycle V(Npre,Npost) cycles per FMG level.  The reference copies RHO into a
// separate RHS array at each FMG level (mg.F90:176) and later overwrites RHS of coarser levels
// with restricted residuals (mg.F90:188).  RHO of a level is never read again after that copy
// within one mg_lin call (the next call restricts RHO afresh from level 0), so RHS and RHO share
// storage here; every kernel reads the same values as in the reference.
// resmax: receives max|resh2| of the last smoothing call (the one mg_cycle's test reads).
// exact (FP32 only, and only when the first level of <= 16^3 cells lies below level 0): the levels
// end at level lb (the first level below level 0 with <= 64^3 cells, or with <= 16^3), solved
// exactly (exact_bottom) where the reference sweeps its coarsest level (mg_slvsml,
// mg.F90:540-546) -- the fused path's coarse-grid solve, so both paths keep the same values.
static int exact_bottom(sg_uniform *c, int mode, int ncycle, int npre, int npost);

template <typename T> static int mg_lin(sg_uniform *c, int ncycle, int npre, int npost, double *resmax,
                                        bool exact = false)
{
    Levels<T> &A = levels<T>(c);
    hipStream_t s = c->stream;
    exact = exact && std::is_same<T, float>::value && c->coarse.exact;
    const int lmax = exact ? c->lb : c->nlev - 1;
    auto bottom = [&]() -> int {
        if constexpr (std::is_same<T, float>::value) {
            if (exact) return exact_bottom(c, 0, ncycle, npre, npost);
        }
        for (int it = 0; it < 2; ++it)
            SG_TRY(launch_relax<T>(s, A.U[lmax], A.RHO[lmax], c->nl[lmax], c->hl[lmax], nullptr));
        return SG_OK;
    };
    for (int l = 1; l <= lmax; ++l) SG_TRY(launch_restrict<T>(s, A.RHO[l], A.RHO[l - 1], c->nl[l]));
    SG_TRY(launch_fill<T>(s, A.U[lmax], n3(c->nl[lmax])));
    SG_TRY(bottom());
    for (int ml = lmax - 1; ml >= 0; --ml) {
        SG_TRY(launch_interp<T>(s, A.U[ml], A.U[ml + 1], c->nl[ml], true, false));
        for (int jc = 0; jc < ncycle; ++jc) {
            for (int v = ml; v <= lmax - 1; ++v) {
                for (int p = 0; p < npre; ++p)
                    SG_TRY(launch_relax<T>(s, A.U[v], A.RHO[v], c->nl[v], c->hl[v], nullptr));
                SG_TRY((launch_resid<T, T>(c, A.RES[v], A.U[v], A.RHO[v], c->nl[v], c->hl[v], nullptr)));
                SG_TRY(launch_restrict<T>(s, A.RHO[v + 1], A.RES[v], c->nl[v + 1]));
                SG_TRY(launch_fill<T>(s, A.U[v + 1], n3(c->nl[v + 1])));
            }
            SG_TRY(bottom());
            for (int v = lmax - 1; v >= ml; --v) {
                SG_TRY(launch_interp<T>(s, A.U[v], A.U[v + 1], c->nl[v], false, true));
                for (int p = 0; p < npost; ++p) {
                    const bool last = (ml == 0 && jc == ncycle - 1 && v == 0 && p == npost - 1);
                    SG_TRY(launch_relax<T>(s, A.U[v], A.RHO[v], c->nl[v], c->hl[v], last ? resmax : nullptr));
                }
            }
        }
    }
    return SG_OK;
}

template <class TL, int NS, bool FS>
static void sweep_variant(hipStream_t s, const float *uin, float *uout, const float *f, int n, float h2, float h2i,
                          bool zero, bool prolong, bool resid, const float *uc, float *rc, double *resmax)
{
    const int ntiles = (n / TL::TX) * (n / TL::TY) * (n / TL::TZ);
    const dim3 grid(8 * ((ntiles + 7) / 8)), block(TL::NT);
    if (resid && zero)
        hipLaunchKernelGGL((k_sweep_tile<TL, NS, true, false, true, FS>), grid, block, 0, s, uin, uout, f, n, h2, h2i, uc, rc, resmax);
    else if (resid)
        hipLaunchKernelGGL((k_sweep_tile<TL, NS, false, false, true, FS>), grid, block, 0, s, uin, uout, f, n, h2, h2i, uc, rc, resmax);
    else if (prolong)
        hipLaunchKernelGGL((k_sweep_tile<TL, NS, false, true, false, FS>), grid, block, 0, s, uin, uout, f, n, h2, h2i, uc, rc, resmax);
    else if (zero)
        hipLaunchKernelGGL((k_sweep_tile<TL, NS, true, false, false, FS>), grid, block, 0, s, uin, uout, f, n, h2, h2i, uc, rc, resmax);
    else
        hipLaunchKernelGGL((k_sweep_tile<TL, NS, false, false, false, FS>), grid, block, 0, s, uin, uout, f, n, h2, h2i, uc, rc, resmax);
}

template <class TL>
static void sweep_ns(hipStream_t s, int ns, bool fs, const float *uin, float *uout, const float *f, int n, float h2,
                     float h2i, bool zero, bool prolong, bool resid, const float *uc, float *rc, double *resmax)
{
    if (ns == 2 && fs) {
        if constexpr (sweep_fits_lds<TL, 2, true>())
            sweep_variant<TL, 2, true>(s, uin, uout, f, n, h2, h2i, zero, prolong, resid, uc, rc, resmax);
    } else if (ns == 2) {
        if constexpr (sweep_fits_lds<TL, 2, false>())
            sweep_variant<TL, 2, false>(s, uin, uout, f, n, h2, h2i, zero, prolong, resid, uc, rc, resmax);
    } else if (fs) {
        if constexpr (sweep_fits_lds<TL, 1, true>())
            sweep_variant<TL, 1, true>(s, uin, uout, f, n, h2, h2i, zero, prolong, resid, uc, rc, resmax);
    } else {
        sweep_variant<TL, 1, false>(s, uin, uout, f, n, h2, h2i, zero, prolong, resid, uc, rc, resmax);
    }
}

// ns (1 or 2) fused red-black sweeps of a fine level (n >= 32) with tile shape cfg, the source
// staged in LDS when fs: uin -> uout.
template <class SS>
static void stream_variant(hipStream_t s, const float *uin, float *uout, const float *f, int n, float h2, float h2i,
                           bool zero, bool prolong, bool resid, const float *uc, float *rc, double *resmax)
{
    const int ntiles = (n / SS::SX) * (n / SS::SY) * (n / SS::ZC);
    const dim3 grid(8 * ((ntiles + 7) / 8)), block(SS::NT);
    if (resid && zero)
        hipLaunchKernelGGL((k_sweep_stream<SS, true, false, true>), grid, block, 0, s, uin, uout, f, n, h2, h2i, uc, rc, resmax);
    else if (resid)
        hipLaunchKernelGGL((k_sweep_stream<SS, false, false, true>), grid, block, 0, s, uin, uout, f, n, h2, h2i, uc, rc, resmax);
    else if (prolong)
        hipLaunchKernelGGL((k_sweep_stream<SS, false, true, false>), grid, block, 0, s, uin, uout, f, n, h2, h2i, uc, rc, resmax);
    else if (zero)
        hipLaunchKernelGGL((k_sweep_stream<SS, true, false, false>), grid, block, 0, s, uin, uout, f, n, h2, h2i, uc, rc, resmax);
    else
        hipLaunchKernelGGL((k_sweep_stream<SS, false, false, false>), grid, block, 0, s, uin, uout, f, n, h2, h2i, uc, rc, resmax);
}

static int launch_sweep_cfg(hipStream_t s, int cfg, int ns, bool fs, const float *uin, float *uout, const float *f,
                            int n, double h, bool zero, bool prolong, bool resid, const float *uc, float *rc,
                            double *resmax)
{
    SG_REQUIRE(cfg >= 0 && cfg < kNumCfg && cfg_fits(cfg, n), "sweep shape %d does not fit n = %d", cfg, n);
    SG_REQUIRE(ns >= 1 && ns <= kMaxSweepsPerLaunch, "bad sweeps per launch %d", ns);
    SG_REQUIRE(cfg_ok(cfg, ns, fs), "sweep shape %d with %d sweeps (staged %d) does not exist", cfg, ns, (int)fs);
    const float h2 = (float)(h * h), h2i = (float)(1.0 / (h * h));
    switch (cfg) {
    case kNumTiles + 0: stream_variant<Stream0>(s, uin, uout, f, n, h2, h2i, zero, prolong, resid, uc, rc, resmax); break;
    case kNumTiles + 1: stream_variant<Stream1>(s, uin, uout, f, n, h2, h2i, zero, prolong, resid, uc, rc, resmax); break;
    case kNumTiles + 2: stream_variant<Stream2>(s, uin, uout, f, n, h2, h2i, zero, prolong, resid, uc, rc, resmax); break;
    case kNumTiles + 3: stream_variant<Stream3>(s, uin, uout, f, n, h2, h2i, zero, prolong, resid, uc, rc, resmax); break;
    case kNumTiles + 4: stream_variant<Stream4>(s, uin, uout, f, n, h2, h2i, zero, prolong, resid, uc, rc, resmax); break;
    case kNumTiles + 5: stream_variant<Stream5>(s, uin, uout, f, n, h2, h2i, zero, prolong, resid, uc, rc, resmax); break;
    case kNumTiles + 6: stream_variant<Stream6>(s, uin, uout, f, n, h2, h2i, zero, prolong, resid, uc, rc, resmax); break;
    case kNumTiles + 7: stream_variant<Stream7>(s, uin, uout, f, n, h2, h2i, zero, prolong, resid, uc, rc, resmax); break;
    case kNumTiles + 8: stream_variant<Stream8>(s, uin, uout, f, n, h2, h2i, zero, prolong, resid, uc, rc, resmax); break;
    case 0: sweep_ns<Tile0>(s, ns, fs, uin, uout, f, n, h2, h2i, zero, prolong, resid, uc, rc, resmax); break;
    case 1: sweep_ns<Tile1>(s, ns, fs, uin, uout, f, n, h2, h2i, zero, prolong, resid, uc, rc, resmax); break;
    case 2: sweep_ns<Tile2>(s, ns, fs, uin, uout, f, n, h2, h2i, zero, prolong, resid, uc, rc, resmax); break;
    case 3: sweep_ns<Tile3>(s, ns, fs, uin, uout, f, n, h2, h2i, zero, prolong, resid, uc, rc, resmax); break;
    default: sweep_ns<Tile4>(s, ns, fs, uin, uout, f, n, h2, h2i, zero, prolong, resid, uc, rc, resmax); break;
    }
    SG_CHECK_LAUNCH();
    return SG_OK;
}

static int level_of(const sg_uniform *c, int n)
{
    for (int l = 0; l < c->nlev; ++l)
        if (c->nl[l] == n) return l;
    return -1;
}

// One launch of ns sweeps on level l with its autotuned tile shape.
static int launch_sweep(sg_uniform *c, int l, int ns, const float *uin, float *uout, bool zero, bool prolong,
                        bool resid, const float *uc, float *rc, double *resmax)
{
    return launch_sweep_cfg(c->stream, c->tile_cfg[l], ns, c->tile_fs[l] != 0, uin, uout, c->f.RHO[l], c->nl[l],
                            c->hl[l], zero, prolong, resid, uc, rc, resmax);
}

// Autotuner: for every fine level, time each fitting (tile shape, sweeps per launch) on the
// level's own arrays as ms per fine-level visit of a V(2,2) cycle -- two sweeps ending with the
// residual restriction, two starting with the prolongated correction, in the launches the cycle
// makes -- and keep the fastest.  Runs once per context; values do not depend on the choice.
static int autotune_tiles(sg_uniform *c)
{
    hipEvent_t a = nullptr, b = nullptr;
    SG_HIP(hipEventCreate(&a));
    SG_HIP(hipEventCreate(&b));
    int rc = SG_OK;
    for (int l = 0; l < c->nlev && rc == SG_OK; ++l) {
        const int n = c->nl[l];
        c->tile_cfg[l] = -1;
        c->tile_ns[l] = 1;
        if (n <= kCoarseMaxN || l + 1 >= c->nlev) continue;
        float best = 1e30f;
        for (int cand = 0; cand < kNumCand && rc == SG_OK; ++cand) {
            int cfg, ns;
            bool fs;
            cand_decode(cand, cfg, ns, fs);
            c->tune_ms[l][cand] = -1.f;
            if (!cfg_fits(cfg, n) || !cfg_ok(cfg, ns, fs)) continue;
            const int reps = n >= 256 ? 3 : 6;  // timed visits, after one untimed
            int flip = 0;
            auto launch = [&](int k, bool prolong, bool resid) {
                const int r = launch_sweep_cfg(c->stream, cfg, k, fs, flip ? c->U1[l] : c->f.U[l],
                                               flip ? c->f.U[l] : c->U1[l], c->f.RHO[l], n, c->hl[l], false, prolong,
                                               resid, prolong ? c->f.U[l + 1] : nullptr,
                                               resid ? c->f.RHO[l + 1] : nullptr, nullptr);
                flip ^= 1;
                return r;
            };
            for (int r = 0; r < reps + 1 && rc == SG_OK; ++r) {
                if (r == 1 && hipEventRecord(a, c->stream) != hipSuccess) rc = SG_ERR_HIP;
                if (rc == SG_OK && ns == 1) {
                    if ((rc = launch(1, false, false)) == SG_OK && (rc = launch(1, false, true)) == SG_OK &&
                        (rc = launch(1, true, false)) == SG_OK)
                        rc = launch(1, false, false);
                } else if (rc == SG_OK) {
                    if ((rc = launch(2, false, true)) == SG_OK) rc = launch(2, true, false);
                }
            }
            if (rc != SG_OK) break;
            float ms = 0.f;
            if (hipEventRecord(b, c->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
                hipEventElapsedTime(&ms, a, b) != hipSuccess) {
                rc = SG_ERR_HIP;
                break;
            }
            ms /= reps;
            c->tune_ms[l][cand] = ms;
            if (ms < best) {
                best = ms;
                c->tile_cfg[l] = cfg;
                c->tile_ns[l] = ns;
                c->tile_fs[l] = fs;
            }
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc == SG_ERR_HIP) set_error("autotune: HIP call failed");
    return rc;
}

// The levels of <= 16^3 cells held by k_coarse (from the first such level, lc) and the Hartley
// basis of level lc for its exact solve, which applies only when lc > 0 (a grid of <= 16^3 cells
// keeps the V-cycle: its smoothing also yields Resmaxg).
static int coarse_levels(sg_uniform *c)
{
    const int lmax = c->nlev - 1;
    int lc = 0;
    while (c->nl[lc] > kCoarseMaxN) ++lc;
    CoarseLevels &L = c->coarse;
    L = CoarseLevels{};
    L.nlev = lmax - lc + 1;
    SG_REQUIRE(L.nlev <= 8, "too many coarse levels");
    int off = 0;
    for (int q = 0; q < L.nlev; ++q) {
        const int l = lc + q;
        L.n[q] = c->nl[l];
        L.lg[q] = 0;
        while ((1 << L.lg[q]) < c->nl[l]) ++L.lg[q];
        L.h2[q] = (float)(c->hl[l] * c->hl[l]);
        L.h2i[q] = (float)(1.0 / (c->hl[l] * c->hl[l]));
        L.off[q] = off;
        off += c->nl[l] * c->nl[l] * c->nl[l];
    }
    SG_REQUIRE(off <= kCoarseCells, "coarse levels exceed LDS capacity");
    L.exact = lc > 0;
    if (L.exact) {
        const int m = c->nl[lc];
        double Q[kCoarseMaxN * kCoarseMaxN], lam[kCoarseMaxN];
        hartley_basis(m, c->hl[lc], Q, lam);
        for (int e = 0; e < m * m; ++e) L.Q[e] = (float)Q[e];
        for (int e = 0; e < m; ++e) L.lam[e] = (float)lam[e];
    }
    c->lc = lc;
    // the exact-coarse cycle ends at the first level below level 0 with <= 64^3 cells; with more
    // than 16^3 cells it is solved by the plane-parallel kernels, else by k_coarse
    int lb = lc;
    if (lc > 0) {
        int l = 1;
        while (c->nl[l] > kPlaneMaxN) ++l;
        if (c->nl[l] > kCoarseMaxN) lb = l;
    }
    c->lb = lb;
    if (lb != lc) {
        const int m = c->nl[lb];
        SG_REQUIRE(m % 16 == 0 && m <= kPlaneMaxN, "plane-parallel exact solve needs n %% 16 == 0, n <= 64");
        std::vector<double> Q((size_t)m * m), lam(m);
        hartley_basis(m, c->hl[lb], Q.data(), lam.data());
        std::vector<float> Qf(Q.begin(), Q.end()), inv((size_t)m * m * m);
        exact_inverse_table(m, lam.data(), inv.data());
        SG_HIP(hipMalloc(&c->d_qx, sizeof(float) * Qf.size()));
        SG_HIP(hipMalloc(&c->d_invx, sizeof(float) * inv.size()));
        SG_HIP(hipMemcpy(c->d_qx, Qf.data(), sizeof(float) * Qf.size(), hipMemcpyHostToDevice));
        SG_HIP(hipMemcpy(c->d_invx, inv.data(), sizeof(float) * inv.size(), hipMemcpyHostToDevice));
    }
    return SG_OK;
}

// k_coarse on the held levels; exact: solve level lc exactly (needs c->coarse.exact).
static int launch_coarse_k(sg_uniform *c, bool exact, const float *rhs, float *u, int mode, int ncycle, int npre,
                           int npost, double *resmax)
{
    SG_REQUIRE(!exact || c->coarse.exact, "exact coarse solve needs a level below the finest");
    CoarseLevels L = c->coarse;
    L.exact = exact ? 1 : 0;
    hipLaunchKernelGGL(k_coarse, dim3(1), dim3(kCoarseThreads), 0, c->stream, rhs, u, L, mode, ncycle, npre, npost,
                       resmax);
    SG_CHECK_LAUNCH();
    return SG_OK;
}

// U[lb] := exact solution for RHO[lb] (FP32): k_coarse when lb is the first level of <= 16^3
// cells, else the three plane-parallel launches with U1[lb] as the intermediate array.
static int exact_bottom(sg_uniform *c, int mode, int ncycle, int npre, int npost)
{
    const int lb = c->lb, n = c->nl[lb];
    SG_REQUIRE(c->coarse.exact, "exact coarse solve needs a level below the finest");
    if (lb == c->lc) return launch_coarse_k(c, true, c->f.RHO[lb], c->f.U[lb], mode, ncycle, npre, npost, nullptr);
    SG_REQUIRE(n <= kPlaneMaxN && n % 16 == 0 && c->d_qx && c->d_invx && c->U1[lb],
               "plane-parallel exact solve not set up (n = %d)", n);
    const dim3 grid(n * plane_groups(n)), block(kSplitThreads);
    hipLaunchKernelGGL(k_exact_xy_split, grid, block, 0, c->stream, (const float *)c->f.RHO[lb], c->U1[lb],
                       (const float *)c->d_qx, n);
    SG_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_exact_z_split, grid, block, 0, c->stream, c->U1[lb], (const float *)c->d_qx,
                       (const float *)c->d_invx, n);
    SG_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_exact_xy_split, grid, block, 0, c->stream, (const float *)c->U1[lb], c->f.U[lb],
                       (const float *)c->d_qx, n);
    SG_CHECK_LAUNCH();
    return SG_OK;
}

// mg_lin (mg.F90:106-218) in FP32 with fused kernels: the levels of <= 16^3 cells run inside one
// single-workgroup kernel (k_coarse); each fine-level visit of a V-cycle is npre + npost fused
// sweeps, the last pre-sweep carrying the residual and its restriction, the first post-sweep the
// prolongated correction.  With `exact` the cycle instead ends at level lb, solved exactly
// (exact_bottom).  Same values as mg_lin<float> with the same `exact`, bit for bit.  *u0 receives
// the buffer that holds the level-0 correction.  rhs1_ready: RHO[1] already holds the restriction
// of RHO[0] (formed with the residual by k_update_resid).
static int fast_mg_lin(sg_uniform *c, int ncycle, int npre, int npost, double *resmax, const float **u0,
                       bool exact, bool rhs1_ready)
{
    SG_REQUIRE(npre >= 1 && npost >= 1, "fused path needs npre, npost >= 1");
    Levels<float> &A = c->f;
    hipStream_t s = c->stream;
    exact = exact && c->coarse.exact;
    const int lc = exact ? c->lb : c->lc;   // level of the coarse-grid solve
    int cur[kMaxLev] = {};
    auto buf = [&](int l, int b) -> float * { return b ? c->U1[l] : A.U[l]; };
    for (int l = rhs1_ready ? 2 : 1; l <= lc; ++l) SG_TRY(launch_restrict<float>(s, A.RHO[l], A.RHO[l - 1], c->nl[l]));
    if (exact) SG_TRY(exact_bottom(c, 0, ncycle, npre, npost));
    else SG_TRY(launch_coarse_k(c, false, A.RHO[lc], A.U[lc], 0, ncycle, npre, npost, lc == 0 ? resmax : nullptr));
    cur[lc] = 0;
    for (int ml = lc - 1; ml >= 0; --ml) {
        SG_TRY(launch_interp_cubic_tile(s, A.U[ml], buf(ml + 1, cur[ml + 1]), c->nl[ml]));
        cur[ml] = 0;
        for (int jc = 0; jc < ncycle; ++jc) {
            for (int v = ml; v <= lc - 1; ++v) {
                // npre sweeps in launches of up to tile_ns[v]; the first may start from zero, the
                // last carries the residual restriction
                for (int r = npre, first = 1; r > 0; first = 0) {
                    const int k = std::min(r, c->tile_ns[v]);
                    const bool last = r == k;
                    SG_TRY(launch_sweep(c, v, k, buf(v, cur[v]), buf(v, 1 - cur[v]), first && v > ml, false, last,
                                        nullptr, last ? A.RHO[v + 1] : nullptr, nullptr));
                    cur[v] ^= 1;
                    r -= k;
                }
            }
            if (exact) SG_TRY(exact_bottom(c, 1, ncycle, npre, npost));
            else SG_TRY(launch_coarse_k(c, false, A.RHO[lc], A.U[lc], 1, ncycle, npre, npost, nullptr));
            cur[lc] = 0;
            for (int v = lc - 1; v >= ml; --v) {
                // npost sweeps; the first launch adds the prolongated correction
                for (int r = npost, first = 1; r > 0; first = 0) {
                    const int k = std::min(r, c->tile_ns[v]);
                    const bool lastcall = (ml == 0 && jc == ncycle - 1 && v == 0 && r == k);
                    SG_TRY(launch_sweep(c, v, k, buf(v, cur[v]), buf(v, 1 - cur[v]), false, first, false,
                                        first ? buf(v + 1, cur[v + 1]) : nullptr, nullptr, lastcall ? resmax : nullptr));
                    cur[v] ^= 1;
                    r -= k;
                }
            }
        }
    }
    *u0 = buf(0, cur[0]);
    return SG_OK;
}

template <typename Tu> static int launch_update(sg_uniform *c, double *psi, const Tu *u)
{
    hipLaunchKernelGGL(k_update<Tu>, dim3(nblocks(n3(c->n))), dim3(kBlock), 0, c->stream, psi, u, n3(c->n));
    SG_CHECK_LAUNCH();
    return SG_OK;
}

// The norms to the host (published: the last kernel already wrote them into hnorm).  The host
// polls an event rather than blocking in a stream synchronize: the outer loop waits on this
// between iterations, and polling wakes it as soon as the values are there (the GPU idles until
// the next iteration is launched).
static int read_norms(sg_uniform *c, bool published = false)
{
    if (!published) SG_HIP(hipMemcpyAsync(c->hnorm, c->dnorm, 5 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SG_HIP(hipEventRecord(c->evn, c->stream));
    hipError_t q;
    while ((q = hipEventQuery(c->evn)) == hipErrorNotReady) {
    }
    SG_HIP(q);
    return SG_OK;
}

// e == nullptr: the first residual (no update); pout, when given, receives a copy of pin, and with
// fnorms max|src| and sum src^2 land in dnorm[3..4].
static int launch_update_resid(sg_uniform *c, const double *pin, const float *e, double *pout, const double *src,
                               bool fnorms = false)
{
    const int n = c->n, lg = ilog2(n);
    const long cells = n3(n);
    // z chunk: about kNormBlocks patches of 512 cells per plane
    int zc = (int)std::max<long>(4, std::min<long>(32, cells / (512L * kNormBlocks)));
    zc = std::min(zc, n);
    const bool wide = n % 256 == 0;  // 256 x 4 patches, else 64 x 8
    const int tx = wide ? 256 : 64, ty = wide ? 4 : 8;
    SG_REQUIRE(n >= 64 && n % tx == 0 && n % ty == 0 && n % zc == 0, "update+residual: bad n = %d", n);
    const dim3 grid(n / tx, n / ty, n / zc), block(tx * ty);
    const int nb = (int)(grid.x * grid.y * grid.z);
    SG_REQUIRE(nb <= c->npart, "norm partial buffer too small");
    SG_REQUIRE(c->nlev > 1 && zc % 2 == 0, "update+residual: no level to restrict into");
    const double h2i = 1.0 / (c->h * c->h);
    // the residual restricted into RHO[1] here: the fused inner solve starts from it
    double *fpmax = fnorms ? c->pmax + c->npart : nullptr, *fpsum = fnorms ? c->psum + c->npart : nullptr;
    if (e && wide)
        hipLaunchKernelGGL((k_update_resid<true, 256, 4>), grid, block, 0, c->stream, pin, e, pout, src, c->f.RHO[0],
                           lg, h2i, zc, c->pmax, c->psum, c->f.RHO[1], fpmax, fpsum);
    else if (e)
        hipLaunchKernelGGL((k_update_resid<true, 64, 8>), grid, block, 0, c->stream, pin, e, pout, src, c->f.RHO[0],
                           lg, h2i, zc, c->pmax, c->psum, c->f.RHO[1], fpmax, fpsum);
    else if (wide)
        hipLaunchKernelGGL((k_update_resid<false, 256, 4>), grid, block, 0, c->stream, pin, e, pout, src, c->f.RHO[0],
                           lg, h2i, zc, c->pmax, c->psum, c->f.RHO[1], fpmax, fpsum);
    else
        hipLaunchKernelGGL((k_update_resid<false, 64, 8>), grid, block, 0, c->stream, pin, e, pout, src, c->f.RHO[0],
                           lg, h2i, zc, c->pmax, c->psum, c->f.RHO[1], fpmax, fpsum);
    SG_CHECK_LAUNCH();
    if (fnorms) {
        hipLaunchKernelGGL(k_norm_finalize, dim3(1), dim3(1024), 0, c->stream, (const double *)fpmax,
                           (const double *)fpsum, nb, c->dnorm + 3, (double *)nullptr);
        SG_CHECK_LAUNCH();
    }
    // the residual's norms, published with Resmaxg and the source norms straight into hnorm
    hipLaunchKernelGGL(k_norm_finalize, dim3(1), dim3(1024), 0, c->stream, c->pmax, c->psum, nb, c->dnorm, c->hnorm_dev);
    SG_CHECK_LAUNCH();
    return SG_OK;
}

// whether the mixed solve uses the fused update+residual step with a second phi buffer
static bool pingpong(const sg_uniform *c, const sg_params *p)
{
    return p->precision == SG_PREC_MIXED && p->fused && c->n >= 64 && c->psi_alt;
}

// One outer iteration after the first residual: inner solve on RHO[0] (= f - L phi), update
// phi += U[0] (from pin into pout; in place when pin == pout), next residual with its norms into
// dnorm[0..1].  Resmaxg lands in dnorm[2].
static int outer_iteration(sg_uniform *c, const double *pin, double *pout, const double *src, const sg_params *p)
{
    // Resmaxg is only read by the reference stopping test; the true-residual test skips it.
    double *resmax = p->stop_mode == SG_STOP_REFERENCE ? c->dnorm + 2 : nullptr;
    SG_TRY(launch_fill<double>(c->stream, c->dnorm + 2, 1));
    if (p->precision == SG_PREC_MIXED) {
        const float *u0 = c->f.U[0];
        const bool exact = p->coarse == SG_COARSE_EXACT;
        // with pin != pout every residual came from launch_update_resid, which restricted it too
        if (p->fused) SG_TRY(fast_mg_lin(c, p->ncycle, p->npre, p->npost, resmax, &u0, exact, pin != pout));
        else SG_TRY(mg_lin<float>(c, p->ncycle, p->npre, p->npost, resmax, exact));
        if (pin != pout) {
            SG_TRY(launch_update_resid(c, pin, u0, pout, src));
        } else {
            SG_TRY(launch_update<float>(c, pout, u0));
            SG_TRY((launch_resid<double, float>(c, c->f.RHO[0], pout, src, c->n, c->h, c->dnorm)));
        }
    } else {
        SG_TRY(mg_lin<double>(c, p->ncycle, p->npre, p->npost, resmax));
        SG_TRY(launch_update<double>(c, pout, c->d.U[0]));
        SG_TRY((launch_resid<double, double>(c, c->d.RHO[0], pout, src, c->n, c->h, c->dnorm)));
    }
    return SG_OK;
}

static void drop_graphs(sg_uniform *c)
{
    for (auto &g : c->graph) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g = sg_uniform::Graph();
    }
}

static int run_outer_iteration(sg_uniform *c, const double *pin, double *pout, const double *src,
                               const sg_params *p, int parity)
{
    if (!(p->use_graph && p->precision == SG_PREC_MIXED)) return outer_iteration(c, pin, pout, src, p);
    sg_uniform::Graph &g = c->graph[parity & 1];
    const bool same = g.exec && g.pin == pin && g.pout == pout && g.src == src && g.ncycle == p->ncycle &&
                      g.npre == p->npre && g.npost == p->npost && g.fused == p->fused && g.stop == p->stop_mode &&
                      g.coarse == p->coarse;
    if (!same) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g = sg_uniform::Graph();
        hipGraph_t graph = nullptr;
        SG_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int rc = outer_iteration(c, pin, pout, src, p);
        const hipError_t e = hipStreamEndCapture(c->stream, &graph);
        if (rc != SG_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        SG_HIP(e);
        const hipError_t e2 = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        SG_HIP(e2);
        g.pin = pin;
        g.pout = pout;
        g.src = src;
        g.ncycle = p->ncycle;
        g.npre = p->npre;
        g.npost = p->npost;
        g.fused = p->fused;
        g.coarse = p->coarse;
        g.stop = p->stop_mode;
    }
    SG_HIP(hipGraphLaunch(g.exec, c->stream));
    return SG_OK;
}

// mg.F90:42-94: FP64 defect correction psi += L^-1 (src - L psi) with the chosen inner solver.
static int uniform_cycle(sg_uniform *c, double *psi, const double *src, const sg_params *p, sg_stats *st)
{
    SG_REQUIRE(c && psi && src, "null context or array");
    SG_REQUIRE(p && st, "null params/stats");
    SG_REQUIRE(p->precision == SG_PREC_FP64 || p->precision == SG_PREC_MIXED, "bad precision %d", p->precision);
    SG_REQUIRE(c->precisions & p->precision, "context was not created with precision %d", p->precision);
    SG_REQUIRE(p->maxit >= 0 && p->maxit <= 1000, "bad maxit %d", p->maxit);
    SG_REQUIRE(p->ncycle >= 1 && p->ncycle <= 64 && p->npre >= 0 && p->npre <= 64 && p->npost >= 0 && p->npost <= 64,
               "bad cycle shape");
    SG_REQUIRE(p->stop_mode == SG_STOP_REFERENCE || p->stop_mode == SG_STOP_TRUE_RESIDUAL, "bad stop_mode");
    SG_REQUIRE(p->tol >= 0.0, "bad tol");
    SG_REQUIRE(!p->fused || p->precision != SG_PREC_MIXED || (p->npre >= 1 && p->npost >= 1),
               "fused mixed path needs npre, npost >= 1");
    SG_REQUIRE(p->coarse == SG_COARSE_EXACT || p->coarse == SG_COARSE_VCYCLE, "bad coarse %d", p->coarse);
    std::memset(st, 0, sizeof(*st));
    const bool mixed = p->precision == SG_PREC_MIXED;
    SG_HIP(hipEventRecord(c->ev0, c->stream));
    const bool pp = pingpong(c, p);
    // With the fused update the iterations alternate between psi and psi_alt, so an odd count ends
    // in psi_alt and costs a final copy.  When the previous solve on this context took an odd
    // count, the first residual step also copies psi into psi_alt and the iterations start there.
    const bool start_alt = pp && c->last_odd;
    if (pp) {
        SG_TRY(launch_update_resid(c, psi, nullptr, start_alt ? c->psi_alt : nullptr, src, true));
    } else {
        SG_TRY(launch_norm(c, src, n3(c->n), c->dnorm + 3));
        if (mixed) SG_TRY((launch_resid<double, float>(c, c->f.RHO[0], psi, src, c->n, c->h, c->dnorm)));
        else SG_TRY((launch_resid<double, double>(c, c->d.RHO[0], psi, src, c->n, c->h, c->dnorm)));
    }
    double *cur = start_alt ? c->psi_alt : psi;  // buffer holding the current phi (psi or psi_alt)
    SG_TRY(read_norms(c, pp));
    const double fmax = c->hnorm[3], fl2 = std::sqrt(c->hnorm[4]);
    const double rhomax = c->hnorm[0];
    int it = 0;
    int status = SG_ERR_NOT_CONVERGED;
    for (;;) {
        const double rmax = c->hnorm[0], rl2 = std::sqrt(c->hnorm[1]);
        st->rel_linf = fmax > 0 ? rmax / fmax : 0.0;
        st->rel_l2 = fl2 > 0 ? rl2 / fl2 : 0.0;
        st->rel_h2 = st->rel_linf;   // one cell width: the h^2 weights cancel
        if (it < SG_HIST) st->hist[it] = st->rel_linf;
        if (!std::isfinite(rmax) || !std::isfinite(rl2)) { status = SG_ERR_NONFINITE; break; }
        if (p->stop_mode == SG_STOP_TRUE_RESIDUAL && st->rel_linf < p->tol) { status = SG_OK; break; }
        if (it > 0 && p->stop_mode == SG_STOP_REFERENCE) {
            // mg.F90:91 Resmaxg / rhomax with Resmaxg = max|resh2| / h^2 (mg_poisson.F90:77)
            const double resmaxg = c->hnorm[2] / (c->h * c->h);
            st->resmaxg = resmaxg;
            if ((rhomax > 0 ? resmaxg / rhomax : 0.0) < p->tol) { status = SG_OK; break; }
        }
        if (it >= p->maxit) break;
        double *next = pp ? (cur == psi ? c->psi_alt : psi) : cur;
        SG_TRY(run_outer_iteration(c, cur, next, src, p, cur == psi ? 0 : 1));
        cur = next;
        SG_TRY(read_norms(c, pp));
        ++it;
        st->resmaxg = c->hnorm[2] / (c->h * c->h);
    }
    if (pp) c->last_odd = it & 1;
    if (cur != psi) SG_HIP(hipMemcpyAsync(psi, cur, sizeof(double) * n3(c->n), hipMemcpyDeviceToDevice, c->stream));
    SG_HIP(hipEventRecord(c->ev1, c->stream));
    SG_HIP(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    SG_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    st->ms = ms;
    st->iterations = it;
    st->status = status;
    return status;
}

template <typename T> static int alloc_levels(sg_uniform *c, Levels<T> &A)
{
    for (int l = 0; l < c->nlev; ++l) {
        const size_t bytes = sizeof(T) * (size_t)n3(c->nl[l]);
        SG_HIP(hipMalloc(&A.U[l], bytes));
        SG_HIP(hipMalloc(&A.RES[l], bytes));
        SG_HIP(hipMalloc(&A.RHO[l], bytes));
        SG_HIP(hipMemsetAsync(A.U[l], 0, bytes, c->stream));
        SG_HIP(hipMemsetAsync(A.RES[l], 0, bytes, c->stream));
        SG_HIP(hipMemsetAsync(A.RHO[l], 0, bytes, c->stream));
    }
    return SG_OK;
}

template <typename T> static void free_levels(Levels<T> &A)
{
    for (int l = 0; l < kMaxLev; ++l) {
        if (A.U[l]) (void)hipFree(A.U[l]);
        if (A.RES[l]) (void)hipFree(A.RES[l]);
        if (A.RHO[l]) (void)hipFree(A.RHO[l]);
        A.U[l] = A.RES[l] = A.RHO[l] = nullptr;
    }
}

static void destroy(sg_uniform *c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graphs(c);
    free_levels(c->d);
    free_levels(c->f);
    for (int l = 0; l < kMaxLev; ++l)
        if (c->U1[l]) (void)hipFree(c->U1[l]);
    if (c->psi_dev) (void)hipFree(c->psi_dev);
    if (c->psi_alt) (void)hipFree(c->psi_alt);
    if (c->d_qx) (void)hipFree(c->d_qx);
    if (c->d_invx) (void)hipFree(c->d_invx);
    if (c->src_dev) (void)hipFree(c->src_dev);
    if (c->adj_src) (void)hipFree(c->adj_src);
    if (c->adj_part) (void)hipFree(c->adj_part);
    if (c->pmax) (void)hipFree(c->pmax);
    if (c->psum) (void)hipFree(c->psum);
    if (c->dnorm) (void)hipFree(c->dnorm);
    if (c->hnorm) (void)hipHostFree(c->hnorm);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->evn) (void)hipEventDestroy(c->evn);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

static int create(sg_uniform **out, int n, double h, int precisions)
{
    SG_REQUIRE(out, "null output pointer");
    *out = nullptr;
    SG_REQUIRE(n >= 2 && n <= kMaxN && (n & (n - 1)) == 0, "n must be a power of two in [2, %d], got %d", kMaxN, n);
    SG_REQUIRE(h > 0.0 && std::isfinite(h), "bad h");
    SG_REQUIRE(precisions > 0 && (precisions & ~(SG_PREC_FP64 | SG_PREC_MIXED)) == 0, "bad precisions mask");
    sg_uniform *c = new sg_uniform();
    c->n = n;
    c->h = h;
    c->precisions = precisions;
    // mg_data.F90:57-75: levels n/2^l down to 1 cell; h doubles per level (mg_data.F90:554-559)
    int lmax = 0;
    while ((1 << lmax) < n) ++lmax;
    c->nlev = lmax + 1;
    for (int l = 0; l <= lmax; ++l) {
        c->nl[l] = n >> l;
        c->hl[l] = h * (double)(1 << l);
    }
    int rc = SG_OK;
    auto fail = [&](int r) { destroy(c); return r; };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("hipStreamCreateWithFlags failed");
        return fail(SG_ERR_HIP);
    }
    if ((rc = [&]() -> int {
            SG_HIP(hipEventCreate(&c->ev0));
            SG_HIP(hipEventCreate(&c->ev1));
            SG_HIP(hipEventCreateWithFlags(&c->evn, hipEventDisableTiming));
            c->npart = (int)nblocks(n3(n));
            // two sets of partials: the residual's and (first residual step) the source's
            SG_HIP(hipMalloc(&c->pmax, sizeof(double) * 2 * c->npart));
            SG_HIP(hipMalloc(&c->psum, sizeof(double) * 2 * c->npart));
            SG_HIP(hipMalloc(&c->dnorm, sizeof(double) * 8));
            SG_HIP(hipMemsetAsync(c->dnorm, 0, sizeof(double) * 8, c->stream));
            SG_HIP(hipHostMalloc(&c->hnorm, sizeof(double) * 8, hipHostMallocDefault));
            SG_HIP(hipHostGetDevicePointer((void **)&c->hnorm_dev, c->hnorm, 0));
            SG_TRY(coarse_levels(c));
            if (precisions & SG_PREC_FP64) SG_TRY(alloc_levels<double>(c, c->d));
            if (precisions & SG_PREC_MIXED) {
                SG_TRY(alloc_levels<float>(c, c->f));
                for (int l = 0; l < c->nlev; ++l) {
                    const size_t bytes = sizeof(float) * (size_t)n3(c->nl[l]);
                    SG_HIP(hipMalloc(&c->U1[l], bytes));
                    SG_HIP(hipMemsetAsync(c->U1[l], 0, bytes, c->stream));
                }
                SG_TRY(autotune_tiles(c));
                SG_HIP(hipMalloc(&c->psi_alt, sizeof(double) * (size_t)n3(n)));
            }
            SG_HIP(hipStreamSynchronize(c->stream));
            return SG_OK;
        }()) != SG_OK)
        return fail(rc);
    *out = c;
    return SG_OK;
}

} // namespace sg

// ----------------------------------------------------------------------------------------------
// C ABI
// ----------------------------------------------------------------------------------------------
// ---- the allocation-free relaxation (sg_uniform_relax_noalloc) ----
// Its only device memory beyond the caller's arrays is the Resmaxg accumulator, kept in this
// file-scope array: kScratchDoubles doubles per GPU, held by every GPU the library runs on until the
// process ends.  In host code `scratch` names a host-side stand-in, not a device address: the
// address the kernels get is the one hipGetSymbolAddress returns for the current device, looked up
// once per device and kept in g_scratch_ptr.
constexpr int kScratchDoubles = 64;
constexpr int kScratchDevices = 64;
static __device__ __align__(256) double scratch[kScratchDoubles];
static double *g_scratch_ptr[kScratchDevices] = {};
static std::mutex g_scratch_mu;

static int scratch_address(double **out)
{
    int dev = 0;
    SG_HIP(hipGetDevice(&dev));
    SG_REQUIRE(dev >= 0 && dev < kScratchDevices, "device %d beyond the %d the static scratch table covers", dev,
               kScratchDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        double *p = nullptr;
        const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            sg::set_error("hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
            return SG_ERR_HIP;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return SG_OK;
}

extern "C" {

const char *sg_last_error(void) { return sg::last_error(); }

int sg_version(void) { return 1; }

void sg_default_params(sg_params *p)
{
    if (!p) return;
    p->tol = 1e-2;             // mg.F90:50 errormax
    p->maxit = 10;             // mg.F90:52 nmg
    p->stop_mode = SG_STOP_REFERENCE;
    p->precision = SG_PREC_FP64;
    p->ncycle = 2;             // mg.F90:123-125
    p->npre = 2;
    p->npost = 2;
    p->use_graph = 1;
    p->fused = 1;
    p->coarse = SG_COARSE_EXACT;
}

int sg_uniform_create(sg_uniform **ctx, int n, double h, int precisions) { return sg::create(ctx, n, h, precisions); }

int sg_uniform_destroy(sg_uniform *ctx)
{
    sg::destroy(ctx);
    return SG_OK;
}

int sg_uniform_cycle_dev(sg_uniform *ctx, double *psi, const double *src, const sg_params *p, sg_stats *st)
{
    return sg::uniform_cycle(ctx, psi, src, p, st);
}

int sg_uniform_cycle_host(sg_uniform *c, double *psi_gh, const double *src_gh, const sg_params *p, sg_stats *st)
{
    SG_REQUIRE(c && psi_gh && src_gh, "null argument");
    const int n = c->n, ng = n + 2;
    const size_t dense = (size_t)n * n * n;
    if (!c->psi_dev) {
        SG_HIP(hipMalloc(&c->psi_dev, sizeof(double) * dense));
        SG_HIP(hipMalloc(&c->src_dev, sizeof(double) * dense));
    }
    std::vector<double> a(dense), b(dense);
    auto gi = [&](int i, int j, int k) { return ((size_t)(k + 1) * ng + (j + 1)) * ng + (i + 1); };
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                a[((size_t)k * n + j) * n + i] = psi_gh[gi(i, j, k)];
                b[((size_t)k * n + j) * n + i] = src_gh[gi(i, j, k)];
            }
    SG_HIP(hipMemcpyAsync(c->psi_dev, a.data(), sizeof(double) * dense, hipMemcpyHostToDevice, c->stream));
    SG_HIP(hipMemcpyAsync(c->src_dev, b.data(), sizeof(double) * dense, hipMemcpyHostToDevice, c->stream));
    const int rc = sg::uniform_cycle(c, c->psi_dev, c->src_dev, p, st);
    if (rc != SG_OK && rc != SG_ERR_NOT_CONVERGED && rc != SG_ERR_NONFINITE) return rc;
    SG_HIP(hipMemcpyAsync(a.data(), c->psi_dev, sizeof(double) * dense, hipMemcpyDeviceToHost, c->stream));
    SG_HIP(hipStreamSynchronize(c->stream));
    // write back with periodic ghosts (fmg_boundary_phys_periodic.F90:80-102 order i, j, k)
    auto wrap = [&](int q) { return q < 0 ? q + n : (q >= n ? q - n : q); };
    for (int k = -1; k <= n; ++k)
        for (int j = -1; j <= n; ++j)
            for (int i = -1; i <= n; ++i)
                psi_gh[gi(i, j, k)] = a[((size_t)wrap(k) * n + wrap(j)) * n + wrap(i)];
    return rc;
}

int sg_uniform_gradient_dev(sg_uniform *c, const double *psi, double *g)
{
    SG_REQUIRE(c && psi && g, "null argument");
    hipLaunchKernelGGL(sg::k_gradient, dim3(sg::nblocks(sg::n3(c->n))), dim3(sg::kBlock), 0, c->stream, g, psi,
                       sg::ilog2(c->n), 1.0 / c->h);
    SG_CHECK_LAUNCH();
    SG_HIP(hipStreamSynchronize(c->stream));
    return SG_OK;
}

int sg_uniform_gradient_sample_dev(sg_uniform *c, const double *psi, const int *cells, int npts, double *out)
{
    SG_REQUIRE(c && psi && cells && out && npts >= 0, "bad argument");
    if (npts == 0) return SG_OK;
    hipLaunchKernelGGL(sg::k_grad_sample, dim3(sg::nblocks(npts)), dim3(sg::kBlock), 0, c->stream, psi, cells, npts,
                       sg::ilog2(c->n), 1.0 / c->h, out);
    SG_CHECK_LAUNCH();
    SG_HIP(hipStreamSynchronize(c->stream));
    return SG_OK;
}

int sg_uniform_gradient_sample_adjoint_dev(sg_uniform *c, const int *cells, const double *w, int npts, double *out)
{
    SG_REQUIRE(c && cells && w && out && npts >= 0, "bad argument");
    SG_HIP(hipMemsetAsync(out, 0, sizeof(double) * sg::n3(c->n), c->stream));
    if (npts > 0) {
        hipLaunchKernelGGL(sg::k_grad_sample_adjoint, dim3(sg::nblocks(npts)), dim3(sg::kBlock), 0, c->stream, cells, w,
                           npts, sg::ilog2(c->n), 1.0 / c->h, out);
        SG_CHECK_LAUNCH();
    }
    SG_HIP(hipStreamSynchronize(c->stream));
    return SG_OK;
}

// d loss / d src of the solve src -> phi (phi = L^+ P src, P the zero-mean projection, L the
// periodic 7-point operator, which is symmetric): grad = P L^+ P adj, by the same multigrid solver.
int sg_uniform_adjoint_dev(sg_uniform *c, const double *adj, double *grad, const sg_params *p, sg_stats *st)
{
    SG_REQUIRE(c && adj && grad && p && st, "null argument");
    const long len = sg::n3(c->n);
    if (!c->adj_src) {
        SG_HIP(hipMalloc(&c->adj_src, sizeof(double) * len));
        SG_HIP(hipMalloc(&c->adj_part, sizeof(double) * sg::kMeanGroups));
    }
    auto project = [&](const double *a, double *out) -> int {
        hipLaunchKernelGGL(sg::k_sum_part, dim3(sg::kMeanGroups), dim3(256), 0, c->stream, a, len, c->adj_part);
        SG_CHECK_LAUNCH();
        hipLaunchKernelGGL(sg::k_sub_mean, dim3(1024), dim3(256), 0, c->stream, a, out, len, (const double *)c->adj_part,
                           sg::kMeanGroups);
        SG_CHECK_LAUNCH();
        return SG_OK;
    };
    SG_TRY(project(adj, c->adj_src));
    SG_HIP(hipMemsetAsync(grad, 0, sizeof(double) * len, c->stream));
    const int rc = sg::uniform_cycle(c, grad, c->adj_src, p, st);
    if (rc != SG_OK && rc != SG_ERR_NOT_CONVERGED) return rc;
    SG_TRY(project(grad, grad));
    SG_HIP(hipStreamSynchronize(c->stream));
    return rc;
}

int sg_uniform_relax_dev(sg_uniform *c, int prec, void *u, const void *rhs, int n, double h, double *resmaxg)
{
    SG_REQUIRE(c && u && rhs, "null argument");
    SG_REQUIRE(n >= 1 && n <= c->n && (n & (n - 1)) == 0, "bad n");
    SG_TRY(sg::launch_fill<double>(c->stream, c->dnorm + 2, 1));
    if (prec == SG_PREC_FP64)
        SG_TRY(sg::launch_relax<double>(c->stream, (double *)u, (const double *)rhs, n, h, c->dnorm + 2));
    else if (prec == SG_PREC_MIXED)
        SG_TRY(sg::launch_relax<float>(c->stream, (float *)u, (const float *)rhs, n, h, c->dnorm + 2));
    else
        SG_REQUIRE(false, "bad precision");
    SG_TRY(sg::read_norms(c));
    if (resmaxg) *resmaxg = c->hnorm[2] / (h * h);
    return SG_OK;
}

int sg_uniform_resid_dev(sg_uniform *c, int prec, void *res, const void *u, const void *rhs, int n, double h)
{
    SG_REQUIRE(c && res && u && rhs, "null argument");
    SG_REQUIRE(n >= 1 && n <= c->n && (n & (n - 1)) == 0, "bad n");
    if (prec == SG_PREC_FP64)
        SG_TRY((sg::launch_resid<double, double>(c, (double *)res, (const double *)u, (const double *)rhs, n, h, nullptr)));
    else if (prec == SG_PREC_MIXED)
        SG_TRY((sg::launch_resid<float, float>(c, (float *)res, (const float *)u, (const float *)rhs, n, h, nullptr)));
    else
        SG_REQUIRE(false, "bad precision");
    SG_HIP(hipStreamSynchronize(c->stream));
    return SG_OK;
}

int sg_uniform_relax_noalloc(int prec, void *u, const void *rhs, int n, double h, double *resmaxg, void *stream)
{
    SG_REQUIRE(u && rhs, "sg_uniform_relax_noalloc: null argument");
    SG_REQUIRE(n >= 1 && n <= sg::kMaxN && (n & (n - 1)) == 0,
               "sg_uniform_relax_noalloc: n must be a power of two in [1, %d], got %d", sg::kMaxN, n);
    SG_REQUIRE(h > 0.0, "sg_uniform_relax_noalloc: h must be positive");
    SG_REQUIRE(prec == SG_PREC_FP64 || prec == SG_PREC_MIXED, "sg_uniform_relax_noalloc: bad precision %d", prec);
    double *acc = nullptr;
    SG_TRY(scratch_address(&acc));
    hipStream_t s = (hipStream_t)stream;
    // the accumulator starts at zero: a fill kernel (no memset node if the caller captures the stream)
    SG_TRY(sg::launch_fill<double>(s, acc, 1));
    if (prec == SG_PREC_FP64)
        SG_TRY(sg::launch_relax<double>(s, (double *)u, (const double *)rhs, n, h, acc));
    else
        SG_TRY(sg::launch_relax<float>(s, (float *)u, (const float *)rhs, n, h, acc));
    double m = 0.0;
    SG_HIP(hipMemcpyAsync(&m, acc, sizeof(double), hipMemcpyDeviceToHost, s));
    SG_HIP(hipStreamSynchronize(s));
    if (resmaxg) *resmaxg = m / (h * h);
    return SG_OK;
}

int sg_uniform_restrict_dev(sg_uniform *c, int prec, void *uc, const void *uf, int nc)
{
    SG_REQUIRE(c && uc && uf, "null argument");
    SG_REQUIRE(nc >= 1 && 2 * nc <= c->n && (nc & (nc - 1)) == 0, "bad nc");
    if (prec == SG_PREC_FP64) SG_TRY(sg::launch_restrict<double>(c->stream, (double *)uc, (const double *)uf, nc));
    else if (prec == SG_PREC_MIXED) SG_TRY(sg::launch_restrict<float>(c->stream, (float *)uc, (const float *)uf, nc));
    else SG_REQUIRE(false, "bad precision");
    SG_HIP(hipStreamSynchronize(c->stream));
    return SG_OK;
}

int sg_uniform_time_smoother(sg_uniform *c, int prec, int reps, double *ms_per_sweep)
{
    SG_REQUIRE(c && ms_per_sweep && reps > 0 && reps <= 100000, "bad argument");
    SG_REQUIRE(c->precisions & prec, "context lacks precision %d", prec);
    const bool fused = prec == SG_PREC_MIXED && c->n > sg::kCoarseMaxN;
    SG_HIP(hipEventRecord(c->ev0, c->stream));
    for (int r = 0; r < reps; ++r) {
        if (prec == SG_PREC_FP64)
            SG_TRY(sg::launch_relax<double>(c->stream, c->d.U[0], c->d.RHO[0], c->n, c->h, nullptr));
        else if (fused) {
            // the mixed path's fine-level sweep: U[0] -> U1[0] and back on alternate repetitions
            float *a = (r & 1) ? c->U1[0] : c->f.U[0], *b = (r & 1) ? c->f.U[0] : c->U1[0];
            SG_TRY(sg::launch_sweep_cfg(c->stream, c->tile_cfg[0], 1, c->tile_fs[0] != 0, a, b, c->f.RHO[0], c->n,
                                        c->h, false, false, false, nullptr, nullptr, nullptr));
        } else
            SG_TRY(sg::launch_relax<float>(c->stream, c->f.U[0], c->f.RHO[0], c->n, c->h, nullptr));
    }
    SG_HIP(hipEventRecord(c->ev1, c->stream));
    SG_HIP(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    SG_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *ms_per_sweep = (double)ms / reps;
    return SG_OK;
}

int sg_uniform_set_tile(sg_uniform *c, int level, int cfg, int ns, int fs)
{
    SG_REQUIRE(c && (c->precisions & SG_PREC_MIXED), "context lacks the mixed-precision arrays");
    SG_REQUIRE(level >= 0 && level < c->nlev && c->nl[level] > sg::kCoarseMaxN,
               "level %d is not a fused-sweep level", level);
    SG_REQUIRE(cfg >= 0 && cfg < sg::kNumCfg && sg::cfg_fits(cfg, c->nl[level]) && sg::cfg_ok(cfg, ns, fs != 0),
               "sweep configuration (%d, %d, %d) does not fit level %d (n = %d)", cfg, ns, fs, level, c->nl[level]);
    c->tile_cfg[level] = cfg;
    c->tile_ns[level] = ns;
    c->tile_fs[level] = fs != 0;
    sg::drop_graphs(c);
    return SG_OK;
}

int sg_uniform_tile_info(sg_uniform *c, int level, int *cfg, int *ns, int *fs, float *ms, int nms)
{
    SG_REQUIRE(c && cfg && ns && fs && level >= 0 && level < c->nlev, "bad argument");
    *cfg = c->tile_cfg[level];
    *ns = c->tile_ns[level];
    *fs = c->tile_fs[level];
    for (int q = 0; ms && q < nms && q < sg::kNumCand; ++q) ms[q] = c->tune_ms[level][q];
    return SG_OK;
}

int sg_uniform_interp_dev(sg_uniform *c, int prec, void *uf, const void *uc, int nf, int cubic)
{
    SG_REQUIRE(c && uf && uc, "null argument");
    SG_REQUIRE(nf >= 2 && nf <= c->n && (nf & (nf - 1)) == 0, "bad nf");
    if (prec == SG_PREC_FP64)
        SG_TRY(sg::launch_interp<double>(c->stream, (double *)uf, (const double *)uc, nf, cubic != 0, false));
    else if (prec == SG_PREC_MIXED)
        SG_TRY(sg::launch_interp<float>(c->stream, (float *)uf, (const float *)uc, nf, cubic != 0, false));
    else SG_REQUIRE(false, "bad precision");
    SG_HIP(hipStreamSynchronize(c->stream));
    return SG_OK;
}

} // extern "C"
