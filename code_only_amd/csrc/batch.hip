// Batched uniform periodic solves: many independent mg_cycle problems (mg.F90:42-94), each its own
// n^3 grid (n = 2^a, or 3 * 2^a with the 3^3 coarsest level solved exactly) and h, advanced
// together with one launch per step of the multigrid schedule.
//
// Schedule.  Problems are aligned at their finest level (level 0), so "level l" means the same
// thing for every problem that has it.  Each step of mg_lin (mg.F90:106-218) becomes one launch
// over the problems the step concerns, with per-problem levels where a step acts on each problem's
// own coarsest level; every problem still sees exactly its own mg_lin order.  The schedule is
// recorded once per (precision, cycle shape): a first pass runs the host schedule without
// launching and stores every launch's item list, which is uploaded once; later passes launch with
// those device lists, eagerly or as a captured graph.  A problem that has converged is switched off
// through a device activity flag and drops out of every later launch.
//
//  * FP64 (plain): one kernel per reference loop nest, as the single-grid plain solver, with
//    general index arithmetic.  Each problem's values are bitwise those of the CPU restatement.
//  * Mixed: FP32 inner mg_lin with the fused LDS-tiled sweeps, tiled tricubic prolongation and the
//    single-workgroup coarse kernel (one workgroup per problem), tile shapes per level chosen by
//    timing the batched launches; FP64 outer residual and update (ping-pong phi buffers).  A
//    power-of-two problem gets bitwise the single-grid mixed solver's values.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "batch_kernels.h"
#include "sg_common.h"

using namespace sg;
using namespace sg::batch;

namespace {

struct Plan {
    bool valid = false;
    int ncycle = -1, npre = -1, npost = -1, coarse = -1;
    std::vector<Item> items;                // item lists of all launches, concatenated
    std::vector<std::pair<int, int>> recs;  // (offset, count) per launch, in schedule order
    size_t iter_begin = 0;                  // first record of the outer iteration
    Item *d_items = nullptr;
};

constexpr int kCubicFXs[3] = {32, 16, 8};

} // namespace

struct sg_batch {
    int np = 0, precisions = 0;
    std::vector<int> n, nlev, lc;
    std::vector<int> lb;   // bottom level of the exact-coarse cycle: lc, or a level of <= 64^3 cells
    std::vector<double> h;
    int maxlev = 0, maxlc = 0, maxlb = 0;
    std::vector<void *> allocs;
    std::vector<PL<double>> h_pl64;
    std::vector<PL<float>> h_pl32;
    std::vector<CoarseDesc> h_desc;
    PL<double> *d_pl64 = nullptr;
    PL<float> *d_pl32 = nullptr;
    CoarseDesc *d_desc = nullptr;
    double *d_q3 = nullptr, *d_lam3 = nullptr;  // exact 3^3 solve: basis, eigenvalues per problem
    Outer *h_outer = nullptr, *d_outer = nullptr;
    int *h_active = nullptr, *d_active = nullptr;
    double *d_pmax = nullptr, *d_psum = nullptr;
    long npart = 0;
    double *d_norm = nullptr, *h_norm = nullptr;    // [2p] max|r|, [2p+1] sum r^2
    double *d_snorm = nullptr, *h_snorm = nullptr;  // same for the source
    int tile_cfg[kBLev] = {}, tile_ns[kBLev] = {}, tile_fs[kBLev] = {}, cubic_fx[kBLev] = {};
    float tune_ms[kBLev][32] = {};
    bool tuned = false;
    Plan plan[2];  // [0] FP64, [1] mixed
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    int graph_plan_gen[2] = {-1, -1};
    int plan_gen = 0;
    // executor state
    Plan *cur = nullptr;
    bool recording = false;
    size_t cursor = 0;
    int ecode = CU;  // mixed: code of the FP32 buffer holding level 0's correction
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace {

using Pairs = std::vector<std::pair<int, int>>;  // (problem, level)

template <class F> Pairs members(const sg_batch *c, F pred_level)
{
    Pairs out;
    for (int p = 0; p < c->np; ++p) {
        const int l = pred_level(p);
        if (l >= 0) out.push_back({p, l});
    }
    return out;
}

long cells(int n) { return (long)n * n * n; }
int nblk_cells(long m) { return (int)((m + kThreads - 1) / kThreads); }

// Item list of one launch: `blocks(p, l)` workgroups per (problem, level); recorded on the first
// pass, taken from the device copy afterwards.  *d == nullptr / *cnt == 0: nothing to launch.
template <class FB>
int take(sg_batch *c, const Pairs &pl, FB blocks, const Item **d, int *cnt, int *total)
{
    std::vector<Item> lst;
    int tot = 0;
    for (auto &q : pl) {
        SG_REQUIRE(q.first >= 0 && q.first < c->np && q.second >= 0 && q.second < c->nlev[q.first],
                   "batch schedule: problem %d has no level %d", q.first, q.second);
        const int nb = blocks(q.first, q.second);
        lst.push_back(Item{q.first, q.second, tot, nb});
        tot += nb;
    }
    Plan &P = *c->cur;
    *d = nullptr;
    *cnt = 0;
    *total = tot;
    if (c->recording) {
        P.recs.push_back({(int)P.items.size(), (int)lst.size()});
        P.items.insert(P.items.end(), lst.begin(), lst.end());
        return SG_OK;
    }
    SG_REQUIRE(c->cursor < P.recs.size(), "batch schedule: more launches than recorded");
    const auto r = P.recs[c->cursor++];
    SG_REQUIRE(r.second == (int)lst.size(), "batch schedule changed between record and replay");
    if (r.second > 0) {
        *d = P.d_items + r.first;
        *cnt = r.second;
    }
    return SG_OK;
}

template <typename T> BL<T> bl(sg_batch *c, const Item *d, int cnt);
template <> BL<double> bl<double>(sg_batch *c, const Item *d, int cnt) { return {d, cnt, c->d_pl64, c->d_active}; }
template <> BL<float> bl<float>(sg_batch *c, const Item *d, int cnt) { return {d, cnt, c->d_pl32, c->d_active}; }

template <typename T> int pl_n(sg_batch *c, int p, int l)
{
    return std::is_same<T, double>::value ? c->h_pl64[p * kBLev + l].n : c->h_pl32[p * kBLev + l].n;
}

// ---- launches of the generic kernels --------------------------------------------------------

template <typename T> int op_relax(sg_batch *c, const Pairs &pl, int cu, int crhs, int ipass)
{
    const Item *d;
    int cnt, tot;
    SG_TRY(take(c, pl, [&](int p, int l) {
        const int n = pl_n<T>(c, p, l);
        return n == 1 ? 1 : nblk_cells(cells(n) / 2);
    }, &d, &cnt, &tot));
    if (!cnt) return SG_OK;
    hipLaunchKernelGGL(kb_relax<T>, dim3(tot), dim3(kThreads), 0, c->stream, bl<T>(c, d, cnt), cu, crhs, ipass);
    SG_CHECK_LAUNCH();
    return SG_OK;
}

// one mg_relax call (both colours) of each listed level
template <typename T> int op_relax2(sg_batch *c, const Pairs &pl, int cu, int crhs)
{
    SG_TRY(op_relax<T>(c, pl, cu, crhs, 1));
    return op_relax<T>(c, pl, cu, crhs, 2);
}

template <typename T> int op_resid(sg_batch *c, const Pairs &pl, int cres, int cu, int crhs)
{
    const Item *d;
    int cnt, tot;
    SG_TRY(take(c, pl, [&](int p, int l) { return nblk_cells(cells(pl_n<T>(c, p, l))); }, &d, &cnt, &tot));
    if (!cnt) return SG_OK;
    hipLaunchKernelGGL(kb_resid<T>, dim3(tot), dim3(kThreads), 0, c->stream, bl<T>(c, d, cnt), cres, cu, crhs);
    SG_CHECK_LAUNCH();
    return SG_OK;
}

// pairs name the coarse level
template <typename T> int op_restrict(sg_batch *c, const Pairs &pl, int cdst, int csrc)
{
    for (auto &q : pl) SG_REQUIRE(q.second >= 1, "batch schedule: restriction into level 0");
    const Item *d;
    int cnt, tot;
    SG_TRY(take(c, pl, [&](int p, int l) { return nblk_cells(cells(pl_n<T>(c, p, l))); }, &d, &cnt, &tot));
    if (!cnt) return SG_OK;
    hipLaunchKernelGGL(kb_restrict<T>, dim3(tot), dim3(kThreads), 0, c->stream, bl<T>(c, d, cnt), cdst, csrc);
    SG_CHECK_LAUNCH();
    return SG_OK;
}

template <typename T> int op_fill0(sg_batch *c, const Pairs &pl, int code)
{
    const Item *d;
    int cnt, tot;
    SG_TRY(take(c, pl, [&](int p, int l) { return nblk_cells(cells(pl_n<T>(c, p, l))); }, &d, &cnt, &tot));
    if (!cnt) return SG_OK;
    hipLaunchKernelGGL(kb_fill0<T>, dim3(tot), dim3(kThreads), 0, c->stream, bl<T>(c, d, cnt), code);
    SG_CHECK_LAUNCH();
    return SG_OK;
}

// mg_interp (mg.F90:608-633) of each listed fine level from the next coarser one: tricubic for
// the FMG step when the coarse grid plus ghosts is at least 4 wide, trilinear otherwise.  The
// tricubic levels that an LDS tile shape (x extent 32, 16 or 8, widest first; 8 x 8 in y, z)
// divides take the tiled form, the others the per-cell form (same values).
template <typename T> int op_interp(sg_batch *c, const Pairs &pl, int cdst, int csrc, bool cubic, bool add)
{
    Pairs pc, pln, pt[3];
    for (auto &q : pl) {
        SG_REQUIRE(q.second + 1 < c->nlev[q.first], "batch schedule: interpolation from a missing level");
        const int nc = pl_n<T>(c, q.first, q.second + 1), n = pl_n<T>(c, q.first, q.second);
        if (!(cubic && nc + 2 >= 4 && !add)) {
            (cubic && nc + 2 >= 4 ? pc : pln).push_back(q);
            continue;
        }
        int g = -1;
        for (int e = 0; e < 3 && g < 0; ++e)
            if (n >= kCubicFXs[e] && n % kCubicFXs[e] == 0 && n % 8 == 0) g = e;
        (g < 0 ? pc : pt[g]).push_back(q);
    }
    for (int e = 0; e < 3; ++e) {
        const int fx = kCubicFXs[e];
        const Item *d;
        int cnt, tot;
        SG_TRY(take(c, pt[e], [&](int p, int l) {
            const int n = pl_n<T>(c, p, l);
            return (n / fx) * (n / 8) * (n / 8);
        }, &d, &cnt, &tot));
        if (!cnt) continue;
        const BL<T> L = bl<T>(c, d, cnt);
        const dim3 grid(8 * ((tot + 7) / 8)), block(256);
        if (fx == 32)
            hipLaunchKernelGGL((kb_interp_cubic_tile_codes<32, 8, 8, 256, T>), grid, block, 0, c->stream, L, tot, cdst, csrc);
        else if (fx == 16)
            hipLaunchKernelGGL((kb_interp_cubic_tile_codes<16, 8, 8, 256, T>), grid, block, 0, c->stream, L, tot, cdst, csrc);
        else
            hipLaunchKernelGGL((kb_interp_cubic_tile_codes<8, 8, 8, 256, T>), grid, block, 0, c->stream, L, tot, cdst, csrc);
        SG_CHECK_LAUNCH();
    }
    for (int pass = 0; pass < 2; ++pass) {
        const Pairs &pp = pass == 0 ? pc : pln;
        const Item *d;
        int cnt, tot;
        SG_TRY(take(c, pp, [&](int p, int l) { return nblk_cells(cells(pl_n<T>(c, p, l))); }, &d, &cnt, &tot));
        if (!cnt) continue;
        const BL<T> L = bl<T>(c, d, cnt);
        if (pass == 0)
            hipLaunchKernelGGL((kb_interp<T, true, false>), dim3(tot), dim3(kThreads), 0, c->stream, L, cdst, csrc);
        else if (add)
            hipLaunchKernelGGL((kb_interp<T, false, true>), dim3(tot), dim3(kThreads), 0, c->stream, L, cdst, csrc);
        else
            hipLaunchKernelGGL((kb_interp<T, false, false>), dim3(tot), dim3(kThreads), 0, c->stream, L, cdst, csrc);
        SG_CHECK_LAUNCH();
    }
    return SG_OK;
}

// mg_slvsml (mg.F90:540-546) on each problem's coarsest level: two sweeps at 1^3 (only the second
// colour has a cell), the exact solve at 3^3.
int op_slvsml64(sg_batch *c, const Pairs &pl)
{
    Pairs p1, p3;
    for (auto &q : pl) (pl_n<double>(c, q.first, q.second) == 3 ? p3 : p1).push_back(q);
    for (int it = 0; it < 2; ++it) SG_TRY(op_relax<double>(c, p1, CU, CRHS, 2));
    const Item *d;
    int cnt, tot;
    SG_TRY(take(c, p3, [](int, int) { return 1; }, &d, &cnt, &tot));
    if (cnt) {
        hipLaunchKernelGGL(kb_exact3<double>, dim3(tot), dim3(64), 0, c->stream, bl<double>(c, d, cnt), CU, CRHS,
                           (const double *)c->d_q3, (const double *)c->d_lam3);
        SG_CHECK_LAUNCH();
    }
    return SG_OK;
}

// ---- FP64 plain mg_lin over the batch (mg.F90:106-218) ----------------------------------------

int plain_mg_lin(sg_batch *c, int ncycle, int npre, int npost)
{
    const int Lmax = c->maxlev - 1;
    auto at = [&](int l, int minlev) {  // problems with more than minlev levels, at level l
        return members(c, [&](int p) { return c->nlev[p] > minlev ? l : -1; });
    };
    for (int l = 1; l <= Lmax; ++l) SG_TRY(op_restrict<double>(c, at(l, l), CRHS, CRHS));
    const Pairs bottom = members(c, [&](int p) { return c->nlev[p] - 1; });
    SG_TRY(op_fill0<double>(c, bottom, CU));
    SG_TRY(op_slvsml64(c, bottom));
    for (int ml = Lmax - 1; ml >= 0; --ml) {
        const Pairs M = at(ml, ml + 1);
        const Pairs Mb = members(c, [&](int p) { return c->nlev[p] > ml + 1 ? c->nlev[p] - 1 : -1; });
        SG_TRY(op_interp<double>(c, M, CU, CU, true, false));
        for (int jc = 0; jc < ncycle; ++jc) {
            for (int v = ml; v <= Lmax - 1; ++v) {
                const Pairs Mv = at(v, v + 1), Mv1 = at(v + 1, v + 1);
                for (int q = 0; q < npre; ++q) SG_TRY(op_relax2<double>(c, Mv, CU, CRHS));
                SG_TRY(op_resid<double>(c, Mv, CRES, CU, CRHS));
                SG_TRY(op_restrict<double>(c, Mv1, CRHS, CRES));
                SG_TRY(op_fill0<double>(c, Mv1, CU));
            }
            SG_TRY(op_slvsml64(c, Mb));
            for (int v = Lmax - 1; v >= ml; --v) {
                const Pairs Mv = at(v, v + 1);
                SG_TRY(op_interp<double>(c, Mv, CU, CU, false, true));
                for (int q = 0; q < npost; ++q) SG_TRY(op_relax2<double>(c, Mv, CU, CRHS));
            }
        }
    }
    return SG_OK;
}

// ---- fused FP32 mg_lin over the batch ---------------------------------------------------------

template <class TL, int NS, bool FS>
void sweep_variant(sg_batch *c, const BL<float> &L, int tot, int cin, int cout, int cuc, bool zero, bool prolong,
                   bool resid)
{
    const dim3 grid(8 * ((tot + 7) / 8)), block(TL::NT);
    if (resid && zero)
        hipLaunchKernelGGL((kb_sweep<TL, NS, true, false, true, FS>), grid, block, 0, c->stream, L, tot, cin, cout, cuc);
    else if (resid)
        hipLaunchKernelGGL((kb_sweep<TL, NS, false, false, true, FS>), grid, block, 0, c->stream, L, tot, cin, cout, cuc);
    else if (prolong)
        hipLaunchKernelGGL((kb_sweep<TL, NS, false, true, false, FS>), grid, block, 0, c->stream, L, tot, cin, cout, cuc);
    else if (zero)
        hipLaunchKernelGGL((kb_sweep<TL, NS, true, false, false, FS>), grid, block, 0, c->stream, L, tot, cin, cout, cuc);
    else
        hipLaunchKernelGGL((kb_sweep<TL, NS, false, false, false, FS>), grid, block, 0, c->stream, L, tot, cin, cout, cuc);
}

template <class TL>
void sweep_ns(sg_batch *c, int ns, bool fs, const BL<float> &L, int tot, int cin, int cout, int cuc, bool zero,
              bool prolong, bool resid)
{
    if (ns == 2 && fs) {
        if constexpr (sweep_fits_lds<TL, 2, true>()) sweep_variant<TL, 2, true>(c, L, tot, cin, cout, cuc, zero, prolong, resid);
    } else if (ns == 2) {
        if constexpr (sweep_fits_lds<TL, 2, false>()) sweep_variant<TL, 2, false>(c, L, tot, cin, cout, cuc, zero, prolong, resid);
    } else if (fs) {
        if constexpr (sweep_fits_lds<TL, 1, true>()) sweep_variant<TL, 1, true>(c, L, tot, cin, cout, cuc, zero, prolong, resid);
    } else {
        sweep_variant<TL, 1, false>(c, L, tot, cin, cout, cuc, zero, prolong, resid);
    }
}

template <class SS>
void stream_variant(sg_batch *c, const BL<float> &L, int tot, int cin, int cout, int cuc, bool zero, bool prolong,
                    bool resid)
{
    const dim3 grid(8 * ((tot + 7) / 8)), block(SS::NT);
    if (resid && zero)
        hipLaunchKernelGGL((kb_stream<SS, true, false, true>), grid, block, 0, c->stream, L, tot, cin, cout, cuc);
    else if (resid)
        hipLaunchKernelGGL((kb_stream<SS, false, false, true>), grid, block, 0, c->stream, L, tot, cin, cout, cuc);
    else if (prolong)
        hipLaunchKernelGGL((kb_stream<SS, false, true, false>), grid, block, 0, c->stream, L, tot, cin, cout, cuc);
    else if (zero)
        hipLaunchKernelGGL((kb_stream<SS, true, false, false>), grid, block, 0, c->stream, L, tot, cin, cout, cuc);
    else
        hipLaunchKernelGGL((kb_stream<SS, false, false, false>), grid, block, 0, c->stream, L, tot, cin, cout, cuc);
}

// ns fused sweeps of level l of the listed problems with tile shape cfg: buffer cin -> cout;
// zero: start from u = 0; prolong: add the trilinear correction from level l + 1's buffer cuc;
// resid: restrict the residual into level l + 1's RHS.
int op_sweep_cfg(sg_batch *c, const Pairs &pl, int cfg, int ns, bool fs, int cin, int cout, int cuc, bool zero,
                 bool prolong, bool resid)
{
    SG_REQUIRE(cfg >= 0 && cfg < kNumCfg && cfg_ok(cfg, ns, fs), "bad sweep configuration %d/%d/%d", cfg, ns, (int)fs);
    const int *td = cfg < kNumTiles ? kTileDims[cfg] : kStreamDims[cfg - kNumTiles];
    for (auto &q : pl) {
        SG_REQUIRE(q.second < c->lc[q.first], "batch schedule: fused sweep on a coarse-kernel level");
        SG_REQUIRE(cfg_fits(cfg, c->h_pl32[q.first * kBLev + q.second].n), "sweep shape %d does not fit n = %d", cfg,
                   c->h_pl32[q.first * kBLev + q.second].n);
        SG_REQUIRE(c->h_pl32[q.first * kBLev + q.second].n > 2 * (2 * ns + 1), "tile halo wider than the grid");
    }
    const Item *d;
    int cnt, tot;
    SG_TRY(take(c, pl, [&](int p, int l) {
        const int n = c->h_pl32[p * kBLev + l].n;
        return (n / td[0]) * (n / td[1]) * (n / td[2]);
    }, &d, &cnt, &tot));
    if (!cnt) return SG_OK;
    const BL<float> L = bl<float>(c, d, cnt);
    switch (cfg) {
    case kNumTiles + 0: stream_variant<Stream0>(c, L, tot, cin, cout, cuc, zero, prolong, resid); break;
    case kNumTiles + 1: stream_variant<Stream1>(c, L, tot, cin, cout, cuc, zero, prolong, resid); break;
    case kNumTiles + 2: stream_variant<Stream2>(c, L, tot, cin, cout, cuc, zero, prolong, resid); break;
    case kNumTiles + 3: stream_variant<Stream3>(c, L, tot, cin, cout, cuc, zero, prolong, resid); break;
    case kNumTiles + 4: stream_variant<Stream4>(c, L, tot, cin, cout, cuc, zero, prolong, resid); break;
    case kNumTiles + 5: stream_variant<Stream5>(c, L, tot, cin, cout, cuc, zero, prolong, resid); break;
    case kNumTiles + 6: stream_variant<Stream6>(c, L, tot, cin, cout, cuc, zero, prolong, resid); break;
    case kNumTiles + 7: stream_variant<Stream7>(c, L, tot, cin, cout, cuc, zero, prolong, resid); break;
    case kNumTiles + 8: stream_variant<Stream8>(c, L, tot, cin, cout, cuc, zero, prolong, resid); break;
    case 0: sweep_ns<Tile0>(c, ns, fs, L, tot, cin, cout, cuc, zero, prolong, resid); break;
    case 1: sweep_ns<Tile1>(c, ns, fs, L, tot, cin, cout, cuc, zero, prolong, resid); break;
    case 2: sweep_ns<Tile2>(c, ns, fs, L, tot, cin, cout, cuc, zero, prolong, resid); break;
    case 3: sweep_ns<Tile3>(c, ns, fs, L, tot, cin, cout, cuc, zero, prolong, resid); break;
    default: sweep_ns<Tile4>(c, ns, fs, L, tot, cin, cout, cuc, zero, prolong, resid); break;
    }
    SG_CHECK_LAUNCH();
    return SG_OK;
}

int op_interp_cubic_tile(sg_batch *c, const Pairs &pl, int fx, int cuc)
{
    for (auto &q : pl) {
        const int n = c->h_pl32[q.first * kBLev + q.second].n;
        SG_REQUIRE(q.second < c->lc[q.first] && n >= fx && n % fx == 0 && n % 8 == 0,
                   "batch schedule: tricubic tile %d does not fit n = %d", fx, n);
    }
    const Item *d;
    int cnt, tot;
    SG_TRY(take(c, pl, [&](int p, int l) {
        const int n = c->h_pl32[p * kBLev + l].n;
        return (n / fx) * (n / 8) * (n / 8);
    }, &d, &cnt, &tot));
    if (!cnt) return SG_OK;
    const BL<float> L = bl<float>(c, d, cnt);
    const dim3 grid(8 * ((tot + 7) / 8)), block(256);
    if (fx == 32) hipLaunchKernelGGL((kb_interp_cubic_tile<32, 8, 8, 256>), grid, block, 0, c->stream, L, tot, cuc);
    else if (fx == 16) hipLaunchKernelGGL((kb_interp_cubic_tile<16, 8, 8, 256>), grid, block, 0, c->stream, L, tot, cuc);
    else hipLaunchKernelGGL((kb_interp_cubic_tile<8, 8, 8, 256>), grid, block, 0, c->stream, L, tot, cuc);
    SG_CHECK_LAUNCH();
    return SG_OK;
}

int op_coarse(sg_batch *c, const Pairs &pl, int mode, int ncycle, int npre, int npost, bool exact)
{
    for (auto &q : pl) SG_REQUIRE(q.second == c->lc[q.first], "batch schedule: coarse kernel off its level");
    const Item *d;
    int cnt, tot;
    SG_TRY(take(c, pl, [](int, int) { return 1; }, &d, &cnt, &tot));
    if (!cnt) return SG_OK;
    hipLaunchKernelGGL(kb_coarse, dim3(tot), dim3(kCoarseThreads), 0, c->stream, bl<float>(c, d, cnt),
                       (const CoarseDesc *)c->d_desc, mode, ncycle, npre, npost, exact ? 1 : 0);
    SG_CHECK_LAUNCH();
    return SG_OK;
}

// U := exact solution for RHS on each pair's level by the plane-parallel kernels, U1 as the
// intermediate array (the problems' bottom levels of more than 16^3 cells).
int op_exact_planes(sg_batch *c, const Pairs &pl)
{
    for (auto &q : pl) {
        const int n = c->h_pl32[q.first * kBLev + q.second].n;
        SG_REQUIRE(q.second == c->lb[q.first] && q.second != c->lc[q.first] && c->h_desc[q.first].qp &&
                       n <= kPlaneMaxN && n % 4 == 0,
                   "batch schedule: plane-parallel exact solve off its level");
    }
    const Item *d;
    int cnt, tot;
    auto planes = [&](int p, int l) { return c->h_pl32[p * kBLev + l].n; };
    const CoarseDesc *D = (const CoarseDesc *)c->d_desc;
    for (int step = 0; step < 3; ++step) {
        SG_TRY(take(c, pl, planes, &d, &cnt, &tot));
        if (!cnt) continue;
        const BL<float> L = bl<float>(c, d, cnt);
        const dim3 grid(8 * ((tot + 7) / 8)), block(kPlaneThreads);
        if (step == 0) hipLaunchKernelGGL(kb_exact_xy, grid, block, 0, c->stream, L, tot, D, CRHS, CU1);
        else if (step == 1) hipLaunchKernelGGL(kb_exact_z, grid, block, 0, c->stream, L, tot, D, CU1);
        else hipLaunchKernelGGL(kb_exact_xy, grid, block, 0, c->stream, L, tot, D, CU1, CU);
        SG_CHECK_LAUNCH();
    }
    return SG_OK;
}

// mg_lin in FP32 with the fused kernels (the single-grid fused mg_lin, stepped over all problems
// with the same fine level index at once).  c->ecode receives the buffer code that holds level 0's
// correction for problems with at least one fine level.  exact: the cycle of a problem with finer
// levels ends at its level lb, solved exactly (as the single-grid solver's SG_COARSE_EXACT).
int fast_mg_lin(sg_batch *c, int ncycle, int npre, int npost, bool exact)
{
    const std::vector<int> &bot = exact ? c->lb : c->lc;  // level of each problem's coarse-grid solve
    const int maxlc = exact ? c->maxlb : c->maxlc;
    auto fine_at = [&](int l, int minlc) {  // problems with bottom level > minlc, at level l
        return members(c, [&](int p) { return bot[p] > minlc ? l : -1; });
    };
    // coarse-grid solve of the problems selected by pred: k_coarse, or the plane-parallel kernels
    auto bottom = [&](auto pred, int mode) -> int {
        auto planes = [&](int p) { return exact && c->lb[p] != c->lc[p]; };
        SG_TRY(op_coarse(c, members(c, [&](int p) { return pred(p) && !planes(p) ? c->lc[p] : -1; }), mode, ncycle,
                         npre, npost, exact));
        if (exact) SG_TRY(op_exact_planes(c, members(c, [&](int p) { return pred(p) && planes(p) ? c->lb[p] : -1; })));
        return SG_OK;
    };
    for (int l = 1; l <= maxlc; ++l)
        SG_TRY(op_restrict<float>(c, members(c, [&](int p) { return bot[p] >= l ? l : -1; }), CRHS, CRHS));
    SG_TRY(bottom([&](int p) { return bot[p] == 0; }, 0));
    int cur[kBLev] = {};
    auto code = [](int b) { return b ? CU1 : CU; };
    for (int ml = maxlc - 1; ml >= 0; --ml) {
        SG_TRY(bottom([&](int p) { return bot[p] == ml + 1; }, 0));
        const Pairs M = fine_at(ml, ml);
        SG_TRY(op_interp_cubic_tile(c, M, c->cubic_fx[ml], code(cur[ml + 1])));
        cur[ml] = 0;
        for (int jc = 0; jc < ncycle; ++jc) {
            for (int v = ml; v <= maxlc - 1; ++v) {
                const Pairs Mv = fine_at(v, v);
                for (int r = npre, first = 1; r > 0; first = 0) {
                    const int k = std::min(r, c->tile_ns[v]);
                    const bool last = r == k;
                    SG_TRY(op_sweep_cfg(c, Mv, c->tile_cfg[v], k, c->tile_fs[v] != 0, code(cur[v]), code(1 - cur[v]),
                                        CU, first && v > ml, false, last));
                    cur[v] ^= 1;
                    r -= k;
                }
            }
            SG_TRY(bottom([&](int p) { return bot[p] > ml; }, 1));
            for (int v = maxlc - 1; v >= ml; --v) {
                const Pairs Mv = fine_at(v, v);
                for (int r = npost, first = 1; r > 0; first = 0) {
                    const int k = std::min(r, c->tile_ns[v]);
                    SG_TRY(op_sweep_cfg(c, Mv, c->tile_cfg[v], k, c->tile_fs[v] != 0, code(cur[v]), code(1 - cur[v]),
                                        code(cur[v + 1]), false, first, false));
                    cur[v] ^= 1;
                    r -= k;
                }
            }
        }
    }
    c->ecode = code(cur[0]);
    return SG_OK;
}

// ---- outer FP64 steps -------------------------------------------------------------------------

int op_src_norms(sg_batch *c)
{
    const Pairs all = members(c, [](int) { return 0; });
    const Item *d;
    int cnt, tot;
    SG_TRY(take(c, all, [&](int p, int) { return nblk_cells(cells(c->n[p])); }, &d, &cnt, &tot));
    if (!cnt) return SG_OK;
    hipLaunchKernelGGL(kb_src_norms, dim3(tot), dim3(kThreads), 0, c->stream, bl<double>(c, d, cnt),
                       (const Outer *)c->d_outer, c->d_pmax, c->d_psum);
    SG_CHECK_LAUNCH();
    hipLaunchKernelGGL(kb_norm_final, dim3(cnt), dim3(1024), 0, c->stream, d, (const int *)c->d_active,
                       (const double *)c->d_pmax, (const double *)c->d_psum, c->d_snorm);
    SG_CHECK_LAUNCH();
    return SG_OK;
}

// r = f - L(pin [+ e]) into level 0's RHS (FP64 or FP32), pout = pin + e, and the norms of r.
template <typename Te> int op_outer_resid(sg_batch *c, int parity, bool upd, int ce)
{
    const Pairs all = members(c, [](int) { return 0; });
    // the z-marching form when every problem's n is a multiple of 32 (chunks of 16 planes)
    bool zform = true;
    for (auto &q : all) zform = zform && c->n[q.first] % kOzX == 0 && c->n[q.first] % 16 == 0;
    const int zc = 16;
    const Item *d;
    int cnt, tot;
    SG_TRY(take(c, all, [&](int p, int) {
        const int n = c->n[p];
        return zform ? (n / kOzX) * (n / kOzY) * (n / zc) : nblk_cells(cells(n));
    }, &d, &cnt, &tot));
    if (!cnt) return SG_OK;
    const BL<Te> L = bl<Te>(c, d, cnt);
    const Outer *O = (const Outer *)c->d_outer;
    const dim3 zgrid(8 * ((tot + 7) / 8)), zblock(kOzX * kOzY);
    if (zform && upd)
        hipLaunchKernelGGL((kb_outer_resid_z<Te, true>), zgrid, zblock, 0, c->stream, L, tot, O, parity, ce, zc,
                           c->d_pmax, c->d_psum);
    else if (zform)
        hipLaunchKernelGGL((kb_outer_resid_z<Te, false>), zgrid, zblock, 0, c->stream, L, tot, O, parity, ce, zc,
                           c->d_pmax, c->d_psum);
    else if (upd)
        hipLaunchKernelGGL((kb_outer_resid<Te, true>), dim3(tot), dim3(kThreads), 0, c->stream, L, O, parity, ce,
                           c->d_pmax, c->d_psum);
    else
        hipLaunchKernelGGL((kb_outer_resid<Te, false>), dim3(tot), dim3(kThreads), 0, c->stream, L, O, parity, ce,
                           c->d_pmax, c->d_psum);
    SG_CHECK_LAUNCH();
    hipLaunchKernelGGL(kb_norm_final, dim3(cnt), dim3(1024), 0, c->stream, d, (const int *)c->d_active,
                       (const double *)c->d_pmax, (const double *)c->d_psum, c->d_norm);
    SG_CHECK_LAUNCH();
    return SG_OK;
}

// The whole schedule of a solve: [source norms, first residual] then one outer iteration, whose
// launches are the same every iteration apart from the phi-buffer parity (a kernel argument).
int run_init(sg_batch *c, bool mixed)
{
    SG_TRY(op_src_norms(c));
    return mixed ? op_outer_resid<float>(c, 0, false, CU) : op_outer_resid<double>(c, 0, false, CU);
}

int run_iteration(sg_batch *c, bool mixed, const sg_params *p, int parity)
{
    if (mixed) {
        SG_TRY(fast_mg_lin(c, p->ncycle, p->npre, p->npost, p->coarse == SG_COARSE_EXACT));
        return op_outer_resid<float>(c, parity, true, c->ecode);
    }
    SG_TRY(plain_mg_lin(c, p->ncycle, p->npre, p->npost));
    return op_outer_resid<double>(c, parity, true, CU);
}

// Record the launch sequence of (precision, cycle shape) once and upload its item lists.
int ensure_plan(sg_batch *c, bool mixed, const sg_params *p)
{
    Plan &P = c->plan[mixed ? 1 : 0];
    if (P.valid && P.ncycle == p->ncycle && P.npre == p->npre && P.npost == p->npost && P.coarse == p->coarse)
        return SG_OK;
    if (P.d_items) (void)hipFree(P.d_items);
    P = Plan();
    c->cur = &P;
    c->recording = true;
    int rc = run_init(c, mixed);
    P.iter_begin = P.recs.size();
    if (rc == SG_OK) rc = run_iteration(c, mixed, p, 0);
    c->recording = false;
    SG_TRY(rc);
    if (!P.items.empty()) {
        SG_HIP(hipMalloc(&P.d_items, sizeof(Item) * P.items.size()));
        SG_HIP(hipMemcpy(P.d_items, P.items.data(), sizeof(Item) * P.items.size(), hipMemcpyHostToDevice));
    }
    P.ncycle = p->ncycle;
    P.npre = p->npre;
    P.npost = p->npost;
    P.coarse = p->coarse;
    P.valid = true;
    ++c->plan_gen;
    return SG_OK;
}

int replay_iteration(sg_batch *c, bool mixed, const sg_params *p, int parity)
{
    Plan &P = c->plan[mixed ? 1 : 0];
    c->cur = &P;
    if (!(mixed && p->use_graph)) {
        c->cursor = P.iter_begin;
        return run_iteration(c, mixed, p, parity);
    }
    if (!c->graph[parity] || c->graph_plan_gen[parity] != c->plan_gen) {
        if (c->graph[parity]) (void)hipGraphExecDestroy(c->graph[parity]);
        c->graph[parity] = nullptr;
        hipGraph_t g = nullptr;
        c->cursor = P.iter_begin;
        SG_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int rc = run_iteration(c, mixed, p, parity);
        const hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc != SG_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        SG_HIP(e);
        const hipError_t e2 = hipGraphInstantiate(&c->graph[parity], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        SG_HIP(e2);
        c->graph_plan_gen[parity] = c->plan_gen;
    }
    SG_HIP(hipGraphLaunch(c->graph[parity], c->stream));
    return SG_OK;
}

// ---- autotuner: tile shape, sweeps per launch and source staging per fine level --------------

int autotune(sg_batch *c)
{
    if (c->tuned) return SG_OK;
    Plan tmp;
    c->cur = &tmp;
    hipEvent_t a = nullptr, b = nullptr;
    SG_HIP(hipEventCreate(&a));
    SG_HIP(hipEventCreate(&b));
    int rc = SG_OK;
    for (int v = 0; v < c->maxlc && rc == SG_OK; ++v) {
        const Pairs Mv = members(c, [&](int p) { return c->lc[p] > v ? v : -1; });
        // tricubic tile: the widest x extent that divides every member's level
        c->cubic_fx[v] = 0;
        for (int fx : kCubicFXs) {
            bool ok = true;
            for (auto &q : Mv) {
                const int n = c->h_pl32[q.first * kBLev + v].n;
                ok = ok && n >= fx && n % fx == 0 && n % 8 == 0;
            }
            if (ok) { c->cubic_fx[v] = fx; break; }
        }
        if (c->cubic_fx[v] == 0) { set_error("no tricubic tile fits level %d", v); rc = SG_ERR_ARG; break; }
        c->tile_cfg[v] = -1;
        float best = 1e30f;
        for (int cand = 0; cand < kNumCand && rc == SG_OK; ++cand) {
            int cfg, ns;
            bool fs;
            cand_decode(cand, cfg, ns, fs);
            bool fits = cfg_ok(cfg, ns, fs);
            for (auto &q : Mv) fits = fits && cfg_fits(cfg, c->h_pl32[q.first * kBLev + v].n);
            c->tune_ms[v][cand] = -1.f;
            if (!fits) continue;
            // one fine-level visit of a V(2,2) cycle in the launches the cycle makes: two sweeps
            // ending with the residual restriction, two starting with the prolongated correction.
            // Record its item lists into a scratch plan, upload, time it.
            int flip = 0;
            auto visit = [&]() -> int {
                auto one = [&](int k, bool prolong, bool resid) {
                    const int r = op_sweep_cfg(c, Mv, cfg, k, fs, flip ? CU1 : CU, flip ? CU : CU1, CU, false, prolong,
                                               resid);
                    flip ^= 1;
                    return r;
                };
                if (ns == 2) {
                    SG_TRY(one(2, false, true));
                    return one(2, true, false);
                }
                SG_TRY(one(1, false, false));
                SG_TRY(one(1, false, true));
                SG_TRY(one(1, true, false));
                return one(1, false, false);
            };
            tmp = Plan();
            c->recording = true;
            rc = visit();
            c->recording = false;
            if (rc != SG_OK) break;
            if (hipMalloc(&tmp.d_items, sizeof(Item) * tmp.items.size()) != hipSuccess ||
                hipMemcpy(tmp.d_items, tmp.items.data(), sizeof(Item) * tmp.items.size(), hipMemcpyHostToDevice) !=
                    hipSuccess) {
                rc = SG_ERR_HIP;
                break;
            }
            const int reps = 3;  // timed visits, after one untimed
            for (int r = 0; r < reps + 1 && rc == SG_OK; ++r) {
                if (r == 1 && hipEventRecord(a, c->stream) != hipSuccess) rc = SG_ERR_HIP;
                c->cursor = 0;
                if (rc == SG_OK) rc = visit();
            }
            float ms = 0.f;
            if (rc == SG_OK && (hipEventRecord(b, c->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
                                hipEventElapsedTime(&ms, a, b) != hipSuccess))
                rc = SG_ERR_HIP;
            (void)hipFree(tmp.d_items);
            tmp.d_items = nullptr;
            if (rc != SG_OK) break;
            ms /= reps;
            c->tune_ms[v][cand] = ms;
            if (ms < best) {
                best = ms;
                c->tile_cfg[v] = cfg;
                c->tile_ns[v] = ns;
                c->tile_fs[v] = fs;
            }
        }
        if (rc == SG_OK && c->tile_cfg[v] < 0) { set_error("no sweep tile fits level %d", v); rc = SG_ERR_ARG; }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    c->cur = nullptr;
    if (rc == SG_ERR_HIP) set_error("batch autotune: HIP call failed");
    if (rc == SG_OK) c->tuned = true;
    return rc;
}

// ---- solve ------------------------------------------------------------------------------------

int read_norms(sg_batch *c, double *dev, double *host)
{
    SG_HIP(hipMemcpyAsync(host, dev, sizeof(double) * 2 * c->np, hipMemcpyDeviceToHost, c->stream));
    SG_HIP(hipStreamSynchronize(c->stream));
    return SG_OK;
}

int upload_active(sg_batch *c)
{
    SG_HIP(hipMemcpyAsync(c->d_active, c->h_active, sizeof(int) * c->np, hipMemcpyHostToDevice, c->stream));
    return SG_OK;
}

int solve(sg_batch *c, double *const *psi, const double *const *src, const sg_params *p, sg_stats *st)
{
    SG_REQUIRE(c && psi && src && p && st, "null argument");
    SG_REQUIRE(p->precision == SG_PREC_FP64 || p->precision == SG_PREC_MIXED, "bad precision %d", p->precision);
    SG_REQUIRE(c->precisions & p->precision, "context was not created with precision %d", p->precision);
    SG_REQUIRE(p->stop_mode == SG_STOP_TRUE_RESIDUAL, "the batched solve stops on the true residual only");
    SG_REQUIRE(p->maxit >= 0 && p->maxit <= 1000, "bad maxit %d", p->maxit);
    SG_REQUIRE(p->tol >= 0.0, "bad tol");
    SG_REQUIRE(p->ncycle >= 1 && p->ncycle <= 64 && p->npre >= 1 && p->npre <= 64 && p->npost >= 1 && p->npost <= 64,
               "bad cycle shape");
    SG_REQUIRE(p->coarse == SG_COARSE_EXACT || p->coarse == SG_COARSE_VCYCLE, "bad coarse %d", p->coarse);
    const bool mixed = p->precision == SG_PREC_MIXED;
    for (int q = 0; q < c->np; ++q) SG_REQUIRE(psi[q] && src[q], "null array of problem %d", q);
    if (mixed) SG_TRY(autotune(c));
    SG_TRY(ensure_plan(c, mixed, p));
    std::memset(st, 0, sizeof(sg_stats) * c->np);
    for (int q = 0; q < c->np; ++q) {
        c->h_outer->psi[q] = psi[q];
        c->h_outer->src[q] = src[q];
        c->h_active[q] = 1;
    }
    SG_HIP(hipEventRecord(c->ev0, c->stream));
    SG_HIP(hipMemcpyAsync(c->d_outer, c->h_outer, sizeof(Outer), hipMemcpyHostToDevice, c->stream));
    SG_TRY(upload_active(c));
    c->cur = &c->plan[mixed ? 1 : 0];
    c->cursor = 0;
    SG_TRY(run_init(c, mixed));
    SG_TRY(read_norms(c, c->d_snorm, c->h_snorm));
    SG_TRY(read_norms(c, c->d_norm, c->h_norm));
    std::vector<int> onalt(c->np, 0), done(c->np, 0);
    int parity = 0;
    for (int it = 0;; ++it) {
        int nactive = 0;
        for (int q = 0; q < c->np; ++q) {
            if (done[q]) continue;
            const double fmax = c->h_snorm[2 * q], fl2 = std::sqrt(c->h_snorm[2 * q + 1]);
            const double rmax = c->h_norm[2 * q], rl2 = std::sqrt(c->h_norm[2 * q + 1]);
            sg_stats &s = st[q];
            s.iterations = it;
            s.rel_linf = fmax > 0 ? rmax / fmax : 0.0;
            s.rel_l2 = fl2 > 0 ? rl2 / fl2 : 0.0;
            s.rel_h2 = s.rel_linf;   // one cell width per problem: the h^2 weights cancel
            if (it < SG_HIST) s.hist[it] = s.rel_linf;
            if (!std::isfinite(rmax) || !std::isfinite(rl2)) { s.status = SG_ERR_NONFINITE; done[q] = 1; }
            else if (s.rel_linf < p->tol) { s.status = SG_OK; done[q] = 1; }
            else if (it >= p->maxit) { s.status = SG_ERR_NOT_CONVERGED; done[q] = 1; }
            c->h_active[q] = !done[q];
            nactive += !done[q];
        }
        if (!nactive) break;
        SG_TRY(upload_active(c));
        SG_TRY(replay_iteration(c, mixed, p, parity));
        for (int q = 0; q < c->np; ++q)
            if (!done[q]) onalt[q] ^= 1;
        parity ^= 1;
        SG_TRY(read_norms(c, c->d_norm, c->h_norm));
    }
    for (int q = 0; q < c->np; ++q)
        if (onalt[q])
            SG_HIP(hipMemcpyAsync(psi[q], c->h_outer->alt[q], sizeof(double) * cells(c->n[q]), hipMemcpyDeviceToDevice,
                                  c->stream));
    SG_HIP(hipEventRecord(c->ev1, c->stream));
    SG_HIP(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    SG_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    int worst = SG_OK;
    for (int q = 0; q < c->np; ++q) {
        st[q].ms = ms;
        if (st[q].status != SG_OK) worst = st[q].status;
    }
    return worst;
}

// ---- context ------------------------------------------------------------------------------------

template <typename T> int dev_alloc(sg_batch *c, T **p, size_t count)
{
    SG_HIP(hipMalloc(p, sizeof(T) * std::max<size_t>(count, 1)));
    c->allocs.push_back(*p);
    return SG_OK;
}

void destroy(sg_batch *c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &g : c->graph)
        if (g) (void)hipGraphExecDestroy(g);
    for (auto &P : c->plan)
        if (P.d_items) (void)hipFree(P.d_items);
    for (void *p : c->allocs) (void)hipFree(p);
    if (c->h_outer) (void)hipHostFree(c->h_outer);
    if (c->h_active) (void)hipHostFree(c->h_active);
    if (c->h_norm) (void)hipHostFree(c->h_norm);
    if (c->h_snorm) (void)hipHostFree(c->h_snorm);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int create(sg_batch **out, int np, const int *ns, const double *hs, int precisions)
{
    SG_REQUIRE(out, "null context pointer");
    *out = nullptr;
    SG_REQUIRE(np >= 1 && np <= kMaxProb, "number of problems must be in [1, %d], got %d", kMaxProb, np);
    SG_REQUIRE(ns && hs, "null size or spacing array");
    SG_REQUIRE(precisions > 0 && (precisions & ~(SG_PREC_FP64 | SG_PREC_MIXED)) == 0, "bad precisions %d", precisions);
    for (int q = 0; q < np; ++q) {
        int m = ns[q], lev = 0;
        SG_REQUIRE(m >= 2 && m <= 1024, "problem %d: n = %d out of range [2, 1024]", q, m);
        while (m % 2 == 0) { m /= 2; ++lev; }
        SG_REQUIRE(m == 1 || m == 3, "problem %d: n = %d is neither 2^a nor 3 * 2^a", q, ns[q]);
        SG_REQUIRE(lev + 1 <= kBLev, "problem %d: too many levels", q);
        SG_REQUIRE(hs[q] > 0.0 && std::isfinite(hs[q]), "problem %d: bad h", q);
    }
    sg_batch *c = new sg_batch();
    c->np = np;
    c->precisions = precisions;
    c->n.assign(ns, ns + np);
    c->h.assign(hs, hs + np);
    c->nlev.resize(np);
    c->lc.resize(np);
    c->lb.resize(np);
    c->h_pl64.assign((size_t)np * kBLev, PL<double>{});
    c->h_pl32.assign((size_t)np * kBLev, PL<float>{});
    c->h_desc.assign(np, CoarseDesc{});
    std::vector<double> lam3(3 * np, 0.0), q3(9, 0.0), lamtmp(3);
    hartley_basis(3, 1.0, q3.data(), lamtmp.data());
    auto fail = [&](int rc) { destroy(c); return rc; };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        set_error("stream/event creation failed");
        return fail(SG_ERR_HIP);
    }
    long part = 0;
    for (int q = 0; q < np; ++q) {
        int nl = 0;
        while (true) {
            const int m = ns[q] >> nl;
            ++nl;
            if (m % 2 != 0) break;
        }
        c->nlev[q] = nl;
        c->maxlev = std::max(c->maxlev, nl);
        int lc = 0;
        while (lc < nl - 1 && (ns[q] >> lc) > kCoarseMaxN) ++lc;
        c->lc[q] = lc;
        c->maxlc = std::max(c->maxlc, lc);
        // exact-coarse cycle: ends at the first level below level 0 with <= 64^3 cells when that
        // has more than 16^3 (plane-parallel solve), else at lc (k_coarse)
        int lbq = lc;
        if (lc > 0) {
            int l = 1;
            while ((ns[q] >> l) > kPlaneMaxN) ++l;
            if ((ns[q] >> l) > kCoarseMaxN) lbq = l;
        }
        c->lb[q] = lbq;
        c->maxlb = std::max(c->maxlb, lbq);
        part += nblk_cells(cells(ns[q]));
        for (int l = 0; l < nl; ++l) {
            const int m = ns[q] >> l;
            const double hl = hs[q] * (double)(1 << l);
            PL<double> &a = c->h_pl64[q * kBLev + l];
            PL<float> &b = c->h_pl32[q * kBLev + l];
            a.n = b.n = m;
            a.lc = b.lc = lc;
            a.dn = b.dn = make_div((unsigned)m);
            a.dnh = b.dnh = make_div((unsigned)std::max(m / 2, 1));
            a.h2 = hl * hl;
            a.h2i = 1.0 / (hl * hl);
            b.h2 = (float)(hl * hl);
            b.h2i = (float)(1.0 / (hl * hl));
            const size_t len = (size_t)cells(m);
            int rc = SG_OK;
            if (precisions & SG_PREC_FP64) {
                double *u, *rhs, *res;
                if ((rc = dev_alloc(c, &u, len)) || (rc = dev_alloc(c, &rhs, len)) || (rc = dev_alloc(c, &res, len)))
                    return fail(rc);
                a.a[CU] = u; a.a[CRHS] = rhs; a.a[CRES] = res; a.a[CU1] = nullptr;
            }
            if (precisions & SG_PREC_MIXED) {
                float *u, *u1, *rhs;
                if ((rc = dev_alloc(c, &u, len)) || (rc = dev_alloc(c, &u1, len)) || (rc = dev_alloc(c, &rhs, len)))
                    return fail(rc);
                b.a[CU] = u; b.a[CU1] = u1; b.a[CRHS] = rhs; b.a[CRES] = nullptr;
            }
            if (m == 3) {
                double Qd[9], ld[3];
                hartley_basis(3, hl, Qd, ld);
                for (int e = 0; e < 3; ++e) lam3[3 * q + e] = ld[e];
            }
        }
        // coarse kernel levels: lc .. nlev-1
        CoarseDesc &D = c->h_desc[q];
        D.nlev = nl - lc;
        int off = 0;
        for (int k = 0; k < D.nlev; ++k) {
            const int l = lc + k, m = ns[q] >> l;
            const double hl = hs[q] * (double)(1 << l);
            D.n[k] = m;
            D.h2[k] = (float)(hl * hl);
            D.h2i[k] = (float)(1.0 / (hl * hl));
            D.off[k] = off;
            D.dn[k] = make_div((unsigned)m);
            D.dnh[k] = make_div((unsigned)std::max(m / 2, 1));
            off += m * m * m;
            if (m == 3) {
                double Qd[9], ld[3];
                hartley_basis(3, hl, Qd, ld);
                for (int e = 0; e < 9; ++e) D.q3[e] = (float)Qd[e];
                for (int e = 0; e < 3; ++e) D.lam3[e] = (float)ld[e];
            }
        }
        // exact solve of held level 0 when the problem has finer levels (SG_COARSE_EXACT)
        D.exact = lc > 0 ? 1 : 0;
        if (D.exact) {
            const int m = ns[q] >> lc;
            if (m > kCoarseMaxN) {
                set_error("problem %d: first coarse level %d^3 exceeds the exact solve's size", q, m);
                return fail(SG_ERR_ARG);
            }
            double Qd[kCoarseMaxN * kCoarseMaxN], ld[kCoarseMaxN];
            hartley_basis(m, hs[q] * (double)(1 << lc), Qd, ld);
            for (int e = 0; e < m * m; ++e) D.qx[e] = (float)Qd[e];
            for (int e = 0; e < m; ++e) D.lamx[e] = (float)ld[e];
        }
        D.qp = D.invp = nullptr;
        if (c->lb[q] != lc && (precisions & SG_PREC_MIXED)) {
            const int m = ns[q] >> c->lb[q];
            std::vector<double> Qd((size_t)m * m), ld(m);
            hartley_basis(m, hs[q] * (double)(1 << c->lb[q]), Qd.data(), ld.data());
            std::vector<float> Qf(Qd.begin(), Qd.end()), inv((size_t)m * m * m);
            exact_inverse_table(m, ld.data(), inv.data());
            float *qp = nullptr, *invp = nullptr;
            int rc = SG_OK;
            if ((rc = dev_alloc(c, &qp, Qf.size())) || (rc = dev_alloc(c, &invp, inv.size()))) return fail(rc);
            if (hipMemcpy(qp, Qf.data(), sizeof(float) * Qf.size(), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(invp, inv.data(), sizeof(float) * inv.size(), hipMemcpyHostToDevice) != hipSuccess) {
                set_error("exact-solve table upload failed");
                return fail(SG_ERR_HIP);
            }
            D.qp = qp;
            D.invp = invp;
        }
        if (off > kCoarseCellsB || D.nlev > 8) {
            set_error("problem %d: coarse levels exceed the coarse kernel's capacity", q);
            return fail(SG_ERR_ARG);
        }
    }
    int rc = SG_OK;
    if ((rc = dev_alloc(c, &c->d_pl64, c->h_pl64.size())) || (rc = dev_alloc(c, &c->d_pl32, c->h_pl32.size())) ||
        (rc = dev_alloc(c, &c->d_desc, c->h_desc.size())) || (rc = dev_alloc(c, &c->d_q3, 9)) ||
        (rc = dev_alloc(c, &c->d_lam3, lam3.size())) || (rc = dev_alloc(c, &c->d_outer, 1)) ||
        (rc = dev_alloc(c, &c->d_active, np)) || (rc = dev_alloc(c, &c->d_pmax, part)) ||
        (rc = dev_alloc(c, &c->d_psum, part)) || (rc = dev_alloc(c, &c->d_norm, 2 * np)) ||
        (rc = dev_alloc(c, &c->d_snorm, 2 * np)))
        return fail(rc);
    c->npart = part;
    if (hipHostMalloc(&c->h_outer, sizeof(Outer)) != hipSuccess || hipHostMalloc(&c->h_active, sizeof(int) * np) != hipSuccess ||
        hipHostMalloc(&c->h_norm, sizeof(double) * 2 * np) != hipSuccess ||
        hipHostMalloc(&c->h_snorm, sizeof(double) * 2 * np) != hipSuccess) {
        set_error("pinned host allocation failed");
        return fail(SG_ERR_HIP);
    }
    std::memset(c->h_outer, 0, sizeof(Outer));
    for (int q = 0; q < np; ++q) {
        double *alt;
        if ((rc = dev_alloc(c, &alt, (size_t)cells(ns[q])))) return fail(rc);
        c->h_outer->alt[q] = alt;
        c->h_outer->h2i[q] = 1.0 / (hs[q] * hs[q]);
        c->h_active[q] = 1;
    }
    if (hipMemcpy(c->d_pl64, c->h_pl64.data(), sizeof(PL<double>) * c->h_pl64.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->d_pl32, c->h_pl32.data(), sizeof(PL<float>) * c->h_pl32.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->d_desc, c->h_desc.data(), sizeof(CoarseDesc) * c->h_desc.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->d_q3, q3.data(), sizeof(double) * 9, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->d_lam3, lam3.data(), sizeof(double) * lam3.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->d_outer, c->h_outer, sizeof(Outer), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->d_active, c->h_active, sizeof(int) * np, hipMemcpyHostToDevice) != hipSuccess) {
        set_error("table upload failed");
        return fail(SG_ERR_HIP);
    }
    // zero the level arrays: a problem's coarsest U is read before any kernel writes it only by
    // the (zero-filled) interpolation of the first FMG step, but keep every array defined
    for (size_t q = 0; q < c->h_pl64.size(); ++q)
        for (int code = 0; code < kCodes; ++code) {
            const int m = c->h_pl64[q].n;
            if (!m) continue;
            if (c->h_pl64[q].a[code] && hipMemset(c->h_pl64[q].a[code], 0, sizeof(double) * cells(m)) != hipSuccess)
                return fail((set_error("memset failed"), SG_ERR_HIP));
            if (c->h_pl32[q].a[code] && hipMemset(c->h_pl32[q].a[code], 0, sizeof(float) * cells(m)) != hipSuccess)
                return fail((set_error("memset failed"), SG_ERR_HIP));
        }
    *out = c;
    return SG_OK;
}

} // namespace

extern "C" {

int sg_batch_create(sg_batch **ctx, int nprob, const int *n, const double *h, int precisions)
{
    return create(ctx, nprob, n, h, precisions);
}

int sg_batch_destroy(sg_batch *ctx)
{
    destroy(ctx);
    return SG_OK;
}

int sg_batch_solve_dev(sg_batch *ctx, double *const *psi, const double *const *src, const sg_params *p,
                       sg_stats *st)
{
    return solve(ctx, psi, src, p, st);
}

int sg_batch_tile_info(sg_batch *c, int level, int *cfg, int *ns, int *fs, int *cubic_fx, float *ms, int nms)
{
    SG_REQUIRE(c && level >= 0 && level < kBLev, "bad context or level");
    SG_REQUIRE(c->tuned, "tiles are chosen on the first mixed-precision solve");
    if (cfg) *cfg = level < c->maxlc ? c->tile_cfg[level] : -1;
    if (ns) *ns = c->tile_ns[level];
    if (fs) *fs = c->tile_fs[level];
    if (cubic_fx) *cubic_fx = c->cubic_fx[level];
    if (ms)
        for (int q = 0; q < nms && q < kNumCand; ++q) ms[q] = c->tune_ms[level][q];
    return SG_OK;
}

int sg_batch_levels(sg_batch *c, int prob, int *nlev, int *first_coarse)
{
    SG_REQUIRE(c && prob >= 0 && prob < c->np, "bad context or problem");
    if (nlev) *nlev = c->nlev[prob];
    if (first_coarse) *first_coarse = c->lc[prob];
    return SG_OK;
}

} // extern "C"
