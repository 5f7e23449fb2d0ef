// Block-structured AMR self-gravity solve on one MI355X: host orchestration of fmg_cycle /
// fmg_lin (fmg.F90:134-377) with the FAS coarse solve vmg_fas (vmg.F90:116-243) and the base-grid
// mg_cycle (mg.F90:42-94), over device-resident block arrays, and the AMR part of the C ABI.
//
// The sequence of operations is the reference's, operation for operation; each operation is one
// kernel over all blocks it applies to (amr_kernels.h), so FP64 results equal a sequential
// execution of the reference's loops in the same operand order.  Block tables (levels, parents,
// children, neighbours and the lists derived from them) are built once per hierarchy at create
// time; the reference rebuilds them on every call (fmg.F90:155-160).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <type_traits>
#include <vector>

#include "amr_kernels.h"
#include "amr_adjoint.h"
#include "base_direct.h"
#include "base_mg.h"
#include "blockmem_kernels.h"
#include "sg_common.h"

namespace sg {
const char *last_error();
}

using namespace sg;
using namespace sg::amr;

namespace {

constexpr int kNfmg = 3;          // FMG levels with 8^3 blocks: 8^3, 4^3, 2^3 (fmg_data.F90:115)
constexpr int kMaxLev = 16;       // AMR levels
constexpr int kThreads = 256;
// launches of go() over fewer items use half-size workgroups: the FMG-level-1 and -2 passes of a
// 4096-block level then spread over twice the CUs (measured 3% off the mixed sphere solve)
constexpr long kSmallWork = 1L << 20;

inline int nof(int L) { return 8 >> L; }
inline long lenof(int L) { long w = nof(L) + 2; return w * w * w; }
inline int pow2ge(int m) { int p = 1; while (p < m) p <<= 1; return p; }

enum Kinds { K_SAME = 1, K_PARENT = 2, K_BOTH = 3 };

// One set of FMG-level block arrays (codes U..PSI) and reflux overrides.
template <typename T> struct Arrays {
    T *A[kNfmg][NCODE] = {};
    T *ov[kNfmg] = {};
};

} // namespace

struct sg_amr {
    int nblk = 0, nlev = 0, bc = 0, zero_average = 0;
    int base_dims[3] = {0, 0, 0};
    double h0 = 0.0, x0[3] = {0, 0, 0};
    std::vector<int> level, ijk, parent, child, nbr, nkind, octant, covered, skip, lev_start;
    std::vector<int> active, phys, cf, cf_slot, ovslot, cov;
    std::vector<int> phys_off, cf_off, cov_off;   // per AMR level, size nlev + 1
    int nslots = 0, nbase = 0;
    // device tables
    int *d_level = nullptr, *d_ijk = nullptr, *d_parent = nullptr, *d_octant = nullptr, *d_nbr = nullptr;
    int *d_nkind = nullptr, *d_covered = nullptr, *d_skip = nullptr, *d_active = nullptr, *d_phys = nullptr;
    int *d_cf = nullptr, *d_cf_slot = nullptr, *d_ovslot = nullptr, *d_cov = nullptr;
    // fused refresh (mixed solve): covered blocks' face layers read by same-level leaf neighbours
    // (entries b*6+f), the child table, and per FMG level whether the hierarchy admits the fused
    // form (build_prep_tables); fused_on is set for the duration of a mixed solve that asks for it
    std::vector<int> covf;
    int *d_child = nullptr;
    // k_prep's source tables per fused FMG level (build_prep_tables): covered face-layer cells
    // {destination, restriction origin}, and per coarse-fine face the sources of the coarse
    // neighbour's face layer; d_cfrec holds {fine block, face + 8 octant, override slot, level}
    int *d_prep_cov[kNfmg] = {}, *d_prep_src[kNfmg] = {}, *d_cfrec = nullptr;
    int prep_ncov[kNfmg] = {};
    int fused_ok[kNfmg] = {};
    // per-block records (kRec ints, amr_kernels.h) of the relaxation / residual kernels, by block
    // number (d_brec) and in the order of the active list (d_arec); cell widths per FMG level
    int *d_brec = nullptr, *d_arec = nullptr;
    LevelScalars hs[kNfmg] = {}, his[kNfmg] = {};
    // vmg_ghostcell at FMG level 2 as one gather per AMR level where the hierarchy allows it
    // (same_fill_pairs): {destination, source or -1} pairs, all levels in one table
    int *d_vfill = nullptr;
    int vfill_off[kMaxLev + 1] = {};
    int vfill_ok[kMaxLev] = {};
    // fmg_ghostcell_fix of the base AMR level at FMG levels 1 and 2 as one gather each (k_fill_pairs;
    // pairs [gfix0_off[L], gfix0_off[L+1]) of d_gfix0), where the hierarchy allows it
    int *d_gfix0 = nullptr;
    int gfix0_off[kNfmg + 1] = {};
    int gfix0_ok[kNfmg] = {};
    // merged vmg_fas steps (vmg_fas_merged): per FMG-level-2 cell what vmg_ghostcell does to it
    // (-2 untouched, -1 zero, else the cell it copies), and per covered block the mask of the 64
    // cells vmg_tau_crop zeroes; usable when every level has its gather table
    int *d_vmap = nullptr;
    unsigned long long *d_crop = nullptr;
    int vmg_merged_ok = 0;
    // autotuned threads per block of the FMG-level-0 colour / residual kernel (k_blk8):
    // [0] relax FP32, [1] residual FP32, [2] relax FP64, [3] residual FP64; candidate timings in ms
    int blk8_nt[4] = {256, 256, 256, 256};
    float blk8_ms[4][3] = {};
    // the inner solve's FMG-level-0 source in the fused mode, compact and colour-split
    // (amr_kernels.h cc_index), 512 values per active block
    float *rhs0c = nullptr;
    // the inner solve's FMG-level-0 correction U in the fused mode, compact (512 values per active
    // block, in active-list order), its residual (the same) and its ghost store (6 x 64 per active
    // block: same-level faces written by the neighbours, coarse-fine ones by the refresh, physical
    // ones zero)
    float *u0c = nullptr;
    float *res0c = nullptr;
    float *gs0 = nullptr;
    // its tables (build_compact_tables), present when cc_ok
    int cc_ok = 0;
    int *d_aidx = nullptr, *d_ptab = nullptr, *d_cov_cc = nullptr, *d_src_cc = nullptr;
    // solves with zero physical ghosts instead of the multipole condition (the adjoint's
    // preconditioner: the composite operator's homogeneous part)
    int homog_bc = 0;
    // the adjoint (build_adjoint): A^T of the homogeneous composite operator and B^T of its physical
    // ghost coupling as CSR over the dense [block][8^3] index space and the ghost items, cell
    // volumes (zero on covered blocks), work vectors
    int adj_ready = 0;
    long adj_ng = 0, adj_nnz = 0;
    long *d_at_ptr = nullptr, *d_bt_ptr = nullptr;
    int *d_at_col = nullptr, *d_bt_col = nullptr;
    double *d_at_val = nullptr, *d_bt_val = nullptr, *d_vol = nullptr, *d_ivol = nullptr, *d_idiag = nullptr;
    double *d_adj = nullptr, *d_adj_part = nullptr, *d_adj_g = nullptr;
    int adj_m = 0;
    int fused_on = 0;
    // set around a refresh that must be the staged one within a fused-mode solve (the fused refresh
    // skipped; the other fused-mode launch forms, which give the staged values, kept)
    int staged_refresh = 0;
    // device state: FP64 arrays (plain solve and the mixed solve's outer loop), FP32 arrays (the
    // mixed solve's inner fmg_lin), allocated on first mixed use
    Arrays<double> d;
    Arrays<float> f;
    int have_f = 0;
    BaseDirect bd{};
    float *bd_mem = nullptr;
    float *bd_grid = nullptr;   // two dense n^3 grids of the multi-workgroup form
    int *bd_inv = nullptr;      // base block number by position
    size_t bd_lds_bytes = 0;
    int bd_ok = 0;
    hipGraphExec_t graph = nullptr;   // mixed solve: one outer iteration, captured once
    int graph_zero_average = -1;      // the zero_average setting the graph was captured with
    int graph_fused = -1;             // and the fused-refresh setting
    int graph_base_chip = -1;         // and the base-solve form
    size_t graph_nodes = 0;           // kernel nodes of the captured outer iteration
    double hl[kNfmg][kMaxLev] = {}, hil[kNfmg][kMaxLev] = {};
    double *d_h[kNfmg] = {}, *d_hi[kNfmg] = {};
    double *d_scal = nullptr;   // [0] Resh2maxg [1] scratch max [2..4] norms [5] mean [8..23] multipole
    double *h_scal = nullptr;   // pinned mirror
    double *d_part = nullptr;   // tree-sum scratch
    double *d_npart = nullptr;  // norm partials [3][max(kNormGroups, active blocks)]
    double *d_slots = nullptr;  // smoother max slots [kMaxSlots]
    int P_all = 1, P_base = 1;
    double *d_dense = nullptr;  // staging for host entry points
    Cubic64 st{};
    // base grid MG
    BaseMG bm{};
    double *bm_mem = nullptr;
    int *d_iters = nullptr;
    int base_mg_ok = 0;
    long bm_lds = 0;            // LDS doubles of the base-grid kernel
    // the base-grid mg_cycle spread over the chip (base_mg_chip; 0: the one-workgroup k_base_mg):
    // bm2 is bm with levels 0 and 1 in global memory (they cross launches)
    int base_chip = 1;
    // the plain FP64 solve with the fused refresh and merged VMG steps (the mixed solve's launch
    // structure, the same values as the reference's staged sequence); off by default
    int fp64_fused = 0;
    BaseMG bm2{};
    int fill0_level = -1;
    // hydro block memory (sg_amr_set_blockmem): its layout and the blocks' pool slots on the device
    sg_blockmem bml{};
    int have_bm = 0;
    int *d_slot_u = nullptr, *d_slot_u1 = nullptr, *d_slot_u2 = nullptr;
    // the hydro steps' item tables (build_hydro_tables, on first use): exchange items per (level,
    // direction), restriction items per fine level, refluxing items per coarse level grouped in
    // rounds of disjoint coarse cells
    int hydro_ready = 0;
    int4 *d_hx = nullptr, *d_fg = nullptr, *d_rf = nullptr;
    std::vector<int> hx_off, fg_off;
    std::vector<std::vector<int>> rf_rounds;   // per coarse level: offsets of its rounds, + end
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace {

template <typename T> inline Arrays<T> &ar(sg_amr &s);
template <> inline Arrays<double> &ar<double>(sg_amr &s) { return s.d; }
template <> inline Arrays<float> &ar<float>(sg_amr &s) { return s.f; }

template <typename K, typename... Args> int go(sg_amr &s, K kern, long work, Args... args)
{
    if (work <= 0) return SG_OK;
    const int nt = work < kSmallWork ? kThreads / 2 : kThreads;
    unsigned grid = (unsigned)((work + nt - 1) / nt);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(nt), 0, s.stream, args...);
    SG_CHECK_LAUNCH();
    return SG_OK;
}

// two independent steps (functors over items [0, n1) and [0, n2)) in one launch
template <class F1, class F2> int go2(sg_amr &s, F1 f1, long n1, F2 f2, long n2)
{
    return go(s, k_pair<F1, F2>, n1 + n2, f1, n1, f2, n2);
}

template <typename K, typename... Args> int go1(sg_amr &s, K kern, Args... args)
{
    hipLaunchKernelGGL(kern, dim3(1), dim3(1024), 0, s.stream, args...);
    SG_CHECK_LAUNCH();
    return SG_OK;
}

__global__ void k_set(double *p, double v) { *p = v; }

int set_scalar(sg_amr &s, int idx, double v)
{
    hipLaunchKernelGGL(k_set, dim3(1), dim3(1), 0, s.stream, s.d_scal + idx, v);
    SG_CHECK_LAUNCH();
    return SG_OK;
}

#define SG_NDISPATCH(L, FN, ...)                        \
    switch (L) {                                        \
    case 0: return FN<8, T>(__VA_ARGS__);               \
    case 1: return FN<4, T>(__VA_ARGS__);               \
    default: return FN<2, T>(__VA_ARGS__);              \
    }

// Every routine below runs on the FP64 arrays (T = double: the plain path, bitwise the reference's
// sequence) or on the FP32 arrays of the mixed path's inner solve (T = float).
template <typename T> constexpr bool kInner = std::is_same<T, float>::value;

// ---- fmg_converge ----
template <int N, typename T> int c2p_lev_n(sg_amr &s, int L, int code, int lc, int cubic)
{
    int b0 = s.lev_start[lc + 1], nb = s.lev_start[lc + 2] - b0;
    long work = (long)nb * (N / 2) * (N / 2) * (N / 2);
    T *A = ar<T>(s).A[L][code];
    if (cubic) return go(s, k_c2p<N, true, T>, work, A, b0, nb, s.d_parent, s.d_octant, s.st);
    return go(s, k_c2p<N, false, T>, work, A, b0, nb, s.d_parent, s.d_octant, s.st);
}
template <typename T> int c2p_lev(sg_amr &s, int L, int code, int lc, int cubic)
{
    SG_NDISPATCH(L, c2p_lev_n, s, L, code, lc, cubic);
}

template <typename T> int c2p(sg_amr &s, int L, int code, int cubic)
{
    for (int lc = s.nlev - 2; lc >= 0; --lc) SG_TRY(c2p_lev<T>(s, L, code, lc, cubic));
    return SG_OK;
}

// ---- fmg_ghostcell ----
template <int N, typename T> int ghost_pass_n(sg_amr &s, int l, int L, int d, int code, int cubic, int kinds)
{
    int b0 = s.lev_start[l], nb = s.lev_start[l + 1] - b0;
    if (l == 0) kinds &= K_SAME;   // fmg_ghfix_parentlev returns at the base level
    if (!kinds) return SG_OK;
    long W = N + 2;
    return go(s, k_ghost<N, T>, (long)nb * 2 * W * W, ar<T>(s).A[L][code], b0, nb, d, cubic, kinds, s.d_nbr, s.d_nkind,
              s.d_octant);
}
template <typename T> int ghost_pass(sg_amr &s, int l, int L, int d, int code, int cubic, int kinds)
{
    SG_NDISPATCH(L, ghost_pass_n, s, l, L, d, code, cubic, kinds);
}

// all three directions of one level in one launch (same values as the three staged passes)
template <int N, typename T> int ghost_level_n(sg_amr &s, int l, int L, int code, int cubic)
{
    int b0 = s.lev_start[l], nb = s.lev_start[l + 1] - b0;
    return go(s, k_ghost_level<N, T>, (long)nb * GhostCells<N>::COUNT, ar<T>(s).A[L][code], b0, nb, cubic, (const int *)s.d_nbr,
              (const int *)s.d_nkind, (const int *)s.d_octant);
}
template <typename T> int ghost_level(sg_amr &s, int l, int L, int code, int cubic)
{
    SG_NDISPATCH(L, ghost_level_n, s, l, L, code, cubic);
}

// fmg_ghostcell_fix (fmg_ghostcell.F90:43-56): levels coarse to fine, one launch per (level,
// direction); in fused mode the base level, whose faces are all same-level or physical, as one
// gather of resolved copies (same values).  `fused` fills each level in one launch instead (same
// values; measured slower on MI355X than the three staged launches at every FMG level, 2^3 blocks
// included, since an edge or corner cell re-evaluates the chain of passes it depends on, so the
// solves do not use it).
template <typename T> int ghostcell_fix(sg_amr &s, int L, int code, int cubic, bool fused = false)
{
    for (int l = 0; l < s.nlev; ++l) {
        if (l == 0 && s.fused_on && s.gfix0_ok[L]) {
            // the base level has same-level faces only: its three passes as one gather
            const int e0 = s.gfix0_off[L], ne = s.gfix0_off[L + 1] - e0;
            SG_TRY(go(s, k_fill_pairs<T>, (long)ne, ar<T>(s).A[L][code], (const int2 *)s.d_gfix0 + e0, ne));
            continue;
        }
        if (fused) SG_TRY(ghost_level<T>(s, l, L, code, cubic));
        else
            for (int d = 0; d < 3; ++d) SG_TRY(ghost_pass<T>(s, l, L, d, code, cubic, K_BOTH));
    }
    return SG_OK;
}

// ---- fmg_boundary ----
template <int N, typename T> int fill0_faces_n(sg_amr &s, int L, int code, int l0, int l1)
{
    int e0 = s.phys_off[l0], ne = s.phys_off[l1] - e0;
    long W = N + 2;
    return go(s, k_fill0_faces<N, T>, (long)ne * W * W, ar<T>(s).A[L][code], s.d_phys + e0, ne);
}
template <typename T> int fill0_faces(sg_amr &s, int L, int code, int l0, int l1)
{
    SG_NDISPATCH(L, fill0_faces_n, s, L, code, l0, l1);
}

// fmg_boundary_fill0 of FMG level Lc = Lf + 1 (array iuc) and of Lf (array iuf) in one launch
template <int NF, typename T> int fill0_two(sg_amr &s, int Lf, int iuf, int iuc)
{
    constexpr int NC = NF / 2;
    const int e0 = s.phys_off[0], ne = s.phys_off[s.nlev] - e0;
    const int *ph = (const int *)s.d_phys + e0;
    return go2(s, Fill0FacesF<NC, T>{ar<T>(s).A[Lf + 1][iuc], ph}, (long)ne * (NC + 2) * (NC + 2),
               Fill0FacesF<NF, T>{ar<T>(s).A[Lf][iuf], ph}, (long)ne * (NF + 2) * (NF + 2));
}

template <int N, typename T> int extrap_faces_n(sg_amr &s, int L, int code, int l0, int l1)
{
    int e0 = s.phys_off[l0], ne = s.phys_off[l1] - e0;
    long W = N + 2;
    if (s.fused_on)   // the three passes in one launch (same values)
        return go(s, k_extrap_all<N, T>, (long)ne * W * W, ar<T>(s).A[L][code], (const int *)s.d_phys + e0, ne,
                  (const int *)s.d_nkind);
    for (int d = 0; d < 3; ++d)
        SG_TRY(go(s, k_extrap_faces<N, T>, (long)ne * W * W, ar<T>(s).A[L][code], s.d_phys + e0, ne, d));
    return SG_OK;
}
template <typename T> int extrap_faces(sg_amr &s, int L, int code, int l0, int l1)
{
    SG_NDISPATCH(L, extrap_faces_n, s, L, code, l0, l1);
}

// ---- fmg_poisson: flux refresh + reflux, relax, resid ----
template <int N, typename T> int reflux_n(sg_amr &s, int L, int code)
{
    int ncf = (int)s.cf.size();
    return go(s, k_reflux<N, T>, (long)ncf * (N / 2) * (N / 2), (const T *)ar<T>(s).A[L][code], ar<T>(s).ov[L],
              (const int *)s.d_cf, (const int *)s.d_cf_slot, ncf, (const double *)s.d_hi[L], (const int *)s.d_level,
              (const int *)s.d_octant);
}
template <typename T> int reflux(sg_amr &s, int L, int code) { SG_NDISPATCH(L, reflux_n, s, L, code); }

// fmg_poisson_flux (fmg_poisson.F90:53-124): the face fluxes themselves are formed by their
// consumers; what remains here is the ghost refresh and the reflux overrides.
// One launch in place of fmg_poisson_flux's staged refresh and reflux (k_prep): same values in
// every cell the following colour pass or residual reads (they read same-level neighbours'
// cells directly).  The FmgLevel_fill0 skip is not needed: from the zero guess it guards, the
// fused refresh computes the same zeros the skipped refresh would leave.  Physical faces keep
// their stored values (zero in the inner solve, whose every refresh zeroes them; the multipole
// values in the outer residual, which refreshes without zeroing).
inline bool fused(const sg_amr &s, int L)
{
    return s.fused_on && !s.staged_refresh && L < kNfmg - 1 && s.fused_ok[L];
}

// The inner solve's FMG-level-0 correction U is held in compact form (u0c with its ghost store gs0;
// amr_kernels.h cc_index, gs_cell) where the fused form applies at FMG level 0 and the compact
// tables could be built: every kernel that reads or writes it in fmg_lin then has a compact
// variant, and the FP64 update adds its interiors.
template <typename T> inline bool compact0(const sg_amr &s, int L, int ju)
{
    if constexpr (std::is_same<T, float>::value) return L == 0 && ju == C_U && fused(s, 0) && s.cc_ok && s.u0c;
    return false;
}

template <int N, typename T> int prep_n(sg_amr &s, int L, int ju)
{
    constexpr int H = N / 2;
    const int ncov = s.prep_ncov[L], ncf = (int)s.cf.size();
    const long work = (long)ncov + (long)ncf * H * H;
    LevelScalars hil{};
    for (int l = 0; l < kMaxLev; ++l) hil.v[l] = s.hil[L][l];
    if constexpr (N == 8 && std::is_same<T, float>::value) {
        if (compact0<T>(s, L, ju))
            return go(s, k_prep<8, float, true>, work, s.u0c, s.f.ov[0], (const int2 *)s.d_cov_cc, ncov,
                      (const int4 *)s.d_cfrec, (const int *)s.d_src_cc, ncf, hil, s.gs0, (const int *)s.d_aidx);
    }
    return go(s, k_prep<N, T>, work, ar<T>(s).A[L][ju], ar<T>(s).ov[L], (const int2 *)s.d_prep_cov[L], ncov,
              (const int4 *)s.d_cfrec, (const int *)s.d_prep_src[L], ncf, hil, (T *)nullptr, (const int *)nullptr);
}


template <typename T> int poisson_flux(sg_amr &s, int L, int ju, int fill0flag)
{
    if (fused(s, L)) {
        s.fill0_level = -1;
        if (L == 0) return prep_n<8, T>(s, L, ju);
        return prep_n<4, T>(s, L, ju);
    }
    if (s.fill0_level != L) {
        SG_TRY(c2p<T>(s, L, ju, 0));
        SG_TRY(ghostcell_fix<T>(s, L, ju, 0));
        if (fill0flag) SG_TRY(fill0_faces<T>(s, L, ju, 0, s.nlev));
    }
    s.fill0_level = -1;
    return reflux<T>(s, L, ju);
}

// one colour pass over nrec blocks given by their records (rec), with the reflux overrides or not
template <int N, typename T>
int relax_pass_n(sg_amr &s, int L, int ju, int jrhs, const int *rec, int nrec, int ipass, bool with_ov,
                 double *resmax, bool halo_free)
{
    long work = (long)nrec * N * N * ((N + 1) / 2);
    return go(s, k_relax<N, T>, work, ar<T>(s).A[L][ju], (const T *)ar<T>(s).A[L][jrhs], rec, nrec, ipass, s.hs[L],
              s.his[L], with_ov ? (const T *)ar<T>(s).ov[L] : (const T *)nullptr, resmax, halo_free ? 1 : 0);
}
template <typename T>
int relax_pass(sg_amr &s, int L, int ju, int jrhs, const int *rec, int nrec, int ipass, bool with_ov,
               double *resmax, bool halo_free = false)
{
    SG_NDISPATCH(L, relax_pass_n, s, L, ju, jrhs, rec, nrec, ipass, with_ov, resmax, halo_free);
}

constexpr int kBlk8Nt[3] = {64, 128, 256};

// one launch of k_blk8 over the active blocks with nt threads per block
// (cast, npart: the residual's FP32 copy and norm partials, k_blk8)
// (gs: the compact level-0 form, FP32 only: Au and Aout compact, gs the ghost store, written
// through the context's push table)
template <bool RELAX, typename T>
int blk8(sg_amr &s, int nt, T *Aout, T *Au, const T *Arhs, int ipass, const T *ov, double *resmax, int halo,
         float *cast = nullptr, double *npart = nullptr, const float *rhsc = nullptr, T *gs = nullptr, int xcd = 0)
{
    const dim3 g((unsigned)s.active.size());
    const int *rec = s.d_arec;
    if constexpr (std::is_same<T, float>::value) {
        if (gs) {
            SG_REQUIRE(rhsc && s.d_ptab, "compact FMG-level-0 pass without its compact source or push table");
            // blocks in contiguous ranges per XCD: the pushes into neighbours' ghost stores mostly
            // stay in the writing XCD's L2
            auto launch = [&](auto kern, int n) {
                hipLaunchKernelGGL(kern, g, dim3(n), 0, s.stream, Aout, Au, Arhs, rec, ipass, s.hs[0], s.his[0], ov,
                                   resmax, 1, cast, npart, rhsc, gs, (const int *)s.d_ptab, 1);
            };
            if (nt == 64) launch(k_blk8<RELAX, float, 64, true>, 64);
            else if (nt == 128) launch(k_blk8<RELAX, float, 128, true>, 128);
            else launch(k_blk8<RELAX, float, 256, true>, 256);
            SG_CHECK_LAUNCH();
            return SG_OK;
        }
    }
    // a halo-free pass reads its same-level neighbours' boundary cells: blocks in contiguous ranges
    // per XCD, so that those reads mostly hit the reading XCD's L2
    xcd = xcd || halo;
    if (nt == 64)
        hipLaunchKernelGGL((k_blk8<RELAX, T, 64>), g, dim3(64), 0, s.stream, Aout, Au, Arhs, rec, ipass, s.hs[0], s.his[0],
                           ov, resmax, halo, cast, npart, rhsc, (T *)nullptr, (const int *)nullptr, xcd);
    else if (nt == 128)
        hipLaunchKernelGGL((k_blk8<RELAX, T, 128>), g, dim3(128), 0, s.stream, Aout, Au, Arhs, rec, ipass, s.hs[0],
                           s.his[0], ov, resmax, halo, cast, npart, rhsc, (T *)nullptr, (const int *)nullptr, xcd);
    else
        hipLaunchKernelGGL((k_blk8<RELAX, T, 256>), g, dim3(256), 0, s.stream, Aout, Au, Arhs, rec, ipass, s.hs[0],
                           s.his[0], ov, resmax, halo, cast, npart, rhsc, (T *)nullptr, (const int *)nullptr, xcd);
    SG_CHECK_LAUNCH();
    return SG_OK;
}

template <typename T> constexpr int kBlk8Slot = std::is_same<T, float>::value ? 0 : 2;

// the inner solve's FMG-level-0 source in compact form, where the fused mode keeps it there (the
// outer residual writes it; RHO at level 0 is read through it)
template <typename T> inline const float *src0c(const sg_amr &s, int L, int jrhs)
{
    if constexpr (std::is_same<T, float>::value) {
        if (s.fused_on && L == 0 && jrhs == C_RHO) return s.rhs0c;
    }
    return nullptr;
}

// The autotuner of the FMG-level-0 kernels: each candidate's time over 16 launches (after one
// warm-up) on the context's own arrays, relaxation and residual apart.  The values it leaves in U
// and RES do not matter: every solve starts from a cleared state.
template <typename T> int tune_blk8(sg_amr &s)
{
    if (s.active.empty()) return SG_OK;
    auto &R = ar<T>(s);
    const int halo = s.fused_ok[0];
    T *U = R.A[0][C_U], *RES = R.A[0][C_RES], *gs = nullptr;
    const float *rc = nullptr;   // the solve's FP32 passes read the compact source
    if constexpr (std::is_same<T, float>::value) {
        rc = s.rhs0c;
        if (halo && s.cc_ok) {   // and, in the fused mode, keep U in compact form
            U = s.u0c;
            RES = s.res0c;
            gs = s.gs0;
        }
    }
    for (int which = 0; which < 2; ++which) {
        float best = 1e30f;
        for (int c = 0; c < 3; ++c) {
            const int nt = kBlk8Nt[c];
            for (int rep = 0; rep < 17; ++rep) {
                if (rep == 1) SG_HIP(hipEventRecord(s.ev0, s.stream));
                if (which == 0)
                    SG_TRY((blk8<true, T>(s, nt, (T *)nullptr, U, R.A[0][C_RHS], 1 + (rep & 1), R.ov[0], nullptr, halo,
                                          nullptr, nullptr, rc, gs)));
                else
                    SG_TRY((blk8<false, T>(s, nt, RES, U, R.A[0][C_RHS], 0, R.ov[0], nullptr, halo, nullptr, nullptr,
                                           rc, gs)));
            }
            SG_HIP(hipEventRecord(s.ev1, s.stream));
            SG_HIP(hipEventSynchronize(s.ev1));
            float ms = 0.f;
            SG_HIP(hipEventElapsedTime(&ms, s.ev0, s.ev1));
            s.blk8_ms[kBlk8Slot<T> + which][c] = ms / 16;
            if (ms < best) {
                best = ms;
                s.blk8_nt[kBlk8Slot<T> + which] = nt;
            }
        }
    }
    return SG_OK;
}

// fmg_poisson_relax (fmg_poisson.F90:130-213).  With want_max, Resh2maxg lands in d_scal[0]; the
// reference overwrites Resh2maxg in every call but reads only the last one of fmg_lin.
// staged_last: the refresh before the second colour runs staged even where the fused form applies,
// so that every ghost cell afterwards holds the reference's value (the fused refresh writes only
// the cells a colour pass reads; a later cubic fmg_converge_c2p reads the others).
template <typename T> int poisson_relax(sg_amr &s, int L, int ju, int jrhs, int fill0flag, bool want_max,
                                        bool staged_last = false)
{
    double *slots = want_max ? s.d_slots : nullptr;
    if (want_max) {
        hipLaunchKernelGGL(k_zero_slots, dim3(1), dim3(64), 0, s.stream, s.d_slots);
        SG_CHECK_LAUNCH();
    }
    for (int ipass = 1; ipass <= 2; ++ipass) {
        if (staged_last && ipass == 2 && s.fused_on) {
            s.staged_refresh = 1;
            const int rc = poisson_flux<T>(s, L, ju, fill0flag);
            s.staged_refresh = 0;
            SG_TRY(rc);
        } else {
            SG_TRY(poisson_flux<T>(s, L, ju, fill0flag));
        }
        const bool hf = fused(s, L);
        if (L == 0 && compact0<T>(s, 0, ju)) {
            SG_TRY((blk8<true, T>(s, s.blk8_nt[kBlk8Slot<T>], (T *)nullptr, (T *)s.u0c, (const T *)nullptr, ipass,
                                  ar<T>(s).ov[0], slots, 1, nullptr, nullptr, src0c<T>(s, 0, jrhs), (T *)s.gs0)));
        } else if (L == 0) {
            SG_TRY((blk8<true, T>(s, s.blk8_nt[kBlk8Slot<T>], (T *)nullptr, ar<T>(s).A[0][ju], ar<T>(s).A[0][jrhs], ipass,
                                  ar<T>(s).ov[0], slots, hf ? 1 : 0, nullptr, nullptr, src0c<T>(s, 0, jrhs))));
        } else {
            SG_TRY(relax_pass<T>(s, L, ju, jrhs, s.d_arec, (int)s.active.size(), ipass, true, slots, hf));
        }
    }
    if (want_max) {
        hipLaunchKernelGGL(k_max_slots, dim3(1), dim3(64), 0, s.stream, (const double *)s.d_slots, s.d_scal + 0);
        SG_CHECK_LAUNCH();
    }
    return SG_OK;
}

template <int N, typename T> int resid_n(sg_amr &s, int L, int jres, int ju, int jrhs)
{
    int na = (int)s.active.size();
    const bool hf = fused(s, L);
    return go(s, k_resid<N, T>, (long)na * N * N * N, ar<T>(s).A[L][jres], (const T *)ar<T>(s).A[L][ju],
              (const T *)ar<T>(s).A[L][jrhs], (const int *)s.d_arec, na, s.his[L], (const T *)ar<T>(s).ov[L],
              hf ? 1 : 0);
}

template <typename T> int poisson_resid(sg_amr &s, int L, int jres, int ju, int jrhs, int fill0flag)
{
    SG_TRY(poisson_flux<T>(s, L, ju, fill0flag));
    if (L == 0 && compact0<T>(s, 0, ju)) {
        // the residual in compact form (res0c), which the restriction to FMG level 1 reads
        SG_REQUIRE(jres == C_RES, "compact FMG-level-0 residual goes to RES");
        return blk8<false, T>(s, s.blk8_nt[kBlk8Slot<T> + 1], (T *)s.res0c, (T *)s.u0c, (const T *)nullptr, 0,
                              ar<T>(s).ov[0], nullptr, 1, nullptr, nullptr, src0c<T>(s, 0, jrhs), (T *)s.gs0);
    }
    if (L == 0) {
        const bool hf = fused(s, 0);
        return blk8<false, T>(s, s.blk8_nt[kBlk8Slot<T> + 1], ar<T>(s).A[0][jres], ar<T>(s).A[0][ju], ar<T>(s).A[0][jrhs], 0,
                              ar<T>(s).ov[0], nullptr, hf ? 1 : 0, nullptr, nullptr, src0c<T>(s, 0, jrhs));
    }
    SG_NDISPATCH(L, resid_n, s, L, jres, ju, jrhs);
}

// ---- fmg.F90 helpers ----
template <typename T> int fmg_rstrct(sg_amr &s, int Lc, int iuc, int iuf)
{
    int na = (int)s.active.size();
    auto &R = ar<T>(s);
    if (Lc == 1 && src0c<T>(s, 0, iuf))
        return go(s, k_rstrct_cc, (long)na * 64, s.f.A[1][iuc], (const float *)s.rhs0c, (const int *)s.d_active, na);
    if (Lc == 1 && iuf == C_RES && compact0<T>(s, 0, C_U))
        return go(s, k_rstrct_cc, (long)na * 64, s.f.A[1][iuc], (const float *)s.res0c, (const int *)s.d_active, na);
    if (Lc == 1)
        return go(s, k_rstrct<4, T>, (long)na * 64, R.A[1][iuc], (const T *)R.A[0][iuf], (const int *)s.d_active, na);
    return go(s, k_rstrct<2, T>, (long)na * 8, R.A[2][iuc], (const T *)R.A[1][iuf], (const int *)s.d_active, na);
}

template <typename T> int interp_blocks(sg_amr &s, int Lf, int iuf, int iuc, int cubic)
{
    int na = (int)s.active.size();
    const int *al = s.d_active;
    auto &R = ar<T>(s);
    if (cubic && na > 0) {
        // tricubic: one workgroup per block, the separable passes staged in LDS
        if (Lf == 0)
            hipLaunchKernelGGL((k_interp_cubic_blk<8, T>), dim3(na), dim3(256), 0, s.stream, R.A[0][iuf],
                               (const T *)R.A[1][iuc], al, na);
        else
            hipLaunchKernelGGL((k_interp_cubic_blk<4, T>), dim3(na), dim3(256), 0, s.stream, R.A[1][iuf],
                               (const T *)R.A[2][iuc], al, na);
        SG_CHECK_LAUNCH();
        return SG_OK;
    }
    if (Lf == 0) return go(s, k_interp<8, false, T>, (long)na * 512, R.A[0][iuf], (const T *)R.A[1][iuc], al, na);
    return go(s, k_interp<4, false, T>, (long)na * 64, R.A[1][iuf], (const T *)R.A[2][iuc], al, na);
}

// fmg_interp (fmg.F90:786-828)
template <typename T> int fmg_interp(sg_amr &s, int Lf, int iuf, int iuc, int cubic)
{
    const int Lc = Lf + 1;
    SG_TRY(c2p<T>(s, Lc, iuc, cubic));
    SG_TRY(ghostcell_fix<T>(s, Lc, iuc, cubic));
    SG_TRY(extrap_faces<T>(s, Lc, iuc, 0, s.nlev));
    if (cubic && compact0<T>(s, Lf, iuf)) {
        // the compact level-0 U: interiors, and the boundary cells into the neighbours' ghost
        // stores; its physical faces are the ghost store's zeros, what fmg_boundary_fill0 leaves
        if (!s.active.empty()) {
            hipLaunchKernelGGL((k_interp_cubic_blk<8, T>), dim3((unsigned)s.active.size()), dim3(256), 0, s.stream,
                               ar<T>(s).A[0][iuf], (const T *)ar<T>(s).A[1][iuc], (const int *)s.d_active,
                               (int)s.active.size(), (T *)s.u0c, (T *)s.gs0, (const int *)s.d_ptab);
            SG_CHECK_LAUNCH();
        }
        return fill0_faces<T>(s, Lc, iuc, 0, s.nlev);
    }
    SG_TRY(interp_blocks<T>(s, Lf, iuf, iuc, cubic));
    if (s.fused_on) return Lf == 0 ? fill0_two<8, T>(s, Lf, iuf, iuc) : fill0_two<4, T>(s, Lf, iuf, iuc);
    SG_TRY(fill0_faces<T>(s, Lc, iuc, 0, s.nlev));
    return fill0_faces<T>(s, Lf, iuf, 0, s.nlev);
}

template <typename T> int fmg_fill0(sg_amr &s, int L, int code)
{
    int na = (int)s.active.size();
    return go(s, k_blocks_fill0<T>, (long)na * lenof(L), ar<T>(s).A[L][code], (const int *)s.d_active, 0, na,
              (int)lenof(L));
}

// fmg_rstrct(Lc, iuc <- iuf), then fmg_fill0(Lc, jz); in fused mode one launch (the zero writes
// neither the restriction's source nor its destination)
template <typename T> int fmg_rstrct_fill0(sg_amr &s, int Lc, int iuc, int iuf, int jz)
{
    if (!s.fused_on || iuc == jz) {
        SG_TRY(fmg_rstrct<T>(s, Lc, iuc, iuf));
        return fmg_fill0<T>(s, Lc, jz);
    }
    const int na = (int)s.active.size();
    auto &R = ar<T>(s);
    const Fill0BlocksF<T> z{R.A[Lc][jz], (const int *)s.d_active, 0, (int)lenof(Lc)};
    const long nz = (long)na * lenof(Lc);
    if constexpr (std::is_same<T, float>::value) {
        if (Lc == 1 && src0c<T>(s, 0, iuf))
            return go2(s, RstrctCCF{s.f.A[1][iuc], (const float *)s.rhs0c, (const int *)s.d_active}, (long)na * 64, z, nz);
        if (Lc == 1 && iuf == C_RES && compact0<T>(s, 0, C_U))
            return go2(s, RstrctCCF{s.f.A[1][iuc], (const float *)s.res0c, (const int *)s.d_active}, (long)na * 64, z, nz);
    }
    if (Lc == 1)
        return go2(s, RstrctF<4, T>{R.A[1][iuc], (const T *)R.A[0][iuf], (const int *)s.d_active}, (long)na * 64, z, nz);
    return go2(s, RstrctF<2, T>{R.A[2][iuc], (const T *)R.A[1][iuf], (const int *)s.d_active}, (long)na * 8, z, nz);
}

template <typename T> int fmg_copy(sg_amr &s, int L, int out, int in, int noskip)
{
    int na = noskip ? s.nblk : (int)s.active.size();
    return go(s, k_blocks_copy<T>, (long)na * lenof(L), ar<T>(s).A[L][out], (const T *)ar<T>(s).A[L][in],
              noskip ? (const int *)nullptr : (const int *)s.d_active, 0, na, (int)lenof(L));
}

template <typename T> int fmg_add(sg_amr &s, int L, int a, int da)
{
    int na = (int)s.active.size();
    return go(s, k_blocks_add<T>, (long)na * lenof(L), ar<T>(s).A[L][a], (const T *)ar<T>(s).A[L][da],
              (const int *)s.d_active, 0, na, (int)lenof(L));
}

template <typename T> int fmg_addint(sg_amr &s, int Lf, int juf, int juc, int jres)
{
    if constexpr (std::is_same<T, float>::value) {
        if (compact0<T>(s, Lf, juf)) {
            // fmg_interp's steps at FMG level 1, then the interpolation added to the compact U in
            // one pass (the staged sequence's level-0 RES is not formed)
            const int Lc = Lf + 1;
            SG_TRY(c2p<T>(s, Lc, juc, 0));
            SG_TRY(ghostcell_fix<T>(s, Lc, juc, 0));
            SG_TRY(extrap_faces<T>(s, Lc, juc, 0, s.nlev));
            if (!s.active.empty()) {
                hipLaunchKernelGGL(k_addint_cc, dim3((unsigned)s.active.size()), dim3(256), 0, s.stream, s.u0c,
                                   (const float *)s.f.A[1][juc], (const int *)s.d_active, s.gs0, (const int *)s.d_ptab);
                SG_CHECK_LAUNCH();
            }
            return fill0_faces<T>(s, Lc, juc, 0, s.nlev);
        }
    }
    SG_TRY(fmg_interp<T>(s, Lf, jres, juc, 0));
    return fmg_add<T>(s, Lf, juf, jres);
}

// ---- vmg.F90 at the coarsest FMG level (2^3 per block) ----
constexpr int LV = kNfmg - 1;
constexpr long LENV = 64;

template <typename T> int lev_copy(sg_amr &s, int l, int out, int in)
{
    int b0 = s.lev_start[l], nb = s.lev_start[l + 1] - b0;
    return go(s, k_blocks_copy<T>, (long)nb * LENV, ar<T>(s).A[LV][out], (const T *)ar<T>(s).A[LV][in],
              (const int *)nullptr, b0, nb, (int)LENV);
}

template <typename T> int lev_add(sg_amr &s, int l, int a, int da, bool covered_only)
{
    auto &R = ar<T>(s);
    if (covered_only) {
        int c0 = s.cov_off[l], nc = s.cov_off[l + 1] - c0;
        return go(s, k_blocks_add<T>, (long)nc * LENV, R.A[LV][a], (const T *)R.A[LV][da], (const int *)(s.d_cov + c0),
                  0, nc, (int)LENV);
    }
    int b0 = s.lev_start[l], nb = s.lev_start[l + 1] - b0;
    return go(s, k_blocks_add<T>, (long)nb * LENV, R.A[LV][a], (const T *)R.A[LV][da], (const int *)nullptr, b0, nb,
              (int)LENV);
}

// vmg_ghostcell (vmg.F90:728-742): same-level faces x, y, z, then zero on physical faces; in the
// mixed solve's fused mode one gather (k_fill_pairs) where the hierarchy allows it
template <typename T> int vmg_ghostcell(sg_amr &s, int l, int code)
{
    if (s.fused_on && s.vfill_ok[l]) {
        const int e0 = s.vfill_off[l], ne = s.vfill_off[l + 1] - e0;
        return go(s, k_fill_pairs<T>, (long)ne, ar<T>(s).A[LV][code], (const int2 *)s.d_vfill + e0, ne);
    }
    for (int d = 0; d < 3; ++d) SG_TRY(ghost_pass<T>(s, l, LV, d, code, 0, K_SAME));
    return fill0_faces<T>(s, LV, code, l, l + 1);
}

// vmg_gridboundary (vmg.F90:708-724): parent-level faces x, y, z (linear), then physical zero; in
// the mixed solve's fused mode one launch (k_parent_fill)
template <typename T> int vmg_gridboundary(sg_amr &s, int l, int code)
{
    if (s.fused_on && l >= 1) {
        const int b0 = s.lev_start[l], nb = s.lev_start[l + 1] - b0;
        const int fill0 = s.phys_off[l + 1] > s.phys_off[l] ? 1 : 0;
        hipLaunchKernelGGL(k_parent_fill<T>, dim3((unsigned)((nb + 7) / 8)), dim3(256), 0, s.stream, ar<T>(s).A[LV][code], b0,
                           nb, fill0, (const int *)s.d_nbr, (const int *)s.d_nkind, (const int *)s.d_octant);
        SG_CHECK_LAUNCH();
        return SG_OK;
    }
    for (int d = 0; d < 3; ++d) SG_TRY(ghost_pass<T>(s, l, LV, d, code, 0, K_PARENT));
    return fill0_faces<T>(s, LV, code, l, l + 1);
}

// vmg_poisson_relax (vmg_poisson.F90:88-145): a ghost refresh (vmg_ghostcell) before each colour.
// The inner (FP32) solve reads same-level neighbours directly instead, so a colour needs no refresh;
// but the ghosts the reference's last refresh leaves are read later (vmg_sub over whole arrays,
// p2c's edge and corner ghosts, the restriction's ghost passes), so with `last` the refresh before
// the final colour is kept: each refresh overwrites all the previous one wrote, so the ghost state
// afterwards is the reference's, with one refresh instead of two per call.
template <typename T> int vmg_relax(sg_amr &s, int l, int ju, int jrhs, bool last = true)
{
    int b0 = s.lev_start[l], nb = s.lev_start[l + 1] - b0;
    for (int ipass = 1; ipass <= 2; ++ipass) {
        if (!kInner<T> || (last && ipass == 2)) SG_TRY(vmg_ghostcell<T>(s, l, ju));
        SG_TRY(relax_pass<T>(s, LV, ju, jrhs, s.d_brec + (long)kRec * b0, nb, ipass, false, nullptr, kInner<T>));
    }
    return SG_OK;
}

template <typename T> int vmg_tau(sg_amr &s, int lc, int jtau, int ju)
{
    const int lf = lc + 1;
    if (lf > s.nlev - 1) return SG_OK;
    auto &R = ar<T>(s);
    SG_TRY(vmg_ghostcell<T>(s, lf, ju));
    int b0 = s.lev_start[lf], nb = s.lev_start[lf + 1] - b0;
    SG_TRY(go(s, k_tau_fine<2, T>, (long)nb * LENV, R.A[LV][jtau], (const T *)R.A[LV][ju], b0, nb, s.hil[LV][lf]));
    SG_TRY(c2p_lev<T>(s, LV, jtau, lc, 0));
    SG_TRY(vmg_ghostcell<T>(s, lc, ju));
    int c0 = s.cov_off[lc], nc = s.cov_off[lc + 1] - c0;
    SG_TRY(go(s, k_tau_add<2, T>, (long)nc * 8, R.A[LV][jtau], (const T *)R.A[LV][ju], (const int *)(s.d_cov + c0), nc,
              s.hil[LV][lc]));
    int e0 = s.cf_off[lf], ne = s.cf_off[lf + 1] - e0;
    return go(s, k_tau_crop<2, T>, (long)ne * 16, R.A[LV][jtau], (const int *)(s.d_cf + e0), ne, (const int *)s.d_parent);
}

template <typename T> int vmg_interp(sg_amr &s, int l, int juf, int juc)
{
    SG_TRY(vmg_ghostcell<T>(s, l - 1, juc));
    SG_TRY(extrap_faces<T>(s, LV, juc, l - 1, l));
    int b0 = s.lev_start[l], nb = s.lev_start[l + 1] - b0;
    SG_TRY(go(s, k_p2c<2, T>, (long)nb * 8, ar<T>(s).A[LV][juf], (const T *)ar<T>(s).A[LV][juc], b0, nb,
              (const int *)s.d_parent, (const int *)s.d_octant));
    return fill0_faces<T>(s, LV, juc, l - 1, l);
}

// The base grid's mg_cycle (mg.F90:42-94, Ncycle 2, Npre = Npost = 2, the reference's stopping
// test Resmaxg / rhomax < tol) with the finest level's steps spread over the chip and the coarser
// levels' stretches in one workgroup each (base_mg.h, k_bm_*): the same operations in the same
// order as the one-workgroup k_base_mg, so the same values.  The stopping test reads two scalars
// back per iteration.
int read_scalars(sg_amr &s);

int base_mg_chip(sg_amr &s)
{
    const BaseMG &M = s.bm2;
    const int n0 = M.n[0], n1 = M.n[1], bc = M.dirichlet;
    const double h0 = M.h[0];
    double *U0 = M.U[0], *RHS0 = M.RHS[0], *RES0 = M.RES[0], *RHO0 = M.RHO[0];
    double *U1 = M.U[1], *RHS1 = M.RHS[1], *RHO1 = M.RHO[1];
    double *acc = s.d_scal + 1;
    auto cube = [](int n) { return (long)(n + 2) * (n + 2) * (n + 2); };
    auto ghosts = [&](double *u, int n, int mode) { return go(s, k_bm_ghosts, cube(n), u, n, mode); };
    auto wrap = [&](double *u, int n) { return bc ? SG_OK : ghosts(u, n, 0); };   // mg_boundary_u
    auto relax0 = [&](double *u, const double *rhs, double *mx) -> int {
        const long ncell = (long)n0 * n0 * ((n0 + 1) / 2);
        for (int ipass = 1; ipass <= 2; ++ipass) {
            SG_TRY(wrap(u, n0));
            SG_TRY(go(s, k_bm_relax, ncell, u, rhs, n0, h0, ipass, mx));
        }
        return SG_OK;
    };
    auto interp0 = [&](double *uf, double *uc, int cubic) -> int {   // bmg_interp for level 0 from level 1
        SG_TRY(bc ? ghosts(uc, n1, 1) : ghosts(uc, n1, 0));
        SG_TRY(go(s, k_bm_interp, (long)n0 * n0 * n0, uf, (const double *)uc, n0, cubic));
        if (bc) {
            SG_TRY(ghosts(uf, n0, 2));
            return ghosts(uc, n1, 2);
        }
        SG_TRY(ghosts(uf, n0, 0));
        return ghosts(uc, n1, 0);
    };
    auto coarse = [&](int part) {
        hipLaunchKernelGGL(k_bm_coarse, dim3(1), dim3(kBaseThreads), sizeof(double) * s.bm_lds, s.stream, M, part);
        SG_CHECK_LAUNCH();
        return SG_OK;
    };
    SG_TRY(bc ? ghosts(M.psi, n0, 2) : ghosts(M.psi, n0, 0));
    double rhomax = 0.0;
    int it = 0;
    for (;;) {
        SG_TRY(wrap(M.psi, n0));
        if (it == 0) SG_TRY(set_scalar(s, 1, 0.0));
        SG_TRY(go(s, k_bm_resid, (long)n0 * n0 * n0, RHO0, (const double *)M.psi, (const double *)M.src, n0, h0,
                  it == 0 ? acc : (double *)nullptr));
        if (it == 0) {
            SG_TRY(read_scalars(s));
            rhomax = s.h_scal[1];
        }
        if (it >= M.maxit) break;
        // mg_lin (mg.F90:106-218)
        SG_TRY(go(s, k_bm_rstrct, (long)n1 * n1 * n1, RHO1, (const double *)RHO0, n1));
        SG_TRY(coarse(0));
        SG_TRY(interp0(U0, U1, 1));
        SG_TRY(go(s, k_bm_blas, cube(n0), RHS0, (const double *)RHO0, cube(n0), 1));
        for (int jc = 0; jc < 2; ++jc) {
            for (int p = 0; p < 2; ++p) SG_TRY(relax0(U0, RHS0, nullptr));
            SG_TRY(wrap(U0, n0));
            SG_TRY(go(s, k_bm_resid, (long)n0 * n0 * n0, RES0, (const double *)U0, (const double *)RHS0, n0, h0,
                      (double *)nullptr));
            SG_TRY(go(s, k_bm_rstrct, (long)n1 * n1 * n1, RHS1, (const double *)RES0, n1));
            SG_TRY(coarse(1));
            SG_TRY(interp0(RES0, U1, 0));
            SG_TRY(go(s, k_bm_blas, cube(n0), U0, (const double *)RES0, cube(n0), 2));
            for (int p = 0; p < 2; ++p) {
                const bool last = jc == 1 && p == 1;
                if (last) SG_TRY(set_scalar(s, 1, 0.0));
                SG_TRY(relax0(U0, RHS0, last ? acc : nullptr));
            }
        }
        SG_TRY(read_scalars(s));
        const double resmaxg = s.h_scal[1] / (h0 * h0);
        SG_TRY(go(s, k_bm_blas, cube(n0), M.psi, (const double *)U0, cube(n0), 2));
        ++it;
        if ((rhomax > 0 ? resmaxg / rhomax : 0.0) < M.tol) break;
    }
    return SG_OK;
}

// vmg_slvsml (vmg.F90:387-432).  FP64: the reference's mg_cycle on the gathered base grid.  Inner
// (FP32) solve: the same base problem solved exactly by separable transforms, spread over the chip
// (base_chip, the default) or in one workgroup (the same values).
template <typename T> int vmg_slvsml(sg_amr &s, int ju, int jrhs)
{
    if (kInner<T> && s.bd_ok && s.bd_grid && s.base_chip) {
        // gather, three forward passes (the third divided by the eigenvalue sums), three back
        // (the last scattered into the base blocks): the values of the one-workgroup form
        const int n = s.bd.n, n3 = n * n * n;
        float *g0 = s.bd_grid, *g1 = s.bd_grid + n3;
        float *Au = s.f.A[LV][ju];
        const int *inv = s.bd_inv;
        SG_TRY(go(s, k_bd_gather, (long)s.nbase * 8, g0, (const float *)s.f.A[LV][jrhs], s.nbase, (const int *)s.d_ijk, n));
        SG_TRY(go(s, k_bd_pass<0, 0>, (long)n3, g1, (const float *)g0, s.bd, Au, inv));
        SG_TRY(go(s, k_bd_pass<1, 0>, (long)n3, g0, (const float *)g1, s.bd, Au, inv));
        SG_TRY(go(s, k_bd_pass<2, 1>, (long)n3, g1, (const float *)g0, s.bd, Au, inv));
        SG_TRY(go(s, k_bd_pass<0, 0>, (long)n3, g0, (const float *)g1, s.bd, Au, inv));
        SG_TRY(go(s, k_bd_pass<1, 0>, (long)n3, g1, (const float *)g0, s.bd, Au, inv));
        return go(s, k_bd_pass<2, 2>, (long)n3, g0, (const float *)g1, s.bd, Au, inv);
    }
    if (kInner<T> && s.bd_ok) {
        hipLaunchKernelGGL(k_base_direct, dim3(1), dim3(kDirectThreads), s.bd_lds_bytes, s.stream, s.f.A[LV][ju],
                           (const float *)s.f.A[LV][jrhs], s.nbase, (const int *)s.d_ijk, s.bd);
        SG_CHECK_LAUNCH();
        return SG_OK;
    }
    if (kInner<T> || !s.base_mg_ok) {
        for (int it = 0; it < 10; ++it) SG_TRY(vmg_relax<T>(s, 0, ju, jrhs, it == 9));
        return SG_OK;
    }
    const int NG = s.bm.n[0];
    SG_TRY(go(s, k_base_gather<2, double>, (long)s.nbase * 8, s.bm.psi, s.bm.src, (const double *)s.d.A[LV][ju],
              (const double *)s.d.A[LV][jrhs], s.nbase, (const int *)s.d_ijk, NG));
    if (s.base_chip && s.bm.nlev >= 3) {
        SG_TRY(base_mg_chip(s));
    } else {
        hipLaunchKernelGGL(k_base_mg, dim3(1), dim3(kBaseThreads), sizeof(double) * s.bm_lds, s.stream, s.bm);
        SG_CHECK_LAUNCH();
    }
    return go(s, k_base_scatter<2, double>, (long)s.nbase * 8, s.d.A[LV][ju], (const double *)s.bm.psi, s.nbase,
              (const int *)s.d_ijk, NG);
}

// vmg_fas with consecutive steps merged into single launches (the mixed solve's fused mode; the
// same operations on every value, amr_kernels.h "merged steps of vmg_fas")
template <typename T> int vmg_fas_merged(sg_amr &s, int jrho)
{
    auto &R = ar<T>(s);
    T *U = R.A[LV][C_U], *RES = R.A[LV][C_RES], *RHS = R.A[LV][C_RHS], *RHO = R.A[LV][C_RHO], *PSI = R.A[LV][C_PSI];
    const long nall = (long)s.nblk * LENV;
    auto nlv = [&](int l) { return s.lev_start[l + 1] - s.lev_start[l]; };
    if (jrho != C_RHO)
        SG_TRY(go(s, k_blocks_copy<T>, nall, RHO, (const T *)R.A[LV][jrho], (const int *)nullptr, 0, s.nblk, (int)LENV));
    for (int lc = s.nlev - 2; lc >= 0; --lc)   // fmg_converge_c2p of U and of RHO, finest first
        SG_TRY(go(s, k_c2p_two<T>, (long)nlv(lc + 1), U, RHO, s.lev_start[lc + 1], nlv(lc + 1), (const int *)s.d_parent,
                  (const int *)s.d_octant));
    // fmg_boundary_fill0 of U on every level with the copy RHS <- RHO in one launch
    const int p0 = s.phys_off[0], np = s.phys_off[s.nlev] - p0;
    SG_TRY(go2(s, Fill0FacesF<2, T>{U, (const int *)s.d_phys + p0}, (long)np * 16, CopyBlocksF<T>{RHS, (const T *)RHO,
               (const int *)nullptr, 0, (int)LENV}, nall));
    bool bounded = false;   // level v's vmg_gridboundary ran with the previous step's tau finish
    for (int v = s.nlev - 1; v >= 1; --v) {
        if (!bounded) SG_TRY(vmg_gridboundary<T>(s, v, C_U));
        bounded = false;
        // the inner solve's relaxation keeps no refresh here: the parent- and same-level passes
        // below rewrite every ghost cell it would write (from the relaxed interiors, the parent
        // level and physical ghosts, none of which it writes)
        for (int p = 0; p < 2; ++p) SG_TRY(vmg_relax<T>(s, v, C_U, C_RHS, p == 1 && !kInner<T>));
        for (int d = 0; d < 3; ++d) SG_TRY(ghost_pass<T>(s, v, LV, d, C_U, 1, K_BOTH));
        // vmg_rstrct(v-1, U, cubic), vmg_copy(v-1, IRUF <- U), vmg_rstrct(v-1, RHS)
        SG_TRY(go(s, k_vmg_rstr<T>, (long)nlv(v - 1) * LENV, U, PSI, RHS, s.lev_start[v - 1], nlv(v - 1),
                  (const int *)s.d_child, (const int *)s.d_covered, s.st));
        // vmg_poisson_tau + vmg_tau_crop (vmg_tau), then vmg_add_onov
        if (s.vfill_ok[v - 1] && s.vfill_ok[v]) {
            // vmg_ghostcell(v) and the coarser level's vmg_ghostcell (adjacent in the pair table) in
            // the same launch: tau reads level v's interiors and face ghosts, on which the gather
            // writes the values the refresh above left (same-level faces: the same neighbour cells;
            // parent-level faces: untouched; physical faces: zero)
            const int e0 = s.vfill_off[v - 1], ne = s.vfill_off[v + 1] - e0;
            SG_TRY(go(s, k_tau_fine_c2p_fill<T>, (long)nlv(v) * LENV + ne, RES, U, s.lev_start[v], nlv(v), s.hil[LV][v],
                      (const int *)s.d_parent, (const int *)s.d_octant, (const int2 *)s.d_vfill + e0, ne));
        } else {
            SG_TRY(vmg_ghostcell<T>(s, v, C_U));
            SG_TRY(go(s, k_tau_fine_c2p<T>, (long)nlv(v) * LENV, RES, (const T *)U, s.lev_start[v], nlv(v), s.hil[LV][v],
                      (const int *)s.d_parent, (const int *)s.d_octant));
            SG_TRY(vmg_ghostcell<T>(s, v - 1, C_U));
        }
        const int c0 = s.cov_off[v - 1], nc = s.cov_off[v] - c0;
        if (v - 1 >= 1 && nlv(v - 1) > 0) {
            // with the next step's vmg_gridboundary(v - 1) in the same launch
            const int l = v - 1, fill0 = s.phys_off[l + 1] > s.phys_off[l] ? 1 : 0;
            const long g = (nlv(l) + 7) / 8 + ((long)nc * LENV + 255) / 256;
            hipLaunchKernelGGL(k_tau_finish_pfill<T>, dim3((unsigned)g), dim3(256), 0, s.stream, RES, RHS, U,
                               (const int *)(s.d_cov + c0), nc, s.hil[LV][l], (const unsigned long long *)(s.d_crop + c0),
                               s.lev_start[l], nlv(l), fill0, (const int *)s.d_nbr, (const int *)s.d_nkind,
                               (const int *)s.d_octant);
            SG_CHECK_LAUNCH();
            bounded = true;
        } else {
            SG_TRY(go(s, k_tau_finish<T>, (long)nc * LENV, RES, RHS, (const T *)U, (const int *)(s.d_cov + c0), nc,
                      s.hil[LV][v - 1], (const unsigned long long *)(s.d_crop + c0)));
        }
    }
    if (s.zero_average) SG_TRY(go1(s, k_zero_average<2, T>, R.A[LV][C_RHS], s.nbase, s.P_base, s.d_part));
    SG_TRY(vmg_slvsml<T>(s, C_U, C_RHS));
    for (int v = 1; v < s.nlev; ++v) {
        // vmg_sub(v-1) with vmg_interp's vmg_ghostcell(v-1, RES), its extrapolation, then
        // vmg_interp_p2c_lev with vmg_add(v), then vmg_interp's fill0
        const long base = (long)s.lev_start[v - 1] * LENV;
        SG_TRY(go(s, k_sub_fill<T>, (long)nlv(v - 1) * LENV, RES, (const T *)U, (const T *)PSI, base,
                  (int)((long)nlv(v - 1) * LENV), (const int *)(s.d_vmap + base)));
        SG_TRY(extrap_faces<T>(s, LV, C_RES, v - 1, v));
        // with vmg_gridboundary(v) in the same launch (it reads only level v's U and level v-1's U,
        // which the fill0 of RES at v-1 after it does not write)
        const int fill0 = s.phys_off[v + 1] > s.phys_off[v] ? 1 : 0;
        if (nlv(v) > 0) {
            hipLaunchKernelGGL(k_p2c_add_pfill<T>, dim3((unsigned)((nlv(v) + 7) / 8)), dim3(256), 0, s.stream, RES, U,
                               s.lev_start[v], nlv(v), fill0, (const int *)s.d_nbr, (const int *)s.d_nkind,
                               (const int *)s.d_parent, (const int *)s.d_octant);
            SG_CHECK_LAUNCH();
        }
        SG_TRY(fill0_faces<T>(s, LV, C_RES, v - 1, v));
        for (int p = 0; p < 2; ++p) SG_TRY(vmg_relax<T>(s, v, C_U, C_RHS, p == 1));
    }
    return SG_OK;
}

// vmg_fas (vmg.F90:116-243): Ncycle = 1, Npre = Npost = 2
template <typename T> int vmg_fas(sg_amr &s, int jrho)
{
    if (s.fused_on && s.vmg_merged_ok) return vmg_fas_merged<T>(s, jrho);
    auto &R = ar<T>(s);
    if (jrho != C_RHO)
        for (int l = 0; l < s.nlev; ++l) SG_TRY(lev_copy<T>(s, l, C_RHO, jrho));
    SG_TRY(c2p<T>(s, LV, C_U, 0));
    SG_TRY(c2p<T>(s, LV, C_RHO, 0));
    SG_TRY(fill0_faces<T>(s, LV, C_U, 0, s.nlev));
    for (int l = 0; l < s.nlev; ++l) SG_TRY(lev_copy<T>(s, l, C_RHS, C_RHO));
    for (int v = s.nlev - 1; v >= 1; --v) {
        SG_TRY(vmg_gridboundary<T>(s, v, C_U));
        for (int p = 0; p < 2; ++p) SG_TRY(vmg_relax<T>(s, v, C_U, C_RHS, p == 1));
        // vmg_rstrct(v-1, U, cubic): per direction parent-level then same-level faces (one
        // launch: the two write disjoint faces and read nothing the other writes)
        for (int d = 0; d < 3; ++d) SG_TRY(ghost_pass<T>(s, v, LV, d, C_U, 1, K_BOTH));
        SG_TRY(c2p_lev<T>(s, LV, C_U, v - 1, 1));
        SG_TRY(lev_copy<T>(s, v - 1, C_PSI, C_U));            // IRUF is the PSI slot (vmg.F90:115)
        SG_TRY(c2p_lev<T>(s, LV, C_RHS, v - 1, 0));
        SG_TRY(vmg_tau<T>(s, v - 1, C_RES, C_U));
        SG_TRY(lev_add<T>(s, v - 1, C_RHS, C_RES, true));     // vmg_add_onov
    }
    if (s.zero_average) SG_TRY(go1(s, k_zero_average<2, T>, R.A[LV][C_RHS], s.nbase, s.P_base, s.d_part));
    SG_TRY(vmg_slvsml<T>(s, C_U, C_RHS));
    for (int v = 1; v < s.nlev; ++v) {
        int b0 = s.lev_start[v - 1], nb = s.lev_start[v] - b0;
        SG_TRY(go(s, k_blocks_sub<T>, (long)nb * LENV, R.A[LV][C_RES], (const T *)R.A[LV][C_U], (const T *)R.A[LV][C_PSI],
                  b0, nb, (int)LENV));
        SG_TRY(vmg_interp<T>(s, v, C_RES, C_RES));
        SG_TRY(lev_add<T>(s, v, C_U, C_RES, false));
        SG_TRY(vmg_gridboundary<T>(s, v, C_U));
        for (int p = 0; p < 2; ++p) SG_TRY(vmg_relax<T>(s, v, C_U, C_RHS, p == 1));
    }
    return SG_OK;
}

// fmg_lin (fmg.F90:270-377)
// In the inner (FP32) solve's fused mode two steps are left out, both without effect on any value
// read later: FMG level 0 reads its source from RHO, of which fmg_copy(0, RHS <- RHO) would make
// a copy (nothing writes RHO or RHS at level 0 within fmg_lin); and the closing
// fmg_converge_c2p(0, U) is not run, since the covered blocks' U it writes is read by no later
// step (the update psi += U covers the leaf blocks, and every refresh forms covered cells from
// their children).
template <typename T> int fmg_lin(sg_amr &s)
{
    const int Lmax = kNfmg - 1;
    const bool lean = kInner<T> && s.fused_on;
    auto rhs = [&](int v) { return lean && v == 0 ? C_RHO : C_RHS; };
    for (int L = 1; L < Lmax; ++L) SG_TRY(fmg_rstrct<T>(s, L, C_RHO, C_RHO));
    SG_TRY(fmg_rstrct_fill0<T>(s, Lmax, C_RHO, C_RHO, C_U));
    SG_TRY(vmg_fas<T>(s, C_RHO));
    for (int Lf = Lmax - 1; Lf >= 0; --Lf) {
        SG_TRY(fmg_interp<T>(s, Lf, C_U, C_U, 1));
        if (!(lean && Lf == 0)) SG_TRY(fmg_copy<T>(s, Lf, C_RHS, C_RHO, 0));
        for (int jc = 0; jc < 2; ++jc) {
            for (int v = Lf; v <= Lmax - 1; ++v) {
                for (int p = 0; p < 2; ++p) SG_TRY(poisson_relax<T>(s, v, C_U, rhs(v), 1, false));
                SG_TRY(poisson_resid<T>(s, v, C_RES, C_U, rhs(v), 1));
                SG_TRY(fmg_rstrct_fill0<T>(s, v + 1, C_RHS, C_RES, C_U));
                s.fill0_level = v + 1;
            }
            SG_TRY(vmg_fas<T>(s, C_RHS));
            for (int v = Lmax - 1; v >= Lf; --v) {
                SG_TRY(fmg_addint<T>(s, v, C_U, C_U, C_RES));
                // the next FMG stage's cubic fmg_converge_c2p at FMG level Lf reads the ghosts of
                // the last relaxation there
                const bool last = jc == 1 && v == Lf && Lf >= 1;
                for (int p = 0; p < 2; ++p)
                    SG_TRY(poisson_relax<T>(s, v, C_U, rhs(v), 1, Lf == 0 && jc == 1 && v == 0 && p == 1,
                                            last && p == 1));
            }
        }
    }
    return lean ? SG_OK : c2p<T>(s, 0, C_U, 0);
}

Geo geo(sg_amr &s) { return Geo{s.d_ijk, s.d_level, s.d_h[0], s.x0[0], s.x0[1], s.x0[2]}; }

// fmg_boundary_physical_grid of the sphere module (fmg_boundary_phys_sphere.F90:80-354)
// squared radius of the sphere inside the domain whose outside the multipole condition masks
double sphere_rad2(const sg_amr &s)
{
    const int n = 8;
    const int NX = s.base_dims[0] * n, NY = s.base_dims[1] * n, NZ = s.base_dims[2] * n;
    double xmin = s.x0[0] + 0.5 * s.h0, ymin = s.x0[1] + 0.5 * s.h0, zmin = s.x0[2] + 0.5 * s.h0;
    double xmax = s.x0[0] + ((double)(NX - 1) + 0.5) * s.h0;
    double ymax = s.x0[1] + ((double)(NY - 1) + 0.5) * s.h0;
    double zmax = s.x0[2] + ((double)(NZ - 1) + 0.5) * s.h0;
    double m = xmax;
    if (ymax < m) m = ymax;
    if (zmax < m) m = zmax;
    if (-xmin < m) m = -xmin;
    if (-ymin < m) m = -ymin;
    if (-zmin < m) m = -zmin;
    return m * m;
}

int sphere_bc(sg_amr &s, int ju, int jrho)
{
    const double rad2 = sphere_rad2(s);
    Geo G = geo(s);
    SG_TRY(go(s, k_sphere_mask<8>, (long)s.nblk * lenof(0), s.d.A[0][jrho], s.nblk, G, rad2));
    const double dv = (s.h0 * s.h0) * s.h0;
    SG_TRY(go(s, k_sphere_moments<8>, (long)s.nbase, (const double *)s.d.A[0][jrho], s.nbase, s.P_base, G, dv, s.d_part));
    SG_TRY(go1(s, k_sphere_coef, s.d_part, s.nbase, s.P_base, s.d_scal + 8));
    int ne = (int)s.phys.size();
    return go(s, k_sphere_ghost<8>, (long)ne * 100, s.d.A[0][ju], (const int *)s.d_phys, ne, G,
              (const double *)(s.d_scal + 8));
}

// fmg_make_physvar mean removal over leaf cells (fmg_data_if.F90:53-78)
int remove_leaf_mean(sg_amr &s, int code)
{
    double *pm = s.d_part, *pv = s.d_part + s.P_all;
    SG_TRY(go(s, k_leaf_mass<8>, (long)s.nblk, (const double *)s.d.A[0][code], s.nblk, (const int *)s.d_covered,
              (const double *)s.d_h[0], (const int *)s.d_level, pm, pv));
    SG_TRY(go1(s, k_tree_mass, pm, pv, s.nblk, s.P_all, s.d_scal + 5));
    return go(s, k_sub_scalar, (long)s.nblk * lenof(0), s.d.A[0][code], (long)s.nblk * lenof(0),
              (const double *)(s.d_scal + 5));
}

// norms of an FMG-level-0 array over the leaf blocks -> d_scal[2..4]
constexpr int kNormGroups = 1024;
int norms(sg_amr &s, int code)
{
    int na = (int)s.active.size();
    hipLaunchKernelGGL(k_norms_part<8>, dim3(kNormGroups), dim3(kThreads), 0, s.stream, (const double *)s.d.A[0][code],
                       (const int *)s.d_active, na, (const double *)s.d_h[0], (const int *)s.d_level, s.d_npart);
    SG_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_norms_final, dim3(1), dim3(1024), 0, s.stream, (const double *)s.d_npart, kNormGroups,
                       s.d_scal + 2);
    SG_CHECK_LAUNCH();
    return SG_OK;
}

int read_scalars(sg_amr &s)
{
    SG_HIP(hipMemcpyAsync(s.h_scal, s.d_scal, 32 * sizeof(double), hipMemcpyDeviceToHost, s.stream));
    SG_HIP(hipStreamSynchronize(s.stream));
    return SG_OK;
}

__global__ void k_dense_to_blocks(double *A, const double *dense, int nblk)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nblk * 512) return;
    int b = (int)(t / 512), c = (int)(t % 512);
    A[(long)b * 1000 + Blk<8>::gi(c % 8, (c / 8) % 8, c / 64)] = dense[t];
}

__global__ void k_blocks_to_dense(double *dense, const double *A, int nblk)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nblk * 512) return;
    int b = (int)(t / 512), c = (int)(t % 512);
    dense[t] = A[(long)b * 1000 + Blk<8>::gi(c % 8, (c / 8) % 8, c / 64)];
}

int zero_state(sg_amr &s)
{
    for (int L = 0; L < kNfmg; ++L)
        for (int c = 0; c < NCODE; ++c)
            SG_HIP(hipMemsetAsync(s.d.A[L][c], 0, sizeof(double) * s.nblk * lenof(L), s.stream));
    s.fill0_level = -1;
    return SG_OK;
}

// fmg_cycle (fmg.F90:134-233).  psi, src, g: device arrays, dense [nblk][8^3] / [nblk][3][8^3].
int solve(sg_amr &s, double *psi, const double *src, double *g, const sg_params &p, sg_stats &st)
{
    std::memset(&st, 0, sizeof(st));
    SG_TRY(zero_state(s));
    SG_HIP(hipEventRecord(s.ev0, s.stream));
    SG_TRY(go(s, k_dense_to_blocks, (long)s.nblk * 512, s.d.A[0][C_U], (const double *)psi, s.nblk));
    SG_TRY(go(s, k_dense_to_blocks, (long)s.nblk * 512, s.d.A[0][C_RHO], src, s.nblk));
    if (s.zero_average) SG_TRY(remove_leaf_mean(s, C_RHO));
    SG_TRY(fmg_copy<double>(s, 0, C_PSI, C_U, 0));
    SG_TRY(fmg_copy<double>(s, 0, C_SRC, C_RHO, 0));
    SG_TRY(c2p<double>(s, 0, C_SRC, 0));
    SG_TRY(c2p<double>(s, 0, C_PSI, 0));
    if (s.bc == 1 && !s.homog_bc) SG_TRY(sphere_bc(s, C_PSI, C_SRC));
    SG_TRY(norms(s, C_SRC));
    SG_TRY(read_scalars(s));
    const double resh2max = s.h_scal[2], srcmax = s.h_scal[3], srcsq = s.h_scal[4];
    // The max norms drop NaN (fmax); the sum of squares keeps it, so test both.
    const bool src_finite = std::isfinite(resh2max) && std::isfinite(srcmax) && std::isfinite(srcsq);
    int nmax = src_finite ? p.maxit : 0;
    if (std::fabs(resh2max) <= 2.2250738585072014e-308) nmax = 0;
    double prev = HUGE_VAL;
    int it = 0, status = SG_ERR_NOT_CONVERGED;
    s.fused_on = s.fp64_fused;   // reset below, before the epilogue reads the stored ghosts
    auto measure = [&](int slot) -> int {
        SG_TRY(norms(s, C_RHO));
        SG_TRY(read_scalars(s));
        st.rel_linf = srcmax > 0 ? s.h_scal[3] / srcmax : 0.0;
        st.rel_l2 = srcsq > 0 ? std::sqrt(s.h_scal[4] / srcsq) : 0.0;
        st.rel_h2 = resh2max > 0 ? s.h_scal[2] / resh2max : 0.0;
        if (!std::isfinite(s.h_scal[3]) || !std::isfinite(s.h_scal[4])) st.rel_linf = st.rel_l2 = NAN;
        if (slot >= 0 && slot < 64) st.hist[slot] = st.rel_linf;
        return SG_OK;
    };
    for (;;) {
        SG_TRY(poisson_resid<double>(s, 0, C_RHO, C_PSI, C_SRC, 0));
        SG_TRY(measure(it));
        if (!src_finite || !std::isfinite(st.rel_l2)) { status = SG_ERR_NONFINITE; break; }
        if (p.stop_mode == SG_STOP_TRUE_RESIDUAL && st.rel_linf < p.tol) { status = SG_OK; break; }
        if (it >= nmax) break;
        SG_TRY(fmg_lin<double>(s));
        SG_TRY(fmg_add<double>(s, 0, C_PSI, C_U));
        ++it;
        SG_TRY(read_scalars(s));
        double ratio = s.h_scal[0] / resh2max;
        st.resmaxg = ratio;
        if (p.stop_mode == SG_STOP_REFERENCE) {
            if (ratio < p.tol) { status = SG_OK; break; }
            if (ratio > prev * 0.5) break;   // stall (fmg.F90:193)
            prev = ratio;
        }
    }
    if (nmax == 0 && src_finite) status = SG_OK;
    if (p.stop_mode == SG_STOP_REFERENCE && status != SG_ERR_NONFINITE) {
        SG_TRY(poisson_resid<double>(s, 0, C_RHO, C_PSI, C_SRC, 0));
        SG_TRY(measure(-1));
        if (!std::isfinite(st.rel_l2)) status = SG_ERR_NONFINITE;
    }
    st.iterations = it;
    st.status = status;
    s.fused_on = 0;
    SG_TRY(fmg_copy<double>(s, 0, C_U, C_PSI, 1));
    SG_TRY(c2p<double>(s, 0, C_U, 0));
    SG_TRY(go(s, k_blocks_to_dense, (long)s.nblk * 512, psi, (const double *)s.d.A[0][C_U], s.nblk));
    if (g) {
        SG_TRY(poisson_flux<double>(s, 0, C_U, 0));
        SG_TRY(go(s, k_psi2g<8>, (long)s.nblk * 512, g, (const double *)s.d.A[0][C_U], s.nblk, (const double *)s.d_hi[0],
                  (const int *)s.d_level, (const int *)s.d_skip, (const int *)s.d_ovslot, (const double *)s.d.ov[0]));
    }
    SG_HIP(hipEventRecord(s.ev1, s.stream));
    SG_HIP(hipStreamSynchronize(s.stream));
    float ms = 0.f;
    SG_HIP(hipEventElapsedTime(&ms, s.ev0, s.ev1));
    st.ms = ms;
    return status;
}

// ---- mixed precision: FP64 defect correction around the FP32 inner fmg_lin ----
int ensure_mixed(sg_amr &s)
{
    if (s.have_f) return SG_OK;
    for (int L = 0; L < kNfmg; ++L) {
        for (int c = 0; c < NCODE; ++c) SG_HIP(hipMalloc((void **)&s.f.A[L][c], sizeof(float) * s.nblk * lenof(L)));
        SG_HIP(hipMalloc((void **)&s.f.ov[L], sizeof(float) * std::max(1, s.nslots) * nof(L) * nof(L)));
    }
    const int bx = s.base_dims[0], n = 2 * bx;
    s.bd_ok = bx == s.base_dims[1] && bx == s.base_dims[2] && n <= kDirectMaxN;
    if (s.bd_ok) {
        const double h = s.hl[LV][0], pi = 3.14159265358979323846;
        std::vector<float> host((size_t)n * n + n);
        for (int m = 0; m < n; ++m) {
            for (int k = 0; k < n; ++k) {
                double q = s.bc == SG_BC_SPHERE
                               ? std::sqrt(2.0 / (n + 1)) * std::sin(pi * (m + 1) * (k + 1) / (n + 1))
                               : (std::cos(2 * pi * m * k / n) + std::sin(2 * pi * m * k / n)) / std::sqrt((double)n);
                host[(size_t)m * n + k] = (float)q;
            }
            double th = s.bc == SG_BC_SPHERE ? pi * (m + 1) / (n + 1) : 2 * pi * m / n;
            host[(size_t)n * n + m] = (float)((2.0 * std::cos(th) - 2.0) / (h * h));
        }
        if (s.bc != SG_BC_SPHERE) host[(size_t)n * n] = 0.f;   // the periodic zero mode
        SG_HIP(hipMalloc((void **)&s.bd_mem, sizeof(float) * host.size()));
        SG_HIP(hipMemcpy(s.bd_mem, host.data(), sizeof(float) * host.size(), hipMemcpyHostToDevice));
        s.bd.n = n;
        s.bd.Q = s.bd_mem;
        s.bd.lam = s.bd_mem + (size_t)n * n;
        s.bd_lds_bytes = base_direct_lds_bytes(n);
        SG_HIP(hipFuncSetAttribute((const void *)k_base_direct, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)s.bd_lds_bytes));
        // multi-workgroup form: scratch grids and the position -> base block map
        std::vector<int> inv((size_t)bx * bx * bx, -1);
        for (int b = 0; b < s.nbase; ++b) inv[((size_t)s.ijk[3 * b + 2] * bx + s.ijk[3 * b + 1]) * bx + s.ijk[3 * b]] = b;
        for (int v : inv) SG_REQUIRE(v >= 0, "base level does not cover its box");
        SG_HIP(hipMalloc((void **)&s.bd_grid, sizeof(float) * 2 * (size_t)n * n * n));
        SG_HIP(hipMalloc((void **)&s.bd_inv, sizeof(int) * inv.size()));
        SG_HIP(hipMemcpy(s.bd_inv, inv.data(), sizeof(int) * inv.size(), hipMemcpyHostToDevice));
    }
    SG_HIP(hipMalloc((void **)&s.rhs0c, sizeof(float) * 512 * std::max<size_t>(1, s.active.size())));
    SG_HIP(hipMemset(s.rhs0c, 0, sizeof(float) * 512 * std::max<size_t>(1, s.active.size())));
    const size_t nac = std::max<size_t>(1, s.active.size());
    SG_HIP(hipMalloc((void **)&s.u0c, sizeof(float) * 512 * nac));
    SG_HIP(hipMemset(s.u0c, 0, sizeof(float) * 512 * nac));
    SG_HIP(hipMalloc((void **)&s.res0c, sizeof(float) * 512 * nac));
    SG_HIP(hipMemset(s.res0c, 0, sizeof(float) * 512 * nac));
    SG_HIP(hipMalloc((void **)&s.gs0, sizeof(float) * 6 * 64 * nac));
    SG_HIP(hipMemset(s.gs0, 0, sizeof(float) * 6 * 64 * nac));
    s.have_f = 1;
    return tune_blk8<float>(s);
}

// The FP64 composite residual of the outer iteration (the reference operator: its refresh and
// reflux) and its norms -> d_scal[2..4].  In the fused mode the residual kernel writes the inner
// solve's FP32 source (the leaf interiors, all the inner solve reads of it) and its blocks' norm
// partials in place of the FP64 residual array, which nothing then reads.
int outer_resid(sg_amr &s)
{
    if (!s.fused_on || s.active.empty()) {
        SG_TRY(poisson_resid<double>(s, 0, C_RHO, C_PSI, C_SRC, 0));
        return norms(s, C_RHO);
    }
    SG_TRY(poisson_flux<double>(s, 0, C_PSI, 0));
    // (only the FP32 copy and the norms are read: the FP64 residual array is not written)
    SG_TRY((blk8<false, double>(s, s.blk8_nt[kBlk8Slot<double> + 1], (double *)nullptr, s.d.A[0][C_PSI],
                                s.d.A[0][C_SRC], 0, s.d.ov[0], nullptr, fused(s, 0) ? 1 : 0, s.rhs0c, s.d_npart)));
    hipLaunchKernelGGL(k_norms_final, dim3(1), dim3(1024), 0, s.stream, (const double *)s.d_npart, (int)s.active.size(),
                       s.d_scal + 2);
    SG_CHECK_LAUNCH();
    return SG_OK;
}

// One outer iteration: FP32 copy of the FP64 residual, inner fmg_lin in FP32, FP64 update
// psi += e, next FP64 composite residual and its norms.  Only launches: it is captured into a
// hipGraph once and replayed.
int mixed_iteration(sg_amr &s)
{
    const int nb = s.nblk;
    if (!s.fused_on)
        SG_TRY(go(s, k_cast_blocks, (long)nb * lenof(0), s.f.A[0][C_RHO], (const double *)s.d.A[0][C_RHO],
                  (const int *)nullptr, nb, (int)lenof(0)));
    SG_TRY(fmg_lin<float>(s));
    const int na = (int)s.active.size();
    if (compact0<float>(s, 0, C_U))
        SG_TRY(go(s, k_add_cast_cc, (long)na * 512, s.d.A[0][C_PSI], (const float *)s.u0c, (const int *)s.d_active, na));
    else
        SG_TRY(go(s, k_add_cast_blocks, (long)na * lenof(0), s.d.A[0][C_PSI], (const float *)s.f.A[0][C_U],
                  (const int *)s.d_active, na, (int)lenof(0)));
    return outer_resid(s);
}

int solve_mixed(sg_amr &s, double *psi, const double *src, double *g, const sg_params &p, sg_stats &st)
{
    std::memset(&st, 0, sizeof(st));
    SG_TRY(ensure_mixed(s));
    SG_TRY(zero_state(s));
    for (int L = 0; L < kNfmg; ++L)
        for (int c = 0; c < NCODE; ++c)
            SG_HIP(hipMemsetAsync(s.f.A[L][c], 0, sizeof(float) * s.nblk * lenof(L), s.stream));
    SG_HIP(hipEventRecord(s.ev0, s.stream));
    SG_TRY(go(s, k_dense_to_blocks, (long)s.nblk * 512, s.d.A[0][C_U], (const double *)psi, s.nblk));
    SG_TRY(go(s, k_dense_to_blocks, (long)s.nblk * 512, s.d.A[0][C_RHO], src, s.nblk));
    if (s.zero_average) SG_TRY(remove_leaf_mean(s, C_RHO));
    SG_TRY(fmg_copy<double>(s, 0, C_PSI, C_U, 0));
    SG_TRY(fmg_copy<double>(s, 0, C_SRC, C_RHO, 0));
    SG_TRY(c2p<double>(s, 0, C_SRC, 0));
    SG_TRY(c2p<double>(s, 0, C_PSI, 0));
    if (s.bc == 1 && !s.homog_bc) SG_TRY(sphere_bc(s, C_PSI, C_SRC));
    SG_TRY(norms(s, C_SRC));
    SG_TRY(read_scalars(s));
    const double resh2max = s.h_scal[2], srcmax = s.h_scal[3], srcsq = s.h_scal[4];
    // The max norms drop NaN (fmax); the sum of squares keeps it, so test both.
    const bool src_finite = std::isfinite(resh2max) && std::isfinite(srcmax) && std::isfinite(srcsq);
    int nmax = src_finite ? p.maxit : 0;
    if (std::fabs(resh2max) <= 2.2250738585072014e-308) nmax = 0;
    // fused refreshes in the outer residual and the inner fmg_lin (same values as the staged ones)
    struct FusedScope {
        sg_amr &s;
        FusedScope(sg_amr &s_, int on) : s(s_) { s.fused_on = on; }
        ~FusedScope() { s.fused_on = 0; }
    } fused_scope(s, p.fused ? 1 : 0);
    SG_TRY(outer_resid(s));
    SG_TRY(read_scalars(s));
    double prev = HUGE_VAL;
    int it = 0, status = SG_ERR_NOT_CONVERGED;
    for (;;) {
        st.rel_linf = srcmax > 0 ? s.h_scal[3] / srcmax : 0.0;
        st.rel_l2 = srcsq > 0 ? std::sqrt(s.h_scal[4] / srcsq) : 0.0;
        st.rel_h2 = resh2max > 0 ? s.h_scal[2] / resh2max : 0.0;
        if (!std::isfinite(s.h_scal[3]) || !std::isfinite(s.h_scal[4])) st.rel_linf = st.rel_l2 = NAN;
        if (it < 64) st.hist[it] = st.rel_linf;
        if (!src_finite || !std::isfinite(st.rel_l2)) { status = SG_ERR_NONFINITE; break; }
        if (p.stop_mode == SG_STOP_TRUE_RESIDUAL && st.rel_linf < p.tol) { status = SG_OK; break; }
        if (it >= nmax) break;
        if (p.use_graph) {
            if (!s.graph || s.graph_zero_average != s.zero_average || s.graph_fused != s.fused_on ||
                s.graph_base_chip != s.base_chip) {
                if (s.graph) (void)hipGraphExecDestroy(s.graph);
                s.graph = nullptr;
                hipGraph_t gr = nullptr;
                SG_HIP(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal));
                int rc = mixed_iteration(s);
                hipError_t e = hipStreamEndCapture(s.stream, &gr);
                if (rc == SG_OK && e == hipSuccess) e = hipGraphGetNodes(gr, nullptr, &s.graph_nodes);
                if (rc == SG_OK && e == hipSuccess) e = hipGraphInstantiate(&s.graph, gr, nullptr, nullptr, 0);
                if (gr) (void)hipGraphDestroy(gr);
                if (rc != SG_OK) return rc;
                SG_HIP(e);
                s.graph_zero_average = s.zero_average;
                s.graph_fused = s.fused_on;
                s.graph_base_chip = s.base_chip;
            }
            SG_HIP(hipGraphLaunch(s.graph, s.stream));
        } else {
            SG_TRY(mixed_iteration(s));
        }
        ++it;
        SG_TRY(read_scalars(s));
        double ratio = s.h_scal[0] / resh2max;
        st.resmaxg = ratio;
        if (p.stop_mode == SG_STOP_REFERENCE) {
            st.rel_linf = srcmax > 0 ? s.h_scal[3] / srcmax : 0.0;
            st.rel_l2 = srcsq > 0 ? std::sqrt(s.h_scal[4] / srcsq) : 0.0;
            st.rel_h2 = resh2max > 0 ? s.h_scal[2] / resh2max : 0.0;
            if (!std::isfinite(s.h_scal[3]) || !std::isfinite(s.h_scal[4])) { status = SG_ERR_NONFINITE; break; }
            if (ratio < p.tol) { status = SG_OK; break; }
            if (ratio > prev * 0.5) break;   // stall (fmg.F90:193)
            prev = ratio;
        }
    }
    if (nmax == 0 && src_finite) status = SG_OK;
    st.iterations = it;
    st.status = status;
    s.fused_on = 0;   // fmg_psi2g reads the stored ghosts: the staged refresh below writes them
    SG_TRY(fmg_copy<double>(s, 0, C_U, C_PSI, 1));
    SG_TRY(c2p<double>(s, 0, C_U, 0));
    SG_TRY(go(s, k_blocks_to_dense, (long)s.nblk * 512, psi, (const double *)s.d.A[0][C_U], s.nblk));
    if (g) {
        SG_TRY(poisson_flux<double>(s, 0, C_U, 0));
        SG_TRY(go(s, k_psi2g<8>, (long)s.nblk * 512, g, (const double *)s.d.A[0][C_U], s.nblk, (const double *)s.d_hi[0],
                  (const int *)s.d_level, (const int *)s.d_skip, (const int *)s.d_ovslot, (const double *)s.d.ov[0]));
    }
    SG_HIP(hipEventRecord(s.ev1, s.stream));
    SG_HIP(hipStreamSynchronize(s.stream));
    float ms = 0.f;
    SG_HIP(hipEventElapsedTime(&ms, s.ev0, s.ev1));
    st.ms = ms;
    return status;
}

// Host mirror of the staged refresh's rule (fmg_converge_c2p, then fmg_ghostcell_fix levels coarse
// to fine, x, y, z): the stored cell the value a refresh leaves in cell (i, j, k) of block b comes
// from, at an FMG level with N cells per block side, encoded as a k_prep source (amr_kernels.h).
// The last direction pass writing a ghost cell is the highest direction among its ghost
// coordinates whose face is not physical; a same-level pass copies the neighbour's layer as the
// earlier passes left it, so the cell is followed through the neighbours from z down to x.  False
// where the fused refresh cannot reproduce the value: a coarse-fine face on the way (the value
// would come from fmg_ghfix_parentlev), or a covered cell whose child there is covered itself (the
// restriction of a restriction).
bool prep_source(const sg_amr &s, int N, int b, int i, int j, int k, int &src)
{
    int p[3] = {i, j, k};
    for (int D = 2; D >= 0; --D) {
        if (p[D] == -1 || p[D] == N) {
            const int f = 6 * b + 2 * D + (p[D] == N ? 1 : 0);
            if (s.nkind[f] == NBR_PARENT) return false;
            if (s.nkind[f] == NBR_SAME) {
                p[D] = p[D] == N ? 0 : N - 1;
                b = s.nbr[f];
            }
        }
    }
    const long W = N + 2, LEN = W * W * W;
    auto gi = [W](int a, int bb, int c) { return ((long)(c + 1) * W + (bb + 1)) * W + (a + 1); };
    const bool interior = p[0] >= 0 && p[0] < N && p[1] >= 0 && p[1] < N && p[2] >= 0 && p[2] < N;
    if (interior && s.covered[b]) {
        const int H = N / 2, o = (p[0] >= H ? 1 : 0) + (p[1] >= H ? 2 : 0) + (p[2] >= H ? 4 : 0);
        const int c = s.child[8 * b + o];
        if (s.covered[c]) return false;
        const long org = (long)c * LEN + gi(2 * (p[0] - (o & 1) * H), 2 * (p[1] - ((o >> 1) & 1) * H),
                                            2 * (p[2] - ((o >> 2) & 1) * H));
        src = -(int)org - 1;
        return true;
    }
    src = (int)((long)b * LEN + gi(p[0], p[1], p[2]));
    return true;
}

// vmg_ghostcell of AMR level l (N cells per block side) as {destination, source} pairs, appended
// to `pairs`: the same-level passes x, y, z copy whole slabs, so the value a ghost cell ends with is
// the stored cell reached by following its ghost coordinates through same-level faces from z down
// to x (the last pass writing it reads what the earlier ones left), and a cell on a physical face
// ends as zero (source -1).  False where a single gather would not reproduce the passes: a source
// that the fill itself writes (a pass would read it before a later pass changes it).
bool same_fill_pairs(const sg_amr &s, int N, int l, std::vector<int> &pairs, bool zero_phys = true)
{
    const int b0 = s.lev_start[l], nb = s.lev_start[l + 1] - b0;
    const long W = N + 2, LEN = W * W * W;
    if ((long)s.nblk * LEN >= (long)INT32_MAX) return false;
    auto addr = [&](int b, int i, int j, int k) { return (long)b * LEN + ((long)(k + 1) * W + (j + 1)) * W + (i + 1); };
    std::vector<char> target((size_t)nb * LEN, 0);
    const size_t first = pairs.size();
    for (int b = b0; b < b0 + nb; ++b)
        for (int k = -1; k <= N; ++k)
            for (int j = -1; j <= N; ++j)
                for (int i = -1; i <= N; ++i) {
                    int p[3] = {i, j, k};
                    bool same = false, phys = false, ghost = false;
                    for (int d = 0; d < 3; ++d)
                        if (p[d] == -1 || p[d] == N) {
                            ghost = true;
                            const int kind = s.nkind[6 * b + 2 * d + (p[d] == N ? 1 : 0)];
                            phys = phys || kind == NBR_NONE;
                            same = same || kind == NBR_SAME;
                        }
                    if (!ghost || !(same || (phys && zero_phys))) continue;
                    const long dst = addr(b, i, j, k);
                    target[(size_t)(dst - (long)b0 * LEN)] = 1;
                    if (phys && zero_phys) {
                        pairs.insert(pairs.end(), {(int)dst, -1});
                        continue;
                    }
                    int m = b;
                    for (int D = 2; D >= 0; --D)
                        if (p[D] == -1 || p[D] == N) {
                            const int f = 6 * m + 2 * D + (p[D] == N ? 1 : 0);
                            if (s.nkind[f] == NBR_SAME) {
                                p[D] = p[D] == N ? 0 : N - 1;
                                m = s.nbr[f];
                            }
                        }
                    pairs.insert(pairs.end(), {(int)dst, (int)addr(m, p[0], p[1], p[2])});
                }
    for (size_t e = first; e < pairs.size(); e += 2) {
        const long rel = (long)pairs[e + 1] - (long)b0 * LEN;
        if (pairs[e + 1] >= 0 && rel >= 0 && rel < (long)nb * LEN && target[(size_t)rel]) {
            pairs.resize(first);
            return false;
        }
    }
    return true;
}

// k_prep's tables at FMG level L; false (the solves then keep the staged refresh there) where some
// value a colour pass or residual reads cannot be reproduced: every coarse value a coarse-fine ghost
// interpolates from, and every covered face-layer cell a same-level leaf reads.
bool build_prep_tables(sg_amr &s, int L, std::vector<int> &cov, std::vector<int> &tab)
{
    const int N = 8 >> L, H = N / 2;
    const long W = N + 2, LEN = W * W * W;
    if ((long)s.nblk * LEN >= (long)INT32_MAX) return false;   // sources are 32-bit offsets
    auto gi = [W](int a, int bb, int c) { return ((long)(c + 1) * W + (bb + 1)) * W + (a + 1); };
    cov.clear();
    tab.clear();
    for (int e : s.covf) {
        const int b = e / 6, f = e % 6, d = f / 2, a = (f & 1) ? N - 1 : 0;
        for (int t2 = 0; t2 < N; ++t2)
            for (int t1 = 0; t1 < N; ++t1) {
                int p[3];
                p[d] = a;
                p[d == 0 ? 1 : 0] = t1;
                p[d == 2 ? 1 : 2] = t2;
                int src = 0;
                if (!prep_source(s, N, b, p[0], p[1], p[2], src) || src >= 0) return false;
                cov.push_back((int)((long)b * LEN + gi(p[0], p[1], p[2])));
                cov.push_back(-src - 1);
            }
    }
    for (size_t e = 0; e < s.cf.size(); ++e) {
        const int b = s.cf[e] / 6, f = s.cf[e] % 6, d = f / 2, lr = f & 1;
        const int m = s.nbr[6 * b + f], o = s.octant[b];
        const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
        const int e1 = ((o >> ta) & 1) * H, e2 = ((o >> tb) & 1) * H;
        for (int y = -1; y <= H; ++y)
            for (int x = -1; x <= H; ++x) {
                int p[3];
                p[d] = lr ? 0 : N - 1;
                p[ta] = x + e1;
                p[tb] = y + e2;
                int src = 0;
                if (!prep_source(s, N, m, p[0], p[1], p[2], src)) return false;
                tab.push_back(src);
            }
    }
    return true;
}

// Tables of the compact FMG-level-0 form (amr_kernels.h gs_cell): the active-list position of each
// block (-1 for skipped ones); per active block and face, the ghost-store face its boundary cells
// go to (6 a' + face of the same-level active neighbour a' across it, else -1); k_prep's tables with
// every cell named by its compact position: covered items {ghost-store cell of the leaf neighbour,
// restriction origin a 512 + i + 8 j + 64 k}, sources v >= 0 a compact interior cell, -1 a physical
// ghost (zero), v <= -2 a restriction with origin -(v + 2).  False where a named cell has no
// compact position.
bool build_compact_tables(const sg_amr &s, const std::vector<int> &cov, const std::vector<int> &src,
                          std::vector<int> &aidx, std::vector<int> &ptab, std::vector<int> &cov_cc,
                          std::vector<int> &src_cc)
{
    constexpr long LEN = 1000;
    const int na = (int)s.active.size();
    aidx.assign(s.nblk, -1);
    for (int a = 0; a < na; ++a) aidx[s.active[a]] = a;
    ptab.assign((size_t)6 * std::max(1, na), -1);
    for (int a = 0; a < na; ++a) {
        const int b = s.active[a];
        for (int f = 0; f < 6; ++f) {
            const int m = s.nbr[6 * b + f];
            if (s.nkind[6 * b + f] == NBR_SAME && aidx[m] >= 0) ptab[6 * a + f] = 6 * aidx[m] + (f ^ 1);
        }
    }
    // a ghosted offset -> (block, i, j, k)
    auto split = [](long o, int &b, int &i, int &j, int &k) {
        b = (int)(o / LEN);
        const int r = (int)(o - (long)b * LEN);
        i = r % 10 - 1;
        j = (r / 10) % 10 - 1;
        k = r / 100 - 1;
    };
    auto inside = [](int i, int j, int k) { return i >= 0 && i < 8 && j >= 0 && j < 8 && k >= 0 && k < 8; };
    auto origin = [&](long o, int &v) {   // restriction origin in compact numbering
        int b, i, j, k;
        split(o, b, i, j, k);
        if (aidx[b] < 0 || !inside(i, j, k) || (i | j | k) & 1) return false;
        v = aidx[b] * 512 + i + 8 * j + 64 * k;
        return true;
    };
    auto cc = [](int i, int j, int k) { return ((i + j + k + 1) & 1) * 256 + (i >> 1) + 4 * j + 32 * k; };
    auto gcell = [](int F, int t1, int t2) { return ((((F & 1) ? 0 : 1) + t1 + t2) & 1) * 32 + (t1 >> 1) + 4 * t2; };
    for (int e : s.cf)
        if (aidx[e / 6] < 0) return false;   // k_prep names a coarse-fine item's block by its position
    cov_cc.clear();
    src_cc.clear();
    for (size_t e = 0; e < s.covf.size(); ++e) {
        const int f = s.covf[e] % 6, m = s.nbr[s.covf[e]];
        if (aidx[m] < 0) return false;
        for (int q = 0; q < 64; ++q) {
            const int t1 = q & 7, t2 = q >> 3;
            int org = 0;
            if (!origin(cov[2 * (e * 64 + q) + 1], org)) return false;
            cov_cc.push_back((6 * aidx[m] + (f ^ 1)) * 64 + gcell(f ^ 1, t1, t2));
            cov_cc.push_back(org);
        }
    }
    for (int v : src) {
        if (v < 0) {
            int org = 0;
            if (!origin(-(long)v - 1, org)) return false;
            src_cc.push_back(-org - 2);
            continue;
        }
        int b, i, j, k;
        split(v, b, i, j, k);
        if (!inside(i, j, k)) {
            src_cc.push_back(-1);
            continue;
        }
        if (aidx[b] < 0) return false;
        src_cc.push_back(aidx[b] * 512 + cc(i, j, k));
    }
    return true;
}

template <typename T> int upload(sg_amr &s, T **dptr, const std::vector<T> &v)
{
    size_t bytes = std::max<size_t>(1, v.size()) * sizeof(T);
    SG_HIP(hipMalloc((void **)dptr, bytes));
    if (!v.empty()) SG_HIP(hipMemcpy(*dptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return SG_OK;
}

void release(sg_amr *s)
{
    if (!s) return;
    int *ip[] = {s->d_level, s->d_ijk, s->d_parent, s->d_octant, s->d_nbr, s->d_nkind, s->d_covered, s->d_skip,
                 s->d_active, s->d_phys, s->d_cf, s->d_cf_slot, s->d_ovslot, s->d_cov, s->d_iters,
                 s->d_child, s->d_cfrec, s->d_brec, s->d_arec, s->d_vfill, s->d_vmap, s->d_gfix0,
                 s->d_slot_u, s->d_slot_u1, s->d_slot_u2, s->d_aidx, s->d_ptab, s->d_cov_cc, s->d_src_cc};
    for (int *q : ip) if (q) (void)hipFree(q);
    if (s->d_crop) (void)hipFree(s->d_crop);
    for (int4 *q : {s->d_hx, s->d_fg, s->d_rf}) if (q) (void)hipFree(q);
    if (s->rhs0c) (void)hipFree(s->rhs0c);
    for (float *q : {s->u0c, s->res0c, s->gs0}) if (q) (void)hipFree(q);
    for (long *q : {s->d_at_ptr, s->d_bt_ptr}) if (q) (void)hipFree(q);
    for (int *q : {s->d_at_col, s->d_bt_col}) if (q) (void)hipFree(q);
    for (double *q : {s->d_at_val, s->d_bt_val, s->d_vol, s->d_ivol, s->d_idiag, s->d_adj, s->d_adj_part, s->d_adj_g})
        if (q) (void)hipFree(q);
    for (int L = 0; L < kNfmg; ++L) {
        if (s->d_prep_cov[L]) (void)hipFree(s->d_prep_cov[L]);
        if (s->d_prep_src[L]) (void)hipFree(s->d_prep_src[L]);
    }
    if (s->graph) (void)hipGraphExecDestroy(s->graph);
    if (s->bd_mem) (void)hipFree(s->bd_mem);
    if (s->bd_grid) (void)hipFree(s->bd_grid);
    if (s->bd_inv) (void)hipFree(s->bd_inv);
    for (int L = 0; L < kNfmg; ++L) {
        for (int c = 0; c < NCODE; ++c) if (s->d.A[L][c]) (void)hipFree(s->d.A[L][c]);
        if (s->d.ov[L]) (void)hipFree(s->d.ov[L]);
        for (int c = 0; c < NCODE; ++c) if (s->f.A[L][c]) (void)hipFree(s->f.A[L][c]);
        if (s->f.ov[L]) (void)hipFree(s->f.ov[L]);
        if (s->d_h[L]) (void)hipFree(s->d_h[L]);
        if (s->d_hi[L]) (void)hipFree(s->d_hi[L]);
    }
    if (s->d_scal) (void)hipFree(s->d_scal);
    if (s->h_scal) (void)hipHostFree(s->h_scal);
    if (s->d_part) (void)hipFree(s->d_part);
    if (s->d_npart) (void)hipFree(s->d_npart);
    if (s->d_slots) (void)hipFree(s->d_slots);
    if (s->d_dense) (void)hipFree(s->d_dense);
    if (s->bm_mem) (void)hipFree(s->bm_mem);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

int create(sg_amr *s, int nblk, int nlev, const int *level, const int *ijk, const int *parent, const int *child,
           const int *nbr, const int base_dims[3], double h0, const double x0[3], int bc)
{
    SG_REQUIRE(nblk > 0 && nlev > 0 && nlev <= kMaxLev, "sg_amr_create: bad sizes nblk=%d nlev=%d", nblk, nlev);
    SG_REQUIRE(bc == SG_BC_PERIODIC || bc == SG_BC_SPHERE, "sg_amr_create: unknown boundary %d", bc);
    SG_REQUIRE(h0 > 0, "sg_amr_create: h0 must be positive");
    for (int a = 0; a < 3; ++a)
        SG_REQUIRE(base_dims[a] > 0 && base_dims[a] <= 4096, "sg_amr_create: bad base_dims[%d] = %d", a, base_dims[a]);
    s->nblk = nblk;
    s->nlev = nlev;
    s->bc = bc;
    s->zero_average = bc == SG_BC_PERIODIC;
    s->h0 = h0;
    for (int a = 0; a < 3; ++a) { s->base_dims[a] = base_dims[a]; s->x0[a] = x0[a]; }
    s->level.assign(level, level + nblk);
    s->ijk.assign(ijk, ijk + 3 * nblk);
    s->parent.assign(parent, parent + nblk);
    s->child.assign(child, child + 8 * nblk);
    s->nbr.assign(nbr, nbr + 6 * nblk);
    s->nkind.assign(6 * nblk, NBR_NONE);
    s->octant.assign(nblk, -1);
    s->covered.assign(nblk, 0);
    s->skip.assign(nblk, 0);
    s->lev_start.assign(nlev + 2, 0);
    for (int b = 0; b < nblk; ++b) {
        SG_REQUIRE(level[b] >= 0 && level[b] < nlev, "sg_amr_create: block %d has level %d", b, level[b]);
        SG_REQUIRE(b == 0 || level[b] >= level[b - 1], "sg_amr_create: blocks must be ordered by level");
        SG_REQUIRE((level[b] == 0) == (parent[b] < 0), "sg_amr_create: block %d parent/level mismatch", b);
        s->lev_start[level[b] + 1] = b + 1;
        for (int a = 0; a < 3; ++a)
            SG_REQUIRE(ijk[3 * b + a] >= 0 && ijk[3 * b + a] < (base_dims[a] << level[b]),
                       "sg_amr_create: block %d position %d out of [0, %d) on axis %d", b, ijk[3 * b + a],
                       base_dims[a] << level[b], a);
        s->covered[b] = child[8 * b] >= 0;
        if (parent[b] >= 0) {
            SG_REQUIRE(parent[b] < nblk && level[parent[b]] == level[b] - 1, "sg_amr_create: bad parent of %d", b);
            s->octant[b] = (ijk[3 * b] & 1) + 2 * (ijk[3 * b + 1] & 1) + 4 * (ijk[3 * b + 2] & 1);
            SG_REQUIRE(child[8 * parent[b] + s->octant[b]] == b, "sg_amr_create: child table disagrees for %d", b);
        }
        for (int o = 0; o < 8; ++o)
            SG_REQUIRE((child[8 * b + o] >= 0) == s->covered[b] && child[8 * b + o] < nblk,
                       "sg_amr_create: block %d must have all or no children", b);
        for (int f = 0; f < 6; ++f) {
            int m = nbr[6 * b + f];
            SG_REQUIRE(m < nblk, "sg_amr_create: neighbour out of range");
            if (m < 0) {
                SG_REQUIRE(bc == SG_BC_SPHERE, "sg_amr_create: periodic hierarchy with an open face (block %d)", b);
                s->nkind[6 * b + f] = NBR_NONE;
            } else if (level[m] == level[b]) {
                s->nkind[6 * b + f] = NBR_SAME;
            } else {
                SG_REQUIRE(level[m] == level[b] - 1, "sg_amr_create: block %d face %d is not properly nested", b, f);
                s->nkind[6 * b + f] = NBR_PARENT;
            }
        }
    }
    for (int l = 1; l <= nlev + 1; ++l) s->lev_start[l] = std::max(s->lev_start[l], s->lev_start[l - 1]);
    for (int l = 0; l < nlev; ++l)
        SG_REQUIRE(s->lev_start[l + 1] > s->lev_start[l], "sg_amr_create: level %d has no blocks", l);
    for (int b = 0; b < nblk; ++b) {
        s->skip[b] = level[b] != nlev - 1 && s->covered[b];
        SG_REQUIRE(!(level[b] == nlev - 1 && s->covered[b]), "sg_amr_create: finest-level block %d has children", b);
        if (!s->skip[b]) s->active.push_back(b);
    }
    s->nbase = s->lev_start[1];
    SG_REQUIRE(s->nbase == base_dims[0] * base_dims[1] * base_dims[2], "sg_amr_create: base level must hold %d blocks",
               base_dims[0] * base_dims[1] * base_dims[2]);
    // per-level lists
    s->phys_off.assign(nlev + 1, 0);
    s->cf_off.assign(nlev + 1, 0);
    s->cov_off.assign(nlev + 1, 0);
    s->ovslot.assign(6 * nblk, -1);
    for (int l = 0; l < nlev; ++l) {
        s->phys_off[l] = (int)s->phys.size();
        s->cf_off[l] = (int)s->cf.size();
        s->cov_off[l] = (int)s->cov.size();
        for (int b = s->lev_start[l]; b < s->lev_start[l + 1]; ++b) {
            if (s->covered[b]) s->cov.push_back(b);
            for (int f = 0; f < 6; ++f) {
                if (s->nkind[6 * b + f] == NBR_NONE) s->phys.push_back(6 * b + f);
                if (s->nkind[6 * b + f] == NBR_PARENT) {
                    int m = nbr[6 * b + f], fc = 6 * m + (f ^ 1);
                    SG_REQUIRE(!s->covered[b], "sg_amr_create: block %d meets a coarser level but has children", b);
                    SG_REQUIRE(!s->covered[m], "sg_amr_create: coarse neighbour %d of %d has children", m, b);
                    if (s->ovslot[fc] < 0) s->ovslot[fc] = s->nslots++;
                    s->cf.push_back(6 * b + f);
                    s->cf_slot.push_back(s->ovslot[fc]);
                }
            }
        }
    }
    s->phys_off[nlev] = (int)s->phys.size();
    s->cf_off[nlev] = (int)s->cf.size();
    s->cov_off[nlev] = (int)s->cov.size();
    for (int b = 0; b < nblk; ++b) {
        if (!s->covered[b]) continue;
        for (int f = 0; f < 6; ++f)
            if (s->nkind[6 * b + f] == NBR_SAME && !s->covered[nbr[6 * b + f]]) s->covf.push_back(6 * b + f);
    }
    std::vector<int> prep_cov[kNfmg], prep_src[kNfmg];
    for (int L = 0; L < kNfmg - 1; ++L) s->fused_ok[L] = build_prep_tables(*s, L, prep_cov[L], prep_src[L]) ? 1 : 0;
    std::vector<int> vfill;
    for (int l = 0; l < nlev; ++l) {
        s->vfill_off[l] = (int)vfill.size() / 2;
        s->vfill_ok[l] = same_fill_pairs(*s, 2, l, vfill) ? 1 : 0;
    }
    s->vfill_off[nlev] = (int)vfill.size() / 2;
    std::vector<int> gfix0;
    for (int L = 1; L < kNfmg; ++L) {
        s->gfix0_off[L] = (int)gfix0.size() / 2;
        s->gfix0_ok[L] = same_fill_pairs(*s, nof(L), 0, gfix0, false) ? 1 : 0;
    }
    s->gfix0_off[kNfmg] = (int)gfix0.size() / 2;
    std::vector<int> vmap;
    std::vector<unsigned long long> crop;
    s->vmg_merged_ok = 1;
    for (int l = 0; l < nlev; ++l) s->vmg_merged_ok &= s->vfill_ok[l];
    if (s->vmg_merged_ok) {
        // vmg_ghostcell's effect per cell, from the gather tables
        vmap.assign((size_t)nblk * lenof(LV), -2);
        for (size_t e = 0; e < vfill.size(); e += 2) vmap[(size_t)vfill[e]] = vfill[e + 1];
        // vmg_tau_crop: a level-l block's face meeting level l-1 zeroes its parent's plane there
        std::vector<int> covidx(nblk, -1);
        for (size_t e = 0; e < s->cov.size(); ++e) covidx[s->cov[e]] = (int)e;
        crop.assign(std::max<size_t>(1, s->cov.size()), 0ULL);
        const int W = nof(LV) + 2;
        for (size_t e = 0; e < s->cf.size(); ++e) {
            const int b = s->cf[e] / 6, f = s->cf[e] % 6, d = f / 2, a = (f & 1) ? nof(LV) - 1 : 0;
            unsigned long long &m = crop[(size_t)covidx[parent[b]]];
            for (int t2 = -1; t2 <= nof(LV); ++t2)
                for (int t1 = -1; t1 <= nof(LV); ++t1) {
                    int p[3];
                    p[d] = a;
                    p[d == 0 ? 1 : 0] = t1;
                    p[d == 2 ? 1 : 2] = t2;
                    m |= 1ULL << (((p[2] + 1) * W + (p[1] + 1)) * W + (p[0] + 1));
                }
        }
    }
    // device tables
    SG_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    SG_HIP(hipEventCreate(&s->ev0));
    SG_HIP(hipEventCreate(&s->ev1));
    SG_TRY(upload(*s, &s->d_level, s->level));
    SG_TRY(upload(*s, &s->d_ijk, s->ijk));
    SG_TRY(upload(*s, &s->d_parent, s->parent));
    SG_TRY(upload(*s, &s->d_octant, s->octant));
    SG_TRY(upload(*s, &s->d_nbr, s->nbr));
    SG_TRY(upload(*s, &s->d_nkind, s->nkind));
    SG_TRY(upload(*s, &s->d_covered, s->covered));
    SG_TRY(upload(*s, &s->d_skip, s->skip));
    SG_TRY(upload(*s, &s->d_active, s->active));
    SG_TRY(upload(*s, &s->d_phys, s->phys));
    SG_TRY(upload(*s, &s->d_cf, s->cf));
    SG_TRY(upload(*s, &s->d_cf_slot, s->cf_slot));
    SG_TRY(upload(*s, &s->d_ovslot, s->ovslot));
    SG_TRY(upload(*s, &s->d_cov, s->cov));
    SG_TRY(upload(*s, &s->d_child, s->child));
    {
        std::vector<int> brec((size_t)kRec * nblk, 0), arec;
        for (int b = 0; b < nblk; ++b) {
            int *r = brec.data() + (size_t)kRec * b;
            r[0] = b;
            r[1] = level[b];
            for (int f = 0; f < 6; ++f) {
                r[2 + f] = s->nkind[6 * b + f] == NBR_SAME ? nbr[6 * b + f] : -1;
                r[8 + f] = s->ovslot[6 * b + f];
            }
        }
        for (int b : s->active) arec.insert(arec.end(), brec.begin() + (size_t)kRec * b, brec.begin() + (size_t)kRec * (b + 1));
        SG_TRY(upload(*s, &s->d_brec, brec));
        SG_TRY(upload(*s, &s->d_arec, arec));
    }
    {
        std::vector<int> rec;
        for (size_t e = 0; e < s->cf.size(); ++e) {
            const int b = s->cf[e] / 6, f = s->cf[e] % 6;
            rec.insert(rec.end(), {b, f + 8 * s->octant[b], s->cf_slot[e], s->level[b]});
        }
        SG_TRY(upload(*s, &s->d_cfrec, rec));
        SG_TRY(upload(*s, &s->d_vfill, vfill));
        SG_TRY(upload(*s, &s->d_gfix0, gfix0));
        if (s->vmg_merged_ok) {
            SG_TRY(upload(*s, &s->d_vmap, vmap));
            SG_TRY(upload(*s, &s->d_crop, crop));
        }
        for (int L = 0; L < kNfmg - 1; ++L) {
            if (!s->fused_ok[L]) continue;
            s->prep_ncov[L] = (int)prep_cov[L].size() / 2;
            SG_TRY(upload(*s, &s->d_prep_cov[L], prep_cov[L]));
            SG_TRY(upload(*s, &s->d_prep_src[L], prep_src[L]));
        }
        // the inner solve's fused form at FMG level 0 is the compact one: without its tables the
        // level keeps the staged refresh (in both precisions)
        std::vector<int> aidx, ptab, cov_cc, src_cc;
        if (s->fused_ok[0] && build_compact_tables(*s, prep_cov[0], prep_src[0], aidx, ptab, cov_cc, src_cc)) {
            s->cc_ok = 1;
            SG_TRY(upload(*s, &s->d_aidx, aidx));
            SG_TRY(upload(*s, &s->d_ptab, ptab));
            SG_TRY(upload(*s, &s->d_cov_cc, cov_cc));
            SG_TRY(upload(*s, &s->d_src_cc, src_cc));
        } else {
            s->fused_ok[0] = 0;
        }
    }
    for (int L = 0; L < kNfmg; ++L) {
        for (int c = 0; c < NCODE; ++c) SG_HIP(hipMalloc((void **)&s->d.A[L][c], sizeof(double) * nblk * lenof(L)));
        SG_HIP(hipMalloc((void **)&s->d.ov[L], sizeof(double) * std::max(1, s->nslots) * nof(L) * nof(L)));
        for (int l = 0; l < nlev; ++l) {
            s->hl[L][l] = std::ldexp(h0, L - l);          // fmg_data.F90:288 CellWidth * 2^fmglev
            s->hil[L][l] = 1.0 / s->hl[L][l];
        }
        for (int l = 0; l < kMaxLev; ++l) {
            s->hs[L].v[l] = s->hl[L][l];
            s->his[L].v[l] = s->hil[L][l];
        }
        std::vector<double> hv(s->hl[L], s->hl[L] + kMaxLev), hiv(s->hil[L], s->hil[L] + kMaxLev);
        SG_TRY(upload(*s, &s->d_h[L], hv));
        SG_TRY(upload(*s, &s->d_hi[L], hiv));
    }
    SG_HIP(hipMalloc((void **)&s->d_scal, 32 * sizeof(double)));
    SG_HIP(hipMemset(s->d_scal, 0, 32 * sizeof(double)));
    SG_HIP(hipHostMalloc((void **)&s->h_scal, 32 * sizeof(double), hipHostMallocDefault));
    s->P_all = pow2ge(nblk);
    s->P_base = pow2ge(s->nbase);
    SG_HIP(hipMalloc((void **)&s->d_part, sizeof(double) * std::max<long>(2L * s->P_all, 16L * s->P_base)));
    SG_HIP(hipMalloc((void **)&s->d_iters, sizeof(int)));
    SG_HIP(hipMalloc((void **)&s->d_npart, sizeof(double) * 3 * std::max<size_t>(kNormGroups, s->active.size())));
    SG_HIP(hipMalloc((void **)&s->d_slots, sizeof(double) * kMaxSlots));
    // cubic restriction weights St (restriction.F90:38-63)
    {
        const double ic = 1.5;
        const int ii[4] = {0, 1, 2, 3};
        double s1[4];
        s1[0] = (ic - ii[1]) * (ic - ii[2]) * (ic - ii[3]) / (double)((ii[0] - ii[1]) * (ii[0] - ii[2]) * (ii[0] - ii[3]));
        s1[1] = (ic - ii[0]) * (ic - ii[2]) * (ic - ii[3]) / (double)((ii[1] - ii[0]) * (ii[1] - ii[2]) * (ii[1] - ii[3]));
        s1[2] = (ic - ii[0]) * (ic - ii[1]) * (ic - ii[3]) / (double)((ii[2] - ii[0]) * (ii[2] - ii[1]) * (ii[2] - ii[3]));
        s1[3] = (ic - ii[0]) * (ic - ii[1]) * (ic - ii[2]) / (double)((ii[3] - ii[0]) * (ii[3] - ii[1]) * (ii[3] - ii[2]));
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b)
                for (int c = 0; c < 4; ++c) s->st.w[(c * 4 + b) * 4 + a] = (s1[a] * s1[b]) * s1[c];
    }
    // base-grid MG (vmg_slvsml): power-of-two cube of base blocks, else 10 base relaxations
    const int bx = base_dims[0];
    s->base_mg_ok = bx == base_dims[1] && bx == base_dims[2] && (bx & (bx - 1)) == 0;
    if (s->base_mg_ok) {
        const int NG = bx * 2;
        int nl = 0;
        while ((1 << nl) < NG) ++nl;
        nl += 1;
        SG_REQUIRE(nl <= kBaseMaxLev && NG <= kBaseMaxN, "sg_amr_create: base grid %d too large", NG);
        long total = 2L * (NG + 2) * (NG + 2) * (NG + 2);
        for (int l = 0; l < nl; ++l) { long w = (NG >> l) + 2; total += 4 * w * w * w; }
        SG_HIP(hipMalloc((void **)&s->bm_mem, sizeof(double) * total));
        SG_HIP(hipMemset(s->bm_mem, 0, sizeof(double) * total));
        double *q = s->bm_mem;
        BaseMG &M = s->bm;
        M.nlev = nl;
        M.dirichlet = bc == SG_BC_SPHERE;
        M.tol = (double)1.e-2f;   // mg.F90:50: the single-precision literal 1.e-2 widened to double
        M.maxit = 10;             // nmg (mg.F90:52)
        M.iters = s->d_iters;
        long w0 = NG + 2;
        M.psi = q; q += w0 * w0 * w0;
        M.src = q; q += w0 * w0 * w0;
        for (int l = 0; l < nl; ++l) {
            M.n[l] = NG >> l;
            M.h[l] = s->hl[LV][0] * (double)(1 << l);   // mg_data.F90:554-559
            long w = M.n[l] + 2, len = w * w * w;
            M.U[l] = q; q += len;
            M.RHS[l] = q; q += len;
            M.RES[l] = q; q += len;
            M.RHO[l] = q; q += len;
            for (int a = 0; a < 4; ++a) M.lds_off[l][a] = -1;
        }
        // coarsest levels first: all four arrays up to 8^3, then U and RHS of larger levels
        for (int l = nl - 1; l >= 1; --l) {
            long w = M.n[l] + 2, len = w * w * w;
            int na = M.n[l] <= 8 ? 4 : 2;
            if ((s->bm_lds + na * len) * 8 > kBaseLdsBytes) break;
            for (int a = 0; a < na; ++a) {
                M.lds_off[l][a] = (int)s->bm_lds;
                s->bm_lds += len;
            }
        }
        SG_HIP(hipFuncSetAttribute((const void *)k_base_mg, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)(s->bm_lds * 8)));
        SG_HIP(hipFuncSetAttribute((const void *)k_bm_coarse, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)(s->bm_lds * 8)));
        s->bm2 = M;
        for (int l = 0; l < std::min(2, nl); ++l)
            for (int a = 0; a < 4; ++a) s->bm2.lds_off[l][a] = -1;
    }
    SG_TRY(tune_blk8<double>(*s));
    SG_HIP(hipDeviceSynchronize());
    return SG_OK;
}

int op(sg_amr &s, int opc, int L, int a, int b, int c, int flag, double *ret)
{
    SG_REQUIRE(L >= 0 && L < kNfmg, "sg_amr_op: FMG level %d out of range", L);
    SG_REQUIRE(a >= 0 && a < NCODE && b >= 0 && b < NCODE && c >= 0 && c < NCODE, "sg_amr_op: bad array code");
    double r = 0.0;
    // relax / resid with flag & 4: the fused refresh of the mixed solve (where the hierarchy admits it)
    struct FusedScope {
        sg_amr &s;
        FusedScope(sg_amr &s_, int on) : s(s_) { s.fused_on = on; }
        ~FusedScope() { s.fused_on = 0; }
    } fused_scope(s, (opc == SG_AMR_OP_RELAX || opc == SG_AMR_OP_RESID || opc == SG_AMR_OP_FMG_LIN) && (flag & 4) ? 1 : 0);
    if (opc == SG_AMR_OP_RELAX || opc == SG_AMR_OP_RESID) flag &= 1;
    switch (opc) {
    case SG_AMR_OP_C2P: SG_TRY(c2p<double>(s, L, a, flag)); break;
    case SG_AMR_OP_GHOSTFIX: SG_TRY(ghostcell_fix<double>(s, L, a, flag & 1, (flag & 2) != 0)); break;
    case SG_AMR_OP_FILL0: SG_TRY(fill0_faces<double>(s, L, a, 0, s.nlev)); break;
    case SG_AMR_OP_EXTRAP: SG_TRY(extrap_faces<double>(s, L, a, 0, s.nlev)); break;
    case SG_AMR_OP_RELAX:
        SG_TRY(poisson_relax<double>(s, L, a, b, flag, true));
        SG_TRY(read_scalars(s));
        r = s.h_scal[0];
        break;
    case SG_AMR_OP_RESID: SG_TRY(poisson_resid<double>(s, L, a, b, c, flag)); break;
    case SG_AMR_OP_RSTRCT:
        SG_REQUIRE(L >= 1, "sg_amr_op: restriction needs FMG level >= 1");
        SG_TRY(fmg_rstrct<double>(s, L, a, b));
        break;
    case SG_AMR_OP_INTERP:
        SG_REQUIRE(L <= 1, "sg_amr_op: interpolation needs FMG level <= 1");
        SG_TRY(fmg_interp<double>(s, L, a, b, flag));
        break;
    case SG_AMR_OP_ADDINT:
        SG_REQUIRE(L <= 1, "sg_amr_op: interpolation needs FMG level <= 1");
        SG_TRY(fmg_addint<double>(s, L, a, b, c));
        break;
    case SG_AMR_OP_VMG_FAS: SG_TRY(vmg_fas<double>(s, a)); break;
    case SG_AMR_OP_FMG_LIN:
        SG_TRY(fmg_lin<double>(s));
        SG_TRY(read_scalars(s));
        r = s.h_scal[0];
        break;
    case SG_AMR_OP_SPHERE_BC:
        SG_REQUIRE(s.bc == SG_BC_SPHERE, "sg_amr_op: not a sphere hierarchy");
        SG_TRY(sphere_bc(s, a, b));
        break;
    case SG_AMR_OP_REMOVE_MEAN: SG_TRY(remove_leaf_mean(s, a)); break;
    case SG_AMR_OP_VMG_TAU:
        SG_REQUIRE(L == LV, "sg_amr_op: vmg_tau runs at the coarsest FMG level %d", LV);
        SG_REQUIRE(flag >= 0 && flag < s.nlev, "sg_amr_op: vmg_tau AMR level %d out of range", flag);
        SG_TRY(vmg_tau<double>(s, flag, a, b));
        break;
    default: SG_REQUIRE(false, "sg_amr_op: unknown operation %d", opc);
    }
    SG_HIP(hipStreamSynchronize(s.stream));
    if (ret) *ret = r;
    return SG_OK;
}

// ---- the adjoint of the solve (sphere boundary) ----
// The converged solve maps the source to the potential linearly:
//   SRC = K C src (C: fmg_converge_c2p of the source, K: the multipole condition's mask),
//   A psi_leaf + B g = SRC_leaf with g = Mp SRC (the multipole ghost values of fmg_boundary_phys_sphere),
//   psi = C psi_leaf (the closing fmg_converge_c2p),
// with A the homogeneous composite operator on the leaf cells (same-level faces, the covered
// neighbours' restriction, the coarse-fine ghost of fmg_ghfix_parentlev and the reflux average of
// fmg_reflux) and B its coupling to the physical ghosts.  Its transpose:
//   src_bar = C^T K (lambda - Mp^T B^T lambda),  A^T lambda = (C^T psi_bar)_leaf.
// build_adjoint assembles A^T and B^T on the host from the hierarchy tables (the ghost sources are
// those of the fused refresh, build_prep_tables); A^T lambda = r is solved by FGMRES, right-
// preconditioned with V A^-1 V^-1 (V: cell volumes; exact where V A is symmetric, i.e. away from
// coarse-fine faces) applied through the forward solve with zero physical ghosts, with damped
// Jacobi sweeps on A^T around it (adj_precond).  With one refinement level this converges in about
// ten iterations; where refined regions nest (three or more AMR levels) the transposed interface
// terms of one level feed the next and the iteration stalls well above tight tolerances, which the
// status reports (SG_ERR_NOT_CONVERGED).

int build_adjoint(sg_amr &s)
{
    if (s.adj_ready) return SG_OK;
    std::vector<int> cov, tab;
    SG_REQUIRE(build_prep_tables(s, 0, cov, tab), "sg_amr_adjoint: the hierarchy's coarse-fine ghosts have no single-pass sources");
    const long n = (long)s.nblk * 512;
    SG_REQUIRE(n < (long)INT32_MAX, "sg_amr_adjoint: %d blocks exceed the 32-bit cell indices of the transposed operator", s.nblk);
    std::vector<int> cfidx((size_t)6 * s.nblk, -1), physidx((size_t)6 * s.nblk, -1);
    for (size_t e = 0; e < s.cf.size(); ++e) cfidx[s.cf[e]] = (int)e;
    for (size_t e = 0; e < s.phys.size(); ++e) physidx[s.phys[e]] = (int)e;
    // reflux override slot and coarse face position -> {coarse-fine item, its coarse cell (c1, c2)}
    std::vector<int> ovmap((size_t)std::max(1, s.nslots) * 64, -1);
    for (size_t e = 0; e < s.cf.size(); ++e) {
        const int fb = s.cf[e] / 6, d = (s.cf[e] % 6) / 2, o = s.octant[fb];
        const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
        const int off1 = ((o >> ta) & 1) * 4, off2 = ((o >> tb) & 1) * 4;
        for (int c2 = 0; c2 < 4; ++c2)
            for (int c1 = 0; c1 < 4; ++c1) ovmap[(size_t)s.cf_slot[e] * 64 + (off1 + c1) + 8 * (off2 + c2)] = (int)e * 16 + c1 + 4 * c2;
    }
    std::vector<std::pair<int, double>> ent;    // one row's leaf entries
    std::vector<std::pair<long, double>> gent;  // one row's ghost-item entries
    bool ok = true;
    // a cell of block b: the leaf cell, or a covered cell's restriction of its children
    std::function<void(int, int, int, int, double)> addcell = [&](int b, int i, int j, int k, double c) {
        if (!s.covered[b]) {
            ent.push_back({b * 512 + i + 8 * j + 64 * k, c});
            return;
        }
        const int o = (i >= 4 ? 1 : 0) + (j >= 4 ? 2 : 0) + (k >= 4 ? 4 : 0), ch = s.child[8 * b + o];
        if (ch < 0) { ok = false; return; }
        for (int dz = 0; dz < 2; ++dz)
            for (int dy = 0; dy < 2; ++dy)
                for (int dx = 0; dx < 2; ++dx) addcell(ch, 2 * (i & 3) + dx, 2 * (j & 3) + dy, 2 * (k & 3) + dz, c * 0.125);
    };
    // a stored physical ghost cell of block b -> its item in the physical-face list (k_sphere_ghost)
    auto ghost_item = [&](int b, const int p[3]) -> long {
        for (int D = 0; D < 3; ++D) {
            if (p[D] != -1 && p[D] != 8) continue;
            const int e = physidx[6 * b + 2 * D + (p[D] == 8 ? 1 : 0)];
            if (e < 0) continue;
            const int t1 = D == 0 ? p[1] : p[0], t2 = D == 2 ? p[1] : p[2];
            return (long)e * 100 + (t1 + 1) + 10 * (t2 + 1);
        }
        return -1;
    };
    // a source of the fused refresh (build_prep_tables' convention) times c
    auto addsrc = [&](int sv, double c) {
        const long o = sv < 0 ? -(long)sv - 1 : sv;
        const int b = (int)(o / 1000), r = (int)(o % 1000);
        int p[3] = {r % 10 - 1, (r / 10) % 10 - 1, r / 100 - 1};
        const bool in = p[0] >= 0 && p[0] < 8 && p[1] >= 0 && p[1] < 8 && p[2] >= 0 && p[2] < 8;
        if (sv < 0) {
            for (int dz = 0; dz < 2; ++dz)
                for (int dy = 0; dy < 2; ++dy)
                    for (int dx = 0; dx < 2; ++dx) addcell(b, p[0] + dx, p[1] + dy, p[2] + dz, c * 0.125);
        } else if (in) {
            addcell(b, p[0], p[1], p[2], c);
        } else {
            const long t = ghost_item(b, p);
            if (t < 0) ok = false;
            else gent.push_back({t, c});
        }
    };
    // c times the fine face flux ffv of coarse-fine item e at fine tangential (2 c1 + q1, 2 c2 + q2):
    // (+-)(g - u0) hi_f with g = (10 u0 + 8 uH - 3 uN) / 15 and uH the (9,3,3,1)/16 bilinear value
    auto add_cf_flux = [&](int e, int c1, int c2, int q1, int q2, double c) {
        const int fb = s.cf[e] / 6, f = s.cf[e] % 6, d = f / 2, lr = f & 1;
        const double hif = 1.0 / std::ldexp(s.h0, -s.level[fb]);
        const double sg = lr ? c * hif : -c * hif;
        const int s1 = 2 * c1 + q1, s2 = 2 * c2 + q2;
        auto at = [&](int a, double w) {
            const int i = d == 0 ? a : s1, j = d == 1 ? a : (d == 0 ? s1 : s2), k = d == 2 ? a : s2;
            addcell(fb, i, j, k, w);
        };
        at(lr ? 7 : 0, sg * (-5.0 / 15.0));
        at(lr ? 6 : 1, sg * (-3.0 / 15.0));
        const double cH = sg * (8.0 / 15.0) / 16.0;
        const int *sr = tab.data() + (size_t)e * 36;
        const int n1 = 2 * q1, n2 = 2 * q2;
        addsrc(sr[(c2 + 1) * 6 + (c1 + 1)], 9.0 * cH);
        addsrc(sr[(c2 + n2) * 6 + (c1 + n1)], cH);
        addsrc(sr[(c2 + n2) * 6 + (c1 + 1)], 3.0 * cH);
        addsrc(sr[(c2 + 1) * 6 + (c1 + n1)], 3.0 * cH);
    };
    // rows of A (leaf cells) as COO, merged per row; B as COO over ghost items
    std::vector<int> ar, ac, br;
    std::vector<long> bc;
    std::vector<double> av, bv;
    std::vector<double> vol(n, 0.0), ivol(n, 0.0), idiag(n, 0.0);
    for (int b : s.active) {
        const double h = std::ldexp(s.h0, -s.level[b]), hi = 1.0 / h, hi2 = hi * hi;
        for (int k = 0; k < 8; ++k)
            for (int j = 0; j < 8; ++j)
                for (int i = 0; i < 8; ++i) {
                    const int row = b * 512 + i + 8 * j + 64 * k;
                    vol[row] = h * h * h;
                    ivol[row] = 1.0 / vol[row];
                    ent.clear();
                    gent.clear();
                    for (int d = 0; d < 3; ++d)
                        for (int lr = 0; lr < 2; ++lr) {
                            int q[3] = {i, j, k};
                            q[d] += lr ? 1 : -1;
                            if (q[d] >= 0 && q[d] < 8) {
                                addcell(b, q[0], q[1], q[2], hi2);
                                addcell(b, i, j, k, -hi2);
                                continue;
                            }
                            const int f = 2 * d + lr, t1 = d == 0 ? j : i, t2 = d == 2 ? j : k;
                            const double sgn = lr ? 1.0 : -1.0;
                            if (s.ovslot[6 * b + f] >= 0) {   // fmg_reflux: the 2x2 average of the fine fluxes
                                const int m = ovmap[(size_t)s.ovslot[6 * b + f] * 64 + t1 + 8 * t2];
                                if (m < 0) { ok = false; continue; }
                                for (int q2 = 0; q2 < 2; ++q2)
                                    for (int q1 = 0; q1 < 2; ++q1) add_cf_flux(m / 16, m % 4, (m / 4) % 4, q1, q2, sgn * hi * 0.25);
                                continue;
                            }
                            const int kind = s.nkind[6 * b + f];
                            if (kind == NBR_SAME) {
                                q[d] = lr ? 0 : 7;
                                addcell(s.nbr[6 * b + f], q[0], q[1], q[2], hi2);
                                addcell(b, i, j, k, -hi2);
                            } else if (kind == NBR_PARENT) {
                                const int e = cfidx[6 * b + f];
                                if (e < 0) { ok = false; continue; }
                                add_cf_flux(e, t1 >> 1, t2 >> 1, t1 & 1, t2 & 1, sgn * hi);
                            } else {
                                q[d] = lr ? 8 : -1;
                                const long t = ghost_item(b, q);
                                if (t < 0) { ok = false; continue; }
                                gent.push_back({t, hi2});
                                addcell(b, i, j, k, -hi2);
                            }
                        }
                    std::sort(ent.begin(), ent.end());
                    for (size_t q = 0; q < ent.size();) {
                        size_t r = q;
                        double v = 0.0;
                        while (r < ent.size() && ent[r].first == ent[q].first) v += ent[r++].second;
                        ar.push_back(row);
                        ac.push_back(ent[q].first);
                        av.push_back(v);
                        if (ent[q].first == row && v != 0.0) idiag[row] = 1.0 / v;
                        q = r;
                    }
                    for (auto &g : gent) {
                        br.push_back(row);
                        bc.push_back(g.first);
                        bv.push_back(g.second);
                    }
                }
    }
    SG_REQUIRE(ok, "sg_amr_adjoint: a cell of the composite operator has no source");
    // A^T and B^T in CSR by counting sort on the column
    const long ng = (long)s.phys.size() * 100;
    std::vector<long> atp(n + 1, 0), btp(ng + 1, 0);
    for (int c : ac) ++atp[c + 1];
    for (long c : bc) ++btp[c + 1];
    for (long q = 0; q < n; ++q) atp[q + 1] += atp[q];
    for (long q = 0; q < ng; ++q) btp[q + 1] += btp[q];
    std::vector<int> atc(ac.size()), btc(bc.size());
    std::vector<double> atv(ac.size()), btv(bc.size());
    {
        std::vector<long> fill(atp.begin(), atp.end() - 1);
        for (size_t q = 0; q < ac.size(); ++q) {
            const long w = fill[ac[q]]++;
            atc[w] = ar[q];
            atv[w] = av[q];
        }
        std::vector<long> fillb(btp.begin(), btp.end() - 1);
        for (size_t q = 0; q < bc.size(); ++q) {
            const long w = fillb[bc[q]]++;
            btc[w] = br[q];
            btv[w] = bv[q];
        }
    }
    SG_TRY(upload(s, &s.d_at_ptr, atp));
    SG_TRY(upload(s, &s.d_at_col, atc));
    SG_TRY(upload(s, &s.d_at_val, atv));
    SG_TRY(upload(s, &s.d_bt_ptr, btp));
    SG_TRY(upload(s, &s.d_bt_col, btc));
    SG_TRY(upload(s, &s.d_bt_val, btv));
    SG_TRY(upload(s, &s.d_vol, vol));
    SG_TRY(upload(s, &s.d_ivol, ivol));
    SG_TRY(upload(s, &s.d_idiag, idiag));
    s.adj_ng = ng;
    s.adj_nnz = (long)atc.size();
    s.adj_ready = 1;
    return SG_OK;
}

constexpr int kAdjM = 20;          // FGMRES restart length
constexpr int kDotGroups = 1024;

double *adj_vec(sg_amr &s, int q) { return s.d_adj + (size_t)q * s.nblk * 512; }

int adj_alloc(sg_amr &s)
{
    if (s.d_adj) return SG_OK;
    // basis V_0..V_m, preconditioned Z_0..Z_{m-1}, w, the forward solve's psi and src, the adjoint's
    // right-hand side and solution, the preconditioner's A^T z and residual
    const size_t nv = 2 * kAdjM + 1 + 7;
    SG_HIP(hipMalloc((void **)&s.d_adj, sizeof(double) * nv * (size_t)s.nblk * 512));
    SG_HIP(hipMalloc((void **)&s.d_adj_part, sizeof(double) * (kDotGroups + 1)));
    // the ghost items' adjoint, the multipole partials (16 per workgroup of 256 items) and cbar
    const long ng = std::max(1L, s.adj_ng);
    SG_HIP(hipMalloc((void **)&s.d_adj_g, sizeof(double) * (ng + 16 * ((ng + 255) / 256) + 16)));
    s.adj_m = kAdjM;
    return SG_OK;
}

int adj_dot(sg_amr &s, const double *x, const double *y, double &out)
{
    const long n = (long)s.nblk * 512;
    const int G = (int)std::min<long>(kDotGroups, (n + 255) / 256);
    hipLaunchKernelGGL(k_dot_part, dim3(G), dim3(256), 0, s.stream, x, y, n, s.d_adj_part);
    SG_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(256), 0, s.stream, (const double *)s.d_adj_part, G, s.d_adj_part + kDotGroups);
    SG_CHECK_LAUNCH();
    SG_HIP(hipMemcpyAsync(&out, s.d_adj_part + kDotGroups, sizeof(double), hipMemcpyDeviceToHost, s.stream));
    SG_HIP(hipStreamSynchronize(s.stream));
    return SG_OK;
}

int at_apply(sg_amr &s, double *y, const double *x)
{
    return go(s, k_csr_spmv, (long)s.nblk * 512, y, x, (const long *)s.d_at_ptr, (const int *)s.d_at_col,
              (const double *)s.d_at_val, (long)s.nblk * 512);
}

// The preconditioner of A^T: z ~ A^-T v in two parts.  V A^-1 V^-1 (V: cell volumes) is A^-T
// wherever V A is symmetric, i.e. away from coarse-fine faces; it is applied through the forward
// solve with zero physical ghosts.  The transposed ghost and reflux terms make A^T act quite
// differently from A on oscillating modes along coarse-fine faces, where that correction is nearly
// blind, so damped Jacobi sweeps on A^T itself come before and after it (a two-level scheme: local
// smoothing, global correction).  pp: the forward solve's precision, tolerance and iteration cap.
constexpr int kAdjSweeps = 2;
constexpr double kAdjOmega = 0.6;

int adj_global(sg_amr &s, double *z, const double *v, const sg_params &pp)   // z = V A^-1 V^-1 v
{
    const long n = (long)s.nblk * 512;
    double *psi = adj_vec(s, 2 * kAdjM + 2), *src = adj_vec(s, 2 * kAdjM + 3);
    SG_TRY(go(s, k_scale_by, n, src, v, (const double *)s.d_ivol, 0, n));
    SG_HIP(hipMemsetAsync(psi, 0, sizeof(double) * n, s.stream));
    sg_stats st;
    s.homog_bc = 1;
    const int rc = pp.precision == SG_PREC_MIXED ? solve_mixed(s, psi, src, nullptr, pp, st) : solve(s, psi, src, nullptr, pp, st);
    s.homog_bc = 0;
    if (rc != SG_OK && rc != SG_ERR_NOT_CONVERGED) return rc;
    return go(s, k_scale_by, n, z, (const double *)psi, (const double *)s.d_vol, 0, n);
}

int adj_precond(sg_amr &s, double *z, const double *v, const sg_params &pp)
{
    const long n = (long)s.nblk * 512;
    double *atz = adj_vec(s, 2 * kAdjM + 6), *res = adj_vec(s, 2 * kAdjM + 7);
    auto sweep = [&]() -> int {
        SG_TRY(at_apply(s, atz, z));
        return go(s, k_jacobi, n, z, v, (const double *)atz, (const double *)s.d_idiag, kAdjOmega, n);
    };
    SG_HIP(hipMemsetAsync(z, 0, sizeof(double) * n, s.stream));
    for (int k = 0; k < kAdjSweeps; ++k) SG_TRY(sweep());
    SG_TRY(at_apply(s, atz, z));
    SG_TRY(go(s, k_axpby, n, atz, v, 1.0, -1.0, n));   // atz = v - A^T z
    SG_TRY(adj_global(s, res, atz, pp));
    SG_TRY(go(s, k_axpby, n, z, (const double *)res, 1.0, 1.0, n));
    for (int k = 0; k < kAdjSweeps; ++k) SG_TRY(sweep());
    return SG_OK;
}

// FGMRES(m) for A^T lam = r (lam starts from zero); tol on ||r - A^T lam|| / ||r||
int adj_fgmres(sg_amr &s, double *lam, const double *r, const sg_params &pp, double tol, int maxit, int &iters,
               double &relres)
{
    const long n = (long)s.nblk * 512;
    const int m = s.adj_m;
    double *w = adj_vec(s, 2 * m + 1);
    auto V = [&](int q) { return adj_vec(s, q); };
    auto Z = [&](int q) { return adj_vec(s, m + 1 + q); };
    double rn = 0.0;
    SG_TRY(adj_dot(s, r, r, rn));
    rn = std::sqrt(rn);
    SG_HIP(hipMemsetAsync(lam, 0, sizeof(double) * n, s.stream));
    iters = 0;
    relres = 0.0;
    if (rn == 0.0) return SG_OK;
    SG_HIP(hipMemcpyAsync(V(0), r, sizeof(double) * n, hipMemcpyDeviceToDevice, s.stream));   // residual of lam = 0
    double beta = rn;
    while (iters < maxit) {
        SG_TRY(go(s, k_axpby, n, V(0), (const double *)V(0), 1.0 / beta, 0.0, n));
        std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m, 0.0), sn(m, 0.0), g(m + 1, 0.0);
        g[0] = beta;
        int j = 0;
        for (; j < m && iters < maxit; ++j, ++iters) {
            SG_TRY(adj_precond(s, Z(j), V(j), pp));
            SG_TRY(at_apply(s, w, Z(j)));
            for (int i = 0; i <= j; ++i) {   // modified Gram-Schmidt
                double hij = 0.0;
                SG_TRY(adj_dot(s, w, V(i), hij));
                H[(size_t)i * m + j] = hij;
                SG_TRY(go(s, k_axpby, n, w, (const double *)V(i), -hij, 1.0, n));
            }
            double hn = 0.0;
            SG_TRY(adj_dot(s, w, w, hn));
            hn = std::sqrt(hn);
            H[(size_t)(j + 1) * m + j] = hn;
            if (hn > 0.0) SG_TRY(go(s, k_axpby, n, V(j + 1), (const double *)w, 1.0 / hn, 0.0, n));
            for (int i = 0; i < j; ++i) {   // earlier Givens rotations on column j
                const double a = H[(size_t)i * m + j], b = H[(size_t)(i + 1) * m + j];
                H[(size_t)i * m + j] = cs[i] * a + sn[i] * b;
                H[(size_t)(i + 1) * m + j] = -sn[i] * a + cs[i] * b;
            }
            const double a = H[(size_t)j * m + j], b = H[(size_t)(j + 1) * m + j];
            const double den = std::sqrt(a * a + b * b);
            cs[j] = den > 0.0 ? a / den : 1.0;
            sn[j] = den > 0.0 ? b / den : 0.0;
            H[(size_t)j * m + j] = den;
            H[(size_t)(j + 1) * m + j] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            if (std::fabs(g[j + 1]) <= tol * rn || hn == 0.0) {
                ++j;
                ++iters;
                break;
            }
        }
        // y = H^-1 g (upper triangular), lam += Z y
        std::vector<double> y(j, 0.0);
        for (int i = j - 1; i >= 0; --i) {
            double t = g[i];
            for (int q = i + 1; q < j; ++q) t -= H[(size_t)i * m + q] * y[q];
            y[i] = t / H[(size_t)i * m + i];
        }
        for (int i = 0; i < j; ++i) SG_TRY(go(s, k_axpby, n, lam, (const double *)Z(i), y[i], 1.0, n));
        // true residual, the next cycle's start
        SG_TRY(at_apply(s, w, lam));
        SG_TRY(go(s, k_axpby, n, w, r, 1.0, -1.0, n));   // w = r - A^T lam
        double res = 0.0;
        SG_TRY(adj_dot(s, w, w, res));
        beta = std::sqrt(res);
        relres = beta / rn;
        if (!std::isfinite(relres)) return SG_ERR_NONFINITE;
        if (relres <= tol) return SG_OK;
        SG_HIP(hipMemcpyAsync(V(0), w, sizeof(double) * n, hipMemcpyDeviceToDevice, s.stream));
    }
    return SG_ERR_NOT_CONVERGED;
}

// src_bar = (d psi / d src)^T psi_bar (dense [nblk][8^3] device arrays); pp: the adjoint's
// tolerance and iteration cap, and the precision of the forward solves that precondition it
int adjoint(sg_amr &s, double *src_bar, const double *psi_bar, const sg_params &pp, sg_stats &st)
{
    std::memset(&st, 0, sizeof(st));
    SG_REQUIRE(s.bc == SG_BC_SPHERE, "sg_amr_adjoint: the adjoint covers the sphere boundary");
    SG_REQUIRE(!s.zero_average, "sg_amr_adjoint: zero averaging is not part of the sphere solve");
    SG_TRY(build_adjoint(s));
    if (pp.precision == SG_PREC_MIXED) SG_TRY(ensure_mixed(s));
    SG_TRY(adj_alloc(s));
    const long n = (long)s.nblk * 512;
    double *a = adj_vec(s, 2 * kAdjM + 4), *lam = adj_vec(s, 2 * kAdjM + 5);
    const auto t0 = std::chrono::steady_clock::now();   // (the forward solves inside use the context's events)
    // C^T psi_bar: the closing c2p's transpose, coarse levels first
    SG_HIP(hipMemcpyAsync(a, psi_bar, sizeof(double) * n, hipMemcpyDeviceToDevice, s.stream));
    for (int l = 0; l + 1 < s.nlev; ++l) {
        const int b0 = s.lev_start[l], nb = s.lev_start[l + 1] - b0;
        SG_TRY(go(s, k_c2p_adjoint, (long)nb * 512, a, b0, nb, (const int *)s.d_covered, (const int *)s.d_child));
    }
    // the inner solve's precision and stopping for the preconditioner
    sg_params inner = pp;
    inner.tol = 1e-6;
    inner.maxit = 10;
    inner.stop_mode = SG_STOP_TRUE_RESIDUAL;
    int iters = 0;
    double relres = 0.0;
    const int rc = adj_fgmres(s, lam, a, inner, pp.tol, pp.maxit, iters, relres);
    st.iterations = iters;
    st.rel_l2 = relres;
    st.status = rc;
    if (rc != SG_OK && rc != SG_ERR_NOT_CONVERGED) return rc;
    // src_bar = lam - Mp^T B^T lam on the leaves, the mask, then C^T of the source's c2p
    double *gbar = s.d_adj_g, *part = s.d_adj_g + std::max(1L, s.adj_ng);
    SG_HIP(hipMemcpyAsync(src_bar, lam, sizeof(double) * n, hipMemcpyDeviceToDevice, s.stream));
    if (s.adj_ng > 0) {
        SG_TRY(go(s, k_csr_spmv, s.adj_ng, gbar, (const double *)lam, (const long *)s.d_bt_ptr, (const int *)s.d_bt_col,
                  (const double *)s.d_bt_val, s.adj_ng));
        SG_TRY(go(s, k_axpby, s.adj_ng, gbar, (const double *)gbar, -1.0, 0.0, s.adj_ng));
        const int G = (int)((s.adj_ng + 255) / 256);
        Geo Gm = geo(s);
        const int *faces = s.d_phys;
        hipLaunchKernelGGL(k_sphere_adj_part, dim3(G), dim3(256), 0, s.stream, (const double *)gbar, s.adj_ng, faces, Gm, part);
        SG_CHECK_LAUNCH();
        double *cbar = part + (size_t)16 * G;
        hipLaunchKernelGGL(k_sphere_adj_coef, dim3(1), dim3(256), 0, s.stream, (const double *)part, G, cbar);
        SG_CHECK_LAUNCH();
        const double dv = (s.h0 * s.h0) * s.h0;
        SG_TRY(go(s, k_sphere_adj_src, (long)s.nbase * 512, src_bar, s.nbase, Gm, dv, (const double *)cbar));
    }
    SG_TRY(go(s, k_sphere_mask_dense, n, src_bar, s.nblk, geo(s), sphere_rad2(s)));
    for (int l = 0; l + 1 < s.nlev; ++l) {
        const int b0 = s.lev_start[l], nb = s.lev_start[l + 1] - b0;
        SG_TRY(go(s, k_c2p_adjoint, (long)nb * 512, src_bar, b0, nb, (const int *)s.d_covered, (const int *)s.d_child));
    }
    SG_HIP(hipStreamSynchronize(s.stream));
    st.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}

} // namespace

extern "C" {

int sg_amr_create(sg_amr **ctx, int nblk, int nlev, const int *level, const int *ijk, const int *parent,
                  const int *child, const int *nbr, const int *base_dims, double h0, const double *x0, int bc)
{
    SG_REQUIRE(ctx && level && ijk && parent && child && nbr && base_dims && x0, "sg_amr_create: null argument");
    *ctx = nullptr;
    sg_amr *s = new sg_amr();
    int rc = create(s, nblk, nlev, level, ijk, parent, child, nbr, base_dims, h0, x0, bc);
    if (rc != SG_OK) {
        release(s);
        return rc;
    }
    *ctx = s;
    return SG_OK;
}

int sg_amr_destroy(sg_amr *ctx)
{
    release(ctx);
    return SG_OK;
}

int sg_amr_set_zero_average(sg_amr *ctx, int on)
{
    SG_REQUIRE(ctx, "sg_amr_set_zero_average: null context");
    ctx->zero_average = on ? 1 : 0;
    // the captured mixed iteration bakes the setting in (vmg_fas's zero average): recapture
    if (ctx->graph && ctx->graph_zero_average != ctx->zero_average) {
        (void)hipGraphExecDestroy(ctx->graph);
        ctx->graph = nullptr;
    }
    return SG_OK;
}

int sg_amr_solve_dev(sg_amr *ctx, double *psi, const double *src, double *g, const sg_params *p, sg_stats *st)
{
    SG_REQUIRE(ctx && psi && src && p && st, "sg_amr_solve_dev: null argument");
    SG_REQUIRE(p->precision == SG_PREC_FP64 || p->precision == SG_PREC_MIXED, "sg_amr_solve_dev: unknown precision %d",
               p->precision);
    SG_REQUIRE(p->maxit >= 0 && p->maxit <= 1000, "sg_amr_solve_dev: bad maxit %d", p->maxit);
    SG_REQUIRE(p->stop_mode == SG_STOP_REFERENCE || p->stop_mode == SG_STOP_TRUE_RESIDUAL,
               "sg_amr_solve_dev: bad stop_mode %d", p->stop_mode);
    SG_REQUIRE(p->tol >= 0.0, "sg_amr_solve_dev: bad tol");
    if (p->precision == SG_PREC_MIXED) return solve_mixed(*ctx, psi, src, g, *p, *st);
    return solve(*ctx, psi, src, g, *p, *st);
}

int sg_amr_solve_host(sg_amr *ctx, double *psi, const double *src, double *g, const sg_params *p, sg_stats *st)
{
    SG_REQUIRE(ctx && psi && src && p && st, "sg_amr_solve_host: null argument");
    sg_amr &s = *ctx;
    const size_t n = (size_t)s.nblk * 512;
    if (!s.d_dense) SG_HIP(hipMalloc((void **)&s.d_dense, sizeof(double) * 5 * n));
    double *dpsi = s.d_dense, *dsrc = dpsi + n, *dg = dsrc + n;
    SG_HIP(hipMemcpyAsync(dpsi, psi, sizeof(double) * n, hipMemcpyHostToDevice, s.stream));
    SG_HIP(hipMemcpyAsync(dsrc, src, sizeof(double) * n, hipMemcpyHostToDevice, s.stream));
    const int rc = sg_amr_solve_dev(ctx, dpsi, dsrc, g ? dg : nullptr, p, st);
    if (rc != SG_OK && rc != SG_ERR_NOT_CONVERGED && rc != SG_ERR_NONFINITE) return rc;
    SG_HIP(hipMemcpyAsync(psi, dpsi, sizeof(double) * n, hipMemcpyDeviceToHost, s.stream));
    if (g) SG_HIP(hipMemcpyAsync(g, dg, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s.stream));
    SG_HIP(hipStreamSynchronize(s.stream));
    return rc;
}

int sg_amr_set_array(sg_amr *ctx, int fmglev, int code, const double *host)
{
    SG_REQUIRE(ctx && host, "sg_amr_set_array: null argument");
    SG_REQUIRE(fmglev >= 0 && fmglev < kNfmg && code >= 0 && code < NCODE, "sg_amr_set_array: bad level/code");
    SG_HIP(hipMemcpyAsync(ctx->d.A[fmglev][code], host, sizeof(double) * ctx->nblk * lenof(fmglev), hipMemcpyHostToDevice,
                          ctx->stream));
    SG_HIP(hipStreamSynchronize(ctx->stream));
    return SG_OK;
}

int sg_amr_get_array(sg_amr *ctx, int fmglev, int code, double *host)
{
    SG_REQUIRE(ctx && host, "sg_amr_get_array: null argument");
    SG_REQUIRE(fmglev >= 0 && fmglev < kNfmg && code >= 0 && code < NCODE, "sg_amr_get_array: bad level/code");
    SG_HIP(hipMemcpyAsync(host, ctx->d.A[fmglev][code], sizeof(double) * ctx->nblk * lenof(fmglev), hipMemcpyDeviceToHost,
                          ctx->stream));
    SG_HIP(hipStreamSynchronize(ctx->stream));
    return SG_OK;
}

int sg_amr_get_fill0_level(sg_amr *ctx, int *v)
{
    SG_REQUIRE(ctx && v, "sg_amr_get_fill0_level: null argument");
    *v = ctx->fill0_level;
    return SG_OK;
}

int sg_amr_set_fill0_level(sg_amr *ctx, int v)
{
    SG_REQUIRE(ctx, "sg_amr_set_fill0_level: null context");
    ctx->fill0_level = v;
    return SG_OK;
}

int sg_amr_info(sg_amr *ctx, int *info, int n)
{
    SG_REQUIRE(ctx && info && n >= 0, "sg_amr_info: null argument");
    long vmg = 0;   // AMR levels whose vmg_ghostcell runs as one gather
    for (int l = 0; l < ctx->nlev; ++l)
        if (ctx->vfill_ok[l]) vmg |= 1L << l;
    long gfix = 0;   // FMG levels whose base-level fmg_ghostcell_fix runs as one gather
    for (int L = 1; L < kNfmg; ++L)
        if (ctx->gfix0_ok[L]) gfix |= 1L << L;
    const long v[6] = {(ctx->fused_ok[0] ? 1 : 0) | (ctx->fused_ok[1] ? 2 : 0), (long)ctx->graph_nodes, ctx->nblk,
                       (long)ctx->cf.size(), vmg, gfix};
    for (int q = 0; q < n; ++q) info[q] = q < 6 ? (int)v[q] : 0;
    return SG_OK;
}

int sg_amr_set_base_solver(sg_amr *ctx, int chip)
{
    SG_REQUIRE(ctx, "sg_amr_set_base_solver: null context");
    SG_REQUIRE(chip == 0 || chip == 1, "sg_amr_set_base_solver: mode must be 0 or 1, got %d", chip);
    ctx->base_chip = chip;
    return SG_OK;
}

int sg_amr_set_fp64_refresh(sg_amr *ctx, int fused)
{
    SG_REQUIRE(ctx, "sg_amr_set_fp64_refresh: null context");
    SG_REQUIRE(fused == 0 || fused == 1, "sg_amr_set_fp64_refresh: mode must be 0 or 1, got %d", fused);
    ctx->fp64_fused = fused;
    return SG_OK;
}

int sg_amr_adjoint_dev(sg_amr *ctx, double *src_bar, const double *psi_bar, const sg_params *p, sg_stats *st)
{
    SG_REQUIRE(ctx && src_bar && psi_bar && p && st, "sg_amr_adjoint_dev: null argument");
    SG_REQUIRE(p->precision == SG_PREC_FP64 || p->precision == SG_PREC_MIXED, "sg_amr_adjoint_dev: unknown precision %d",
               p->precision);
    SG_REQUIRE(p->maxit > 0 && p->maxit <= 10000, "sg_amr_adjoint_dev: bad maxit %d", p->maxit);
    SG_REQUIRE(p->tol > 0.0, "sg_amr_adjoint_dev: bad tol");
    return adjoint(*ctx, src_bar, psi_bar, *p, *st);
}

int sg_amr_adjoint_apply(sg_amr *ctx, double *y, const double *x)
{
    SG_REQUIRE(ctx && y && x, "sg_amr_adjoint_apply: null argument");
    SG_REQUIRE(ctx->bc == SG_BC_SPHERE, "sg_amr_adjoint_apply: the adjoint covers the sphere boundary");
    SG_TRY(build_adjoint(*ctx));
    SG_TRY(at_apply(*ctx, y, x));
    SG_HIP(hipStreamSynchronize(ctx->stream));
    return SG_OK;
}

int sg_amr_time_smoother(sg_amr *ctx, int prec, int reps, double *ms_per_sweep)
{
    SG_REQUIRE(ctx && ms_per_sweep && reps > 0 && reps <= 100000, "sg_amr_time_smoother: bad argument");
    SG_REQUIRE(prec == SG_PREC_FP64 || prec == SG_PREC_MIXED, "sg_amr_time_smoother: unknown precision %d", prec);
    sg_amr &s = *ctx;
    if (prec == SG_PREC_MIXED) SG_TRY(ensure_mixed(s));
    // the solve's FMG-level-0 colour passes: same-level faces read directly where the mixed solve's
    // fused mode does (FP32), from the stored ghosts in the plain FP64 path
    const int halo = prec == SG_PREC_MIXED ? s.fused_ok[0] : 0;
    SG_HIP(hipEventRecord(s.ev0, s.stream));
    for (int r = 0; r < reps; ++r)
        for (int ipass = 1; ipass <= 2; ++ipass) {
            if (prec == SG_PREC_MIXED) {   // (the compact form of U, as the fused mode keeps it)
                const bool cc = halo && s.cc_ok;
                SG_TRY((blk8<true, float>(s, s.blk8_nt[0], (float *)nullptr, cc ? s.u0c : s.f.A[0][C_U],
                                          s.f.A[0][C_RHS], ipass, s.f.ov[0], nullptr, halo, nullptr, nullptr, s.rhs0c,
                                          cc ? s.gs0 : nullptr)));
            } else {
                SG_TRY((blk8<true, double>(s, s.blk8_nt[2], (double *)nullptr, s.d.A[0][C_U], s.d.A[0][C_RHS], ipass,
                                           s.d.ov[0], nullptr, 0)));
            }
        }
    SG_HIP(hipEventRecord(s.ev1, s.stream));
    SG_HIP(hipEventSynchronize(s.ev1));
    float ms = 0.f;
    SG_HIP(hipEventElapsedTime(&ms, s.ev0, s.ev1));
    *ms_per_sweep = (double)ms / reps;
    return SG_OK;
}

int sg_amr_tile_info(sg_amr *ctx, int which, int *choice, float *ms, int nms)
{
    SG_REQUIRE(ctx && choice && (ms || nms == 0) && nms >= 0, "sg_amr_tile_info: null argument");
    SG_REQUIRE(which >= 0 && which < 4, "sg_amr_tile_info: unknown kernel %d", which);
    *choice = ctx->blk8_nt[which];
    for (int q = 0; q < nms; ++q) ms[q] = q < 3 ? ctx->blk8_ms[which][q] : 0.f;
    return SG_OK;
}

int sg_amr_op(sg_amr *ctx, int opc, int fmglev, int a, int b, int c, int flag, double *ret)
{
    SG_REQUIRE(ctx, "sg_amr_op: null context");
    return op(*ctx, opc, fmglev, a, b, c, flag, ret);
}

// ---- hydro steps: item tables ----
namespace {
int upload4(sg_amr &s, int4 **dptr, const std::vector<int4> &v)
{
    SG_HIP(hipMalloc((void **)dptr, std::max<size_t>(1, v.size()) * sizeof(int4)));
    if (!v.empty()) SG_HIP(hipMemcpy(*dptr, v.data(), v.size() * sizeof(int4), hipMemcpyHostToDevice));
    return SG_OK;
}

int build_hydro_tables(sg_amr &s)
{
    if (s.hydro_ready) return SG_OK;
    const int nl = s.nlev;
    std::vector<int4> hx, fg, rf;
    s.hx_off.assign(3 * nl + 1, 0);
    for (int l = 0; l < nl; ++l)
        for (int d = 0; d < 3; ++d) {
            s.hx_off[3 * l + d] = (int)hx.size();
            for (int b = s.lev_start[l]; b < s.lev_start[l + 1]; ++b)
                for (int lr = 0; lr < 2; ++lr) {
                    const int f = 2 * d + lr, m = s.nbr[6 * b + f];
                    if (m >= 0 && s.level[m] == l) {
                        hx.push_back(make_int4(b, lr, 0, m));
                    } else if (l > 0) {
                        const int g = s.nbr[6 * s.parent[b] + f];
                        if (g >= 0 && s.level[g] == l - 1) hx.push_back(make_int4(b, lr, 1 | (s.octant[b] << 1), g));
                    }
                }
        }
    s.hx_off[3 * nl] = (int)hx.size();
    s.fg_off.assign(nl + 1, 0);
    for (int l = 0; l < nl; ++l) {
        s.fg_off[l] = (int)fg.size();
        if (l > 0)
            for (int b = s.lev_start[l]; b < s.lev_start[l + 1]; ++b) fg.push_back(make_int4(b, s.parent[b], s.octant[b], 0));
    }
    s.fg_off[nl] = (int)fg.size();
    // refluxing: items in the reference's order, each in the round after the last one that
    // corrected any of its coarse cells
    s.rf_rounds.assign(nl, {});
    for (int l = 0; l + 1 < nl; ++l) {
        std::vector<int4> items;
        std::vector<int> round;
        std::vector<int> last;   // per (coarse block of level l, cell)
        const int c0 = s.lev_start[l], nc = s.lev_start[l + 1] - c0;
        last.assign((size_t)nc * 512, 0);
        int nround = 0;
        for (int b = s.lev_start[l + 1]; b < s.lev_start[l + 2]; ++b)
            for (int f = 0; f < 6; ++f) {
                const int m = s.nbr[6 * b + f];
                if (m >= 0 && s.level[m] == l + 1) continue;
                const int g = s.nbr[6 * s.parent[b] + f];
                if (g < 0 || s.level[g] != l) continue;
                const int d = f >> 1, lr = f & 1, o = s.octant[b];
                const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
                std::vector<size_t> cells;
                int r = 0;
                for (int c2 = 0; c2 < 4; ++c2)
                    for (int c1 = 0; c1 < 4; ++c1) {
                        int pc[3];
                        pc[d] = lr == 0 ? 7 : 0;
                        pc[ta] = c1 + ((o >> ta) & 1) * 4;
                        pc[tb] = c2 + ((o >> tb) & 1) * 4;
                        const size_t id = (size_t)(g - c0) * 512 + (size_t)((pc[2] * 8 + pc[1]) * 8 + pc[0]);
                        cells.push_back(id);
                        r = std::max(r, last[id]);
                    }
                ++r;
                for (size_t id : cells) last[id] = r;
                items.push_back(make_int4(b, f + 8 * o, g, 0));
                round.push_back(r);
                nround = std::max(nround, r);
            }
        std::vector<int> &off = s.rf_rounds[l];
        for (int r = 1; r <= nround; ++r) {
            off.push_back((int)rf.size());
            for (size_t q = 0; q < items.size(); ++q)
                if (round[q] == r) rf.push_back(items[q]);
        }
        off.push_back((int)rf.size());
    }
    SG_TRY(upload4(s, &s.d_hx, hx));
    SG_TRY(upload4(s, &s.d_fg, fg));
    SG_TRY(upload4(s, &s.d_rf, rf));
    s.hydro_ready = 1;
    return SG_OK;
}

int hydro_checks(sg_amr &s, const char *who)
{
    SG_REQUIRE(s.have_bm, "%s: no block memory layout (sg_amr_set_blockmem)", who);
    return SG_OK;
}

int is_hd9(const sg_blockmem &L) { return L.ncomp == 9; }
} // namespace

// ---- hydro block memory on the device ----
int sg_amr_set_blockmem(sg_amr *ctx, const sg_blockmem *lay, long long pool_slots, const int *slot_u,
                        const int *slot_u1, const int *slot_u2)
{
    SG_REQUIRE(ctx && lay && slot_u, "sg_amr_set_blockmem: null argument");
    sg_amr &s = *ctx;
    const sg_blockmem L = *lay;
    SG_REQUIRE(L.ng >= 1 && L.ng <= 8, "sg_amr_set_blockmem: ng = %d outside 1..8", L.ng);
    SG_REQUIRE(L.ncomp >= 9 && L.ncomp <= 4096, "sg_amr_set_blockmem: ncomp = %d (needs >= 9 labels)", L.ncomp);
    const long long W = 8 + 2 * L.ng;
    SG_REQUIRE(L.stride >= (long long)L.ncomp * W * W * W, "sg_amr_set_blockmem: stride %lld < ncomp (8 + 2 ng)^3",
               L.stride);
    const int lab[9] = {L.mrho, L.mvx, L.mvy, L.mvz, L.mp, L.mpsi, L.mgx, L.mgy, L.mgz};
    for (int a = 0; a < 9; ++a) {
        SG_REQUIRE(lab[a] >= 0 && lab[a] < L.ncomp, "sg_amr_set_blockmem: label %d = %d outside 0..ncomp-1", a, lab[a]);
        for (int c = 0; c < a; ++c) SG_REQUIRE(lab[a] != lab[c], "sg_amr_set_blockmem: labels %d and %d coincide", c, a);
    }
    SG_REQUIRE(pool_slots >= 1 && pool_slots <= (1LL << 31) && (double)pool_slots * (double)L.stride < 9.0e18,
               "sg_amr_set_blockmem: bad pool_slots %lld", pool_slots);
    const int *tabs[3] = {slot_u, slot_u1, slot_u2};
    int **dst[3] = {&s.d_slot_u, &s.d_slot_u1, &s.d_slot_u2};
    for (int q = 0; q < 3; ++q) {
        if (*dst[q]) (void)hipFree(*dst[q]);
        *dst[q] = nullptr;
    }
    s.have_bm = 0;
    std::vector<int> all;   // every slot named, to check that no two arrays share one
    for (int q = 0; q < 3; ++q) {
        if (!tabs[q]) continue;
        for (int b = 0; b < s.nblk; ++b)
            SG_REQUIRE(tabs[q][b] >= 0 && tabs[q][b] < pool_slots, "sg_amr_set_blockmem: slot table %d, block %d: %d",
                       q, b, tabs[q][b]);
        all.insert(all.end(), tabs[q], tabs[q] + s.nblk);
    }
    std::sort(all.begin(), all.end());
    SG_REQUIRE(std::adjacent_find(all.begin(), all.end()) == all.end(),
               "sg_amr_set_blockmem: a pool slot is named twice (each grid's current, predictor and past arrays "
               "need slots of their own)");
    for (int q = 0; q < 3; ++q)
        if (tabs[q]) SG_TRY(upload(s, dst[q], std::vector<int>(tabs[q], tabs[q] + s.nblk)));
    s.bml = L;
    s.have_bm = 1;
    return SG_OK;
}

int sg_amr_solve_blockmem_dev(sg_amr *ctx, double *pool, const sg_timeslice *ts, double pi4g, const sg_params *p,
                              sg_stats *st)
{
    SG_REQUIRE(ctx && pool && ts && p && st, "sg_amr_solve_blockmem_dev: null argument");
    sg_amr &s = *ctx;
    SG_REQUIRE(s.have_bm, "sg_amr_solve_blockmem_dev: no block memory layout (sg_amr_set_blockmem)");
    SG_REQUIRE(ts->source >= 0 && ts->source <= 3, "sg_amr_solve_blockmem_dev: unknown time slice %d", ts->source);
    SG_REQUIRE(!(ts->source == 1 || ts->source == 3) || s.d_slot_u1,
               "sg_amr_solve_blockmem_dev: time slice %d needs the predictor slots", ts->source);
    SG_REQUIRE(!(ts->source == 2 || ts->source == 3) || s.d_slot_u2,
               "sg_amr_solve_blockmem_dev: time slice %d needs the past slots", ts->source);
    SliceCoef C{};
    C.source = ts->source;
    for (int l = 0; l < s.nlev; ++l) {
        const double stp = ts->step_max, dstp = ts->dstep[l];
        C.fine[l] = ts->step[l] == stp ? 1 : 0;
        C.stp1[l] = stp - (ts->step[l] - ts->dstep[l]);
        C.stp2[l] = ts->step[l] - stp;
        C.dstp[l] = dstp;
    }
    const size_t n = (size_t)s.nblk * 512;
    if (!s.d_dense) SG_HIP(hipMalloc((void **)&s.d_dense, sizeof(double) * 5 * n));
    double *dpsi = s.d_dense, *dsrc = dpsi + n, *dg = dsrc + n;
    SG_TRY(go(s, k_bm_stage, (long)n, dpsi, dsrc, (const double *)pool, (const int *)s.d_slot_u,
              (const int *)s.d_slot_u1, (const int *)s.d_slot_u2, s.nblk, (const int *)s.d_level, s.bml, C, pi4g));
    const int rc = sg_amr_solve_dev(ctx, dpsi, dsrc, dg, p, st);
    if (rc != SG_OK && rc != SG_ERR_NOT_CONVERGED && rc != SG_ERR_NONFINITE) return rc;
    SG_TRY(go(s, k_bm_restore, (long)s.nblk * 1000, pool, (const int *)s.d_slot_u, s.nblk, s.bml,
              (const double *)s.d.A[0][C_U], (const double *)dg));
    SG_HIP(hipStreamSynchronize(s.stream));
    return rc;
}

int sg_amr_source_g_dev(sg_amr *ctx, double *pool, const double *dtime, double gamma)
{
    SG_REQUIRE(ctx && pool && dtime, "sg_amr_source_g_dev: null argument");
    sg_amr &s = *ctx;
    SG_REQUIRE(s.have_bm && s.d_slot_u1, "sg_amr_source_g_dev: no block memory layout with predictor slots");
    SG_REQUIRE(std::isfinite(gamma) && gamma != 1.0, "sg_amr_source_g_dev: bad gamma");
    LevelScalars dv{}, dt{};
    for (int l = 0; l < s.nlev; ++l) {
        const double h = s.hl[0][l];
        dv.v[l] = h * h * h;   // get_dv (grid.F90:405-409)
        dt.v[l] = dtime[l];
    }
    const long W = 8 + 2 * s.bml.ng;
    SG_TRY(go(s, k_source_g, (long)s.nblk * W * W * W, pool, (const int *)s.d_slot_u, (const int *)s.d_slot_u1,
              (const int *)s.d_level, s.nblk, s.bml, dv, dt, gamma));
    SG_HIP(hipStreamSynchronize(s.stream));
    return SG_OK;
}

int sg_amr_hydro_boundary_dev(sg_amr *ctx, double *pool, int level, int mode)
{
    SG_REQUIRE(ctx && pool, "sg_amr_hydro_boundary_dev: null argument");
    sg_amr &s = *ctx;
    SG_TRY(hydro_checks(s, "sg_amr_hydro_boundary_dev"));
    SG_REQUIRE(level >= 0 && level < s.nlev, "sg_amr_hydro_boundary_dev: level %d outside 0..%d", level, s.nlev - 1);
    SG_REQUIRE(mode >= 0 && mode <= 2, "sg_amr_hydro_boundary_dev: unknown step mode %d", mode);
    SG_REQUIRE(s.bml.ng >= 2 && s.bml.ng % 2 == 0, "sg_amr_hydro_boundary_dev: ng = %d (needs an even ng >= 2)",
               s.bml.ng);
    SG_REQUIRE(level == 0 || (s.d_slot_u1 && s.d_slot_u2),
               "sg_amr_hydro_boundary_dev: the parent level's pass needs the predictor and past slots");
    SG_TRY(build_hydro_tables(s));
    const long W = 8 + 2 * s.bml.ng, slab = (long)s.bml.ng * W * W * s.bml.ncomp;
    SG_REQUIRE(3 * slab <= 65535L * 256, "sg_amr_hydro_boundary_dev: ncomp too large for the launch shape");
    auto pass = [&](int l, int d, int ntab, int same_only) -> int {
        const int a = s.hx_off[3 * l + d], n = s.hx_off[3 * l + d + 1] - a;
        if (n == 0) return SG_OK;
        const dim3 grid((unsigned)n, (unsigned)((slab * ntab + 255) / 256));
        hipLaunchKernelGGL(k_hx_pass, grid, dim3(256), 0, s.stream, pool, (const int4 *)(s.d_hx + a), n,
                           (const int *)s.d_slot_u, (const int *)s.d_slot_u1, (const int *)s.d_slot_u2, ntab, same_only,
                           mode, s.bml, d);
        SG_CHECK_LAUNCH();
        return SG_OK;
    };
    if (level >= 1)
        for (int d = 0; d < 3; ++d) SG_TRY(pass(level - 1, d, 3, 1));
    for (int d = 0; d < 3; ++d) SG_TRY(pass(level, d, 1, 0));
    SG_HIP(hipStreamSynchronize(s.stream));
    return SG_OK;
}

int sg_amr_fg2cg_dev(sg_amr *ctx, double *pool, int level, double gamma)
{
    SG_REQUIRE(ctx && pool, "sg_amr_fg2cg_dev: null argument");
    sg_amr &s = *ctx;
    SG_TRY(hydro_checks(s, "sg_amr_fg2cg_dev"));
    SG_REQUIRE(is_hd9(s.bml), "sg_amr_fg2cg_dev: the ideal-gas restriction takes the nine-component layout");
    SG_REQUIRE(level >= 0 && level < s.nlev, "sg_amr_fg2cg_dev: level %d outside 0..%d", level, s.nlev - 1);
    SG_REQUIRE(std::isfinite(gamma) && gamma != 1.0, "sg_amr_fg2cg_dev: bad gamma");
    if (level < 1) return SG_OK;   // fg2cg.F90:25
    SG_TRY(build_hydro_tables(s));
    const int a = s.fg_off[level], n = s.fg_off[level + 1] - a;
    SG_TRY(go(s, k_fg2cg, (long)n * 64, pool, (const int4 *)(s.d_fg + a), n, (const int *)s.d_slot_u, s.bml, gamma));
    SG_HIP(hipStreamSynchronize(s.stream));
    return SG_OK;
}

int sg_amr_fluxcorrection_dev(sg_amr *ctx, double *pool, const double *fb, int level, const double *dtime, double gamma)
{
    SG_REQUIRE(ctx && pool && fb && dtime, "sg_amr_fluxcorrection_dev: null argument");
    sg_amr &s = *ctx;
    SG_TRY(hydro_checks(s, "sg_amr_fluxcorrection_dev"));
    SG_REQUIRE(is_hd9(s.bml), "sg_amr_fluxcorrection_dev: the ideal-gas refluxing takes the nine-component layout");
    SG_REQUIRE(std::isfinite(gamma) && gamma != 1.0, "sg_amr_fluxcorrection_dev: bad gamma");
    if (level < 0 || level >= s.nlev - 1) return SG_OK;   // reflux.F90:118-119
    SG_TRY(build_hydro_tables(s));
    const double hf = s.hl[0][level + 1], hc = s.hl[0][level];
    const double dsf = hf * hf, dsc = hc * hc, dv = hc * hc * hc;   // get_ds, get_dv (grid.F90:405-421)
    const std::vector<int> &off = s.rf_rounds[level];
    for (size_t r = 0; r + 1 < off.size(); ++r) {
        const int n = off[r + 1] - off[r];
        SG_TRY(go(s, k_reflux_round, (long)n * 16, pool, (const int4 *)(s.d_rf + off[r]), n, (const int *)s.d_slot_u, fb,
                  s.bml, dsf, dsc, dv, dtime[level], gamma));
    }
    SG_HIP(hipStreamSynchronize(s.stream));
    return SG_OK;
}

} // extern "C"
