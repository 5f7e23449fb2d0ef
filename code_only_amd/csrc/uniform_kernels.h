// Uniform periodic-grid stencil kernels (gfx950).  Dense n^3 arrays, index (k*n + j)*n + i, with
// the periodic ghost layer of the reference (fmg_boundary_phys_periodic.F90:71-104) replaced by
// index wrap.  Arithmetic follows the reference's operand order; the library is compiled with
// -ffp-contract=off so the FP64 instantiations round exactly like the CPU restatement.
#pragma once
#include "sg_common.h"

namespace sg {

__device__ __forceinline__ int wrap_p(int i, int n) { return (i + 1 == n) ? 0 : i + 1; }
__device__ __forceinline__ int wrap_m(int i, int n) { return (i == 0) ? n - 1 : i - 1; }
// Coarse ghost index -1..n mapped to its periodic image.
__device__ __forceinline__ int wrap_g(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// floor(i/2) (config.h:249 IJKC with origin 0).
__device__ __forceinline__ int ijkc(int i) { return (i >= 0) ? (i >> 1) : -((1 - i) >> 1); }

template <typename T> struct Consts;
template <> struct Consts<double> {
    static __device__ __forceinline__ double sixth() { return 1.0 / 6.0; }
};
template <> struct Consts<float> {
    static __device__ __forceinline__ float sixth() { return 1.0f / 6.0f; }
};

// ---- per-cell arithmetic shared by every kernel (plain, fused, single-workgroup) -------------
// Operand order is the reference's: x+, x-, y+, y-, z+, z- neighbours, then -6u, then the source.

// mg_poisson.F90:60-66: returns the updated value, *resh2 = the local h^2-weighted residual.
template <typename T>
__device__ __forceinline__ T relax_value(T uc, T xp, T xm, T yp, T ym, T zp, T zm, T f, T h2, T *resh2)
{
    T s = xp + xm;
    s = s + yp;
    s = s + ym;
    s = s + zp;
    s = s + zm;
    s = s - T(6) * uc;
    s = s - h2 * f;
    *resh2 = s;
    return uc + Consts<T>::sixth() * s;
}

// mg_poisson.F90:20-24: res = -h2i * (sum6 - 6u) + rhs.
template <typename T>
__device__ __forceinline__ T resid_value(T uc, T xp, T xm, T yp, T ym, T zp, T zm, T f, T h2i)
{
    T s = xp + xm;
    s = s + yp;
    s = s + ym;
    s = s + zp;
    s = s + zm;
    s = s - T(6) * uc;
    return -(h2i * s) + f;
}

// restriction.F90:139-143: (000 + 100 + 010 + 001 + 110 + 101 + 011 + 111) * 1/8.
template <typename T>
__device__ __forceinline__ T restrict_value(T a000, T a100, T a010, T a001, T a110, T a101, T a011, T a111)
{
    T s = a000 + a100;
    s = s + a010;
    s = s + a001;
    s = s + a110;
    s = s + a101;
    s = s + a011;
    s = s + a111;
    return s * T(0.125);
}

// interpolation.F90:236-241: (27 c + 9 (x + y + z) + 3 (yz + xz + xy) + xyz) / 64, where x, y, z
// are the face neighbours toward the fine cell, etc.
template <typename T>
__device__ __forceinline__ T trilinear_value(T c, T x, T y, T z, T yz, T xz, T xy, T xyz)
{
    T b = x + y;
    b = b + z;
    T e = yz + xz;
    e = e + xy;
    T s = T(27) * c + T(9) * b;
    s = s + T(3) * e;
    s = s + xyz;
    return s / T(64);
}

// Dense periodic n^3 grids with n = 2^lg: cell c = (k << 2lg) | (j << lg) | i, neighbours by mask.

// mg_poisson.F90:55-74, one colour (ipass 1 = cells with i+j+k odd, 2 = even).  One thread per
// colour cell (n = 1: the single cell belongs to the second pass).  resmax (nullable) receives
// max |resh2| by one atomic max per block.
template <typename T>
__global__ void __launch_bounds__(256) k_relax_color(T *__restrict__ u, const T *__restrict__ rhs,
                                                     int lg, T h2, int ipass, double *resmax)
{
    const int n = 1 << lg, m = n - 1;
    const int lgh = lg > 0 ? lg - 1 : 0;
    const int total = lg > 0 ? (1 << (3 * lg - 1)) : 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double amax = 0.0;
    if (t < total && (lg > 0 || ipass == 2)) {
        const int ih = t & ((1 << lgh) - 1);
        const int j = (t >> lgh) & m, k = t >> (lgh + lg);
        const int i = lg > 0 ? 2 * ih + ((j + k + ipass) & 1) : 0;
        const int row = (k << (2 * lg)) | (j << lg);
        const int c = row | i;
        T s;
        u[c] = relax_value<T>(u[c], u[row | ((i + 1) & m)], u[row | ((i - 1) & m)],
                              u[(k << (2 * lg)) | (((j + 1) & m) << lg) | i],
                              u[(k << (2 * lg)) | (((j - 1) & m) << lg) | i],
                              u[(((k + 1) & m) << (2 * lg)) | (j << lg) | i],
                              u[(((k - 1) & m) << (2 * lg)) | (j << lg) | i], rhs[c], h2, &s);
        amax = fabs((double)s);
    }
    if (resmax) block_atomic_max(resmax, amax);
}

// mg_poisson.F90:17-27: res = -(h2i * (sum6 - 6u)) + rhs, computed in Tin, stored as Tout.
// Grid-stride over the cells; pmax/psum (nullable): per-block partial max|res| and sum res^2 for a
// deterministic two-stage norm (fixed grid, fixed per-thread order).
template <typename Tin, typename Tout>
__global__ void __launch_bounds__(256) k_resid(Tout *__restrict__ res, const Tin *__restrict__ u,
                                               const Tin *__restrict__ rhs, int lg, Tin h2i,
                                               double *pmax, double *psum)
{
    const int n = 1 << lg, m = n - 1, n3 = 1 << (3 * lg);
    double mx = 0.0, sm = 0.0;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n3; c += gridDim.x * blockDim.x) {
        const int i = c & m, j = (c >> lg) & m, k = c >> (2 * lg);
        const int row = c - i;
        const Tin r = resid_value<Tin>(u[c], u[row | ((i + 1) & m)], u[row | ((i - 1) & m)],
                                       u[(k << (2 * lg)) | (((j + 1) & m) << lg) | i],
                                       u[(k << (2 * lg)) | (((j - 1) & m) << lg) | i],
                                       u[(((k + 1) & m) << (2 * lg)) | (j << lg) | i],
                                       u[(((k - 1) & m) << (2 * lg)) | (j << lg) | i], rhs[c], h2i);
        res[c] = (Tout)r;
        mx = fmax(mx, fabs((double)r));
        sm += (double)r * (double)r;
    }
    if (pmax) {
        block_max_sum(mx, sm);
        if (threadIdx.x == 0) { pmax[blockIdx.x] = mx; psum[blockIdx.x] = sm; }
    }
}

// Outer FP64 step of the mixed solve, fused: pout = pin + e (mg.F90:87) and
// r = f - L pout (mg_poisson.F90:17-27) stored as FP32 for the next inner solve, with the norm
// partials of r.  Neighbour values are formed as pin + e, exactly the values pout receives, so the
// result equals the separate update and residual kernels bit for bit; pin and pout must differ.
// UPD = false: r = f - L pin only, and pout (when given) a copy of pin.  With fpmax, the partials
// of max|f| and sum f^2 as well.  With rc, the FP32 residual is also restricted into rc (the
// next level's right-hand side, restriction.F90:132-147 operand order, as k_restrict forms it):
// each plane's FP32 residuals go to one of four LDS planes, and the restriction of planes
// (k - 2, k - 1) runs at plane k, after its barrier.  A workgroup owns a 64 x 8 patch of (i, j) columns and marches
// through zc planes: each thread forms its cell of the next plane (and, on the patch edge, one
// halo cell in x and/or y) while the current plane sits in a double-buffered LDS plane with a
// one-cell halo, so every value is loaded from memory once and one barrier per plane suffices; the
// z neighbours and the next plane's source value stay in registers.  n = 2^lg >= 64.  Patches of
// 64 x 8 columns, or 256 x 4 when n % 256 == 0: an x-halo cell costs a whole cache line of pin
// and of e, so patches as wide as the grid (whose x neighbours wrap into the patch) trade those
// for a second y-halo row per four.
template <bool UPD, int TX, int TY>
__global__ void __launch_bounds__(TX * TY) k_update_resid(const double *__restrict__ pin, const float *__restrict__ e,
                                                          double *__restrict__ pout, const double *__restrict__ f,
                                                          float *__restrict__ r32, int lg, double h2i, int zc,
                                                          double *pmax, double *psum, float *__restrict__ rc,
                                                          double *fpmax, double *fpsum)
{
    constexpr int SP = TX + 2;
    static_assert(TY % 2 == 0 && TX % 2 == 0 && TX * TY <= 1024, "update+residual patch");
    __shared__ double sp[2][(TY + 2) * SP];
    __shared__ float sr[4][TY * TX];
    const int n = 1 << lg, m = n - 1;
    const int tx = (int)threadIdx.x % TX, ty = (int)threadIdx.x / TX;
    // XCD-aware patch numbering: the hardware deals consecutive workgroups round-robin to the 8
    // XCDs; renumbered, each XCD takes a contiguous run of patches (neighbours in y), so the halo
    // rows a patch loads are mostly lines its XCD's L2 already holds
    const int gx = (int)gridDim.x, gy = (int)gridDim.y, nbk = gx * gy * (int)gridDim.z;
    const int hb = ((int)blockIdx.z * gy + (int)blockIdx.y) * gx + (int)blockIdx.x;
    const int vb = nbk % 8 == 0 ? (hb % 8) * (nbk / 8) + hb / 8 : hb;
    const int pbx = vb % gx, pby = (vb / gx) % gy, pbz = vb / (gx * gy);
    const int i = pbx * TX + tx, j = pby * TY + ty;
    const int z0 = pbz * zc;
    auto at = [&](int ii, int jj, int kk) { return (kk << (2 * lg)) | (jj << lg) | ii; };
    // pin, e, f and pout stream through once per outer iteration: non-temporal, so that the
    // last-level cache keeps r32, which the next inner solve reads first
    auto P = [&](int q) -> double {
        return UPD ? __builtin_nontemporal_load(pin + q) + (double)__builtin_nontemporal_load(e + q)
                   : __builtin_nontemporal_load(pin + q);
    };
    // this thread's halo cells: x (left edge of the patch: i - 1, right edge: i + 1), y likewise
    const bool hx = tx == 0 || tx == TX - 1, hy = ty == 0 || ty == TY - 1;
    const int ih = tx == 0 ? (i - 1) & m : (i + 1) & m, jh = ty == 0 ? (j - 1) & m : (j + 1) & m;
    const int sx = tx == 0 ? (ty + 1) * SP : (ty + 1) * SP + TX + 1;  // its slot in the LDS plane
    const int sy = ty == 0 ? tx + 1 : (TY + 1) * SP + tx + 1;
    const int sc = (ty + 1) * SP + tx + 1;
    double zm = P(at(i, j, (z0 - 1) & m)), zcen = P(at(i, j, z0));
    double vx = hx ? P(at(ih, j, z0)) : 0.0, vy = hy ? P(at(i, jh, z0)) : 0.0;
    double fc = __builtin_nontemporal_load(f + at(i, j, z0));
    double mx = 0.0, sm = 0.0, fmx = 0.0, fsm = 0.0;
    // restriction of the FP32 residuals of planes (k0, k0 + 1) of the patch into rc
    const int lgc = lg - 1;
    auto restrict_pair = [&](int k0) {
        const int t = (int)threadIdx.x;
        if (t < (TX / 2) * (TY / 2)) {
            const int cx = t % (TX / 2), cy = t / (TX / 2);
            const float *a = sr[k0 & 3] + 2 * cy * TX + 2 * cx, *b = sr[(k0 + 1) & 3] + 2 * cy * TX + 2 * cx;
            const int ic = pbx * (TX / 2) + cx, jc = pby * (TY / 2) + cy, kc = (k0 & m) >> 1;
            rc[(kc << (2 * lgc)) | (jc << lgc) | ic] =
                restrict_value<float>(a[0], a[1], a[TX], b[0], a[TX + 1], b[1], b[TX], b[TX + 1]);
        }
    };
    for (int k = z0; k < z0 + zc; ++k) {
        double *s = sp[k & 1];
        s[sc] = zcen;
        if (hx) s[sx] = vx;
        if (hy) s[sy] = vy;
        // the next plane's values, loaded while this plane is used
        const int kn = (k + 1) & m;
        const double zp = P(at(i, j, kn));
        double vxn = 0.0, vyn = 0.0, fn = 0.0;
        if (k + 1 < z0 + zc) {
            if (hx) vxn = P(at(ih, j, kn));
            if (hy) vyn = P(at(i, jh, kn));
            fn = __builtin_nontemporal_load(f + at(i, j, kn));
        }
        const int c = at(i, j, k);
        __syncthreads();
        const double rv = resid_value<double>(zcen, s[sc + 1], s[sc - 1], s[sc + SP], s[sc - SP], zp, zm, fc, h2i);
        r32[c] = (float)rv;
        if (pout) __builtin_nontemporal_store(zcen, pout + c);
        if (fpmax) {
            fmx = fmax(fmx, fabs(fc));
            fsm += fc * fc;
        }
        if (rc) {
            sr[k & 3][ty * TX + tx] = (float)rv;
            if (((k - z0) & 1) == 0 && k >= z0 + 2) restrict_pair(k - 2);
        }
        mx = fmax(mx, fabs(rv));
        sm += rv * rv;
        zm = zcen;
        zcen = zp;
        vx = vxn;
        vy = vyn;
        fc = fn;
    }
    if (rc) {
        __syncthreads();
        restrict_pair(z0 + zc - 2);
    }
    const int b = vb;
    block_max_sum(mx, sm);
    if (threadIdx.x == 0) {
        pmax[b] = mx;
        psum[b] = sm;
    }
    if (fpmax) {
        block_max_sum(fmx, fsm);
        if (threadIdx.x == 0) {
            fpmax[b] = fmx;
            fpsum[b] = fsm;
        }
    }
}

// Second stage of the norm: one block reduces np partials in a fixed order -> out[0]=max, out[1]=sum.
// With host (pinned host memory): host[0..1] = out[0..1] and host[2..4] = out[2..4] as well, so
// the host reads the five norm values without a separate copy (out[2..4] are final by then).
static __global__ void __launch_bounds__(1024) k_norm_finalize(const double *pmax, const double *psum, int np,
                                                        double *out, double *host = nullptr)
{
    double mx = 0.0, sm = 0.0;
    for (int q = threadIdx.x; q < np; q += blockDim.x) { mx = fmax(mx, pmax[q]); sm += psum[q]; }
    block_max_sum(mx, sm);
    if (threadIdx.x == 0) {
        out[0] = mx;
        out[1] = sm;
        if (host) {
            host[0] = mx;
            host[1] = sm;
            host[2] = out[2];
            host[3] = out[3];
            host[4] = out[4];
        }
    }
}

// restriction.F90:132-147, one thread per coarse cell (nc = 2^lgc).
template <typename T>
__global__ void __launch_bounds__(256) k_restrict(T *__restrict__ uc, const T *__restrict__ uf, int lgc)
{
    const int nc = 1 << lgc, c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (1 << (3 * lgc))) return;
    const int ic = c & (nc - 1), jc = (c >> lgc) & (nc - 1), kc = c >> (2 * lgc);
    const int lgf = lgc + 1;
    const int i = 2 * ic, j = 2 * jc, k = 2 * kc;
    auto F = [&](int a, int b, int d) { return uf[((k + d) << (2 * lgf)) | ((j + b) << lgf) | (i + a)]; };
    uc[c] = restrict_value<T>(F(0, 0, 0), F(1, 0, 0), F(0, 1, 0), F(0, 0, 1), F(1, 1, 0), F(1, 0, 1), F(0, 1, 1),
                              F(1, 1, 1));
}

// interpolation.F90:226-245 trilinear prolongation of a periodic coarse grid (nf = 2^lgf); ADD
// selects uf = uf + P uc (mg_addint, mg.F90:727-731) instead of uf = P uc.
template <typename T, bool ADD>
__global__ void __launch_bounds__(256) k_interp_linear(T *__restrict__ uf, const T *__restrict__ uc, int lgf)
{
    const int nf = 1 << lgf, c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (1 << (3 * lgf))) return;
    const int lgc = lgf - 1, nc = nf >> 1;
    const int i = c & (nf - 1), jf = (c >> lgf) & (nf - 1), kf = c >> (2 * lgf);
    const int ic = i >> 1, jc = jf >> 1, kc = kf >> 1;
    const int ic1 = (ic + 2 * (i & 1) - 1) & (nc - 1);
    const int jc1 = (jc + 2 * (jf & 1) - 1) & (nc - 1);
    const int kc1 = (kc + 2 * (kf & 1) - 1) & (nc - 1);
    auto C = [&](int a, int b, int d) { return uc[(d << (2 * lgc)) | (b << lgc) | a]; };
    const T v = trilinear_value<T>(C(ic, jc, kc), C(ic1, jc, kc), C(ic, jc1, kc), C(ic, jc, kc1), C(ic, jc1, kc1),
                                   C(ic1, jc, kc1), C(ic1, jc1, kc), C(ic1, jc1, kc1));
    if (ADD) uf[c] = uf[c] + v;
    else uf[c] = v;
}

// One axis of the tricubic stencil (interpolation.F90:106-127): coarse indices shifted to stay
// inside the ghosted coarse range lo..hi, and the four Lagrange weights, in FP64 as there.
__device__ __forceinline__ void cubic_stencil(int f, int lo, int hi, int idx[4], double a[4])
{
    const int cp = ijkc(f);
    const int cn = cp + 2 * (f & 1) - 1;
    const int c1 = cp < cn ? cp : cn;
    const int c2 = cp < cn ? cn : cp;
    int c0 = 2 * c1 - c2, c3 = 2 * c2 - c1;
    int nshift = 0;
    if (c3 > hi) nshift = c3 - hi;
    if (c0 < lo) nshift = c0 - lo;
    c0 -= nshift;
    const int d1 = c1 - nshift, d2 = c2 - nshift;
    c3 -= nshift;
    const double xc = cp + 0.25 * (double)(2 * (f & 1) - 1);
    a[0] = (xc - d1) * (xc - d2) * (xc - c3) / (double)((c0 - d1) * (c0 - d2) * (c0 - c3));
    a[1] = (xc - c0) * (xc - d2) * (xc - c3) / (double)((d1 - c0) * (d1 - d2) * (d1 - c3));
    a[2] = (xc - c0) * (xc - d1) * (xc - c3) / (double)((d2 - c0) * (d2 - d1) * (d2 - c3));
    a[3] = (xc - c0) * (xc - d1) * (xc - d2) / (double)((c3 - c0) * (c3 - d1) * (c3 - d2));
    idx[0] = c0; idx[1] = d1; idx[2] = d2; idx[3] = c3;
}

// interpolation.F90:98-217 tricubic prolongation of a periodic coarse grid (nc >= 2): for each
// fine cell the x pass (bufx), y pass (bufxy) and z pass are evaluated in the same operand order
// as the reference's three separable sweeps, so the result is bitwise that of the sweeps.
template <typename T>
__global__ void __launch_bounds__(256) k_interp_cubic(T *__restrict__ uf, const T *__restrict__ uc, int lgf)
{
    const int nf = 1 << lgf, c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (1 << (3 * lgf))) return;
    const int nc = nf >> 1;
    const int i = c & (nf - 1), jf = (c >> lgf) & (nf - 1), kf = c >> (2 * lgf);
    int ix[4], iy[4], iz[4];
    double ax[4], ay[4], az[4];
    cubic_stencil(i, -1, nc, ix, ax);
    cubic_stencil(jf, -1, nc, iy, ay);
    cubic_stencil(kf, -1, nc, iz, az);
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
        for (int jq = 0; jq < 4; ++jq) {
            const T *row = uc + ((long)iz[kq] * nc + iy[jq]) * nc;
            T s = wx[0] * row[ix[0]] + wx[1] * row[ix[1]];
            s = s + wx[2] * row[ix[2]];
            s = s + wx[3] * row[ix[3]];
            by[jq] = s;
        }
        T s = wy[0] * by[0] + wy[1] * by[1];
        s = s + wy[2] * by[2];
        s = s + wy[3] * by[3];
        bz[kq] = s;
    }
    T s = wz[0] * bz[0] + wz[1] * bz[1];
    s = s + wz[2] * bz[2];
    s = s + wz[3] * bz[3];
    uf[c] = s;
}

template <typename T>
__global__ void __launch_bounds__(256) k_fill(T *__restrict__ p, long len, T v)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < len) p[c] = v;
}

// mg.F90:87 psi = psi + u, with u in the inner solve's precision.
template <typename Tu>
__global__ void __launch_bounds__(256) k_update(double *__restrict__ psi, const Tu *__restrict__ u, long len)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < len) psi[c] = psi[c] + (double)u[c];
}

// fmg_poisson.F90:29-31: g_d = -(u(+1) - u(-1)) * hi * 0.5, periodic.
static __global__ void __launch_bounds__(256) k_gradient(double *__restrict__ g, const double *__restrict__ u, int lg,
                                                  double hi)
{
    const int n = 1 << lg, n3 = 1 << (3 * lg), c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n3) return;
    const int i = c & (n - 1), j = (c >> lg) & (n - 1), k = c >> (2 * lg);
    const int row = (k * n + j) * n;
    // component offsets in 64 bits: 2 n^3 exceeds INT_MAX at n = 1024
    const size_t o1 = (size_t)n3, o2 = 2 * (size_t)n3;
    g[c] = -(u[row + wrap_p(i, n)] - u[row + wrap_m(i, n)]) * hi * 0.5;
    g[o1 + c] = -(u[(k * n + wrap_p(j, n)) * n + i] - u[(k * n + wrap_m(j, n)) * n + i]) * hi * 0.5;
    g[o2 + c] = -(u[(wrap_p(k, n) * n + j) * n + i] - u[(wrap_m(k, n) * n + j) * n + i]) * hi * 0.5;
}

// ---- adjoint helpers (periodic grid, n = 2^lg) ----
// Sampled centred gradient of phi at cells[p] (linear index (k n + j) n + i):
// out[3p + d] = (phi(c + e_d) - phi(c - e_d)) / h * 0.5, the negative of fmg_psi2g's g
// (fmg_poisson.F90:29-31).
static __global__ void k_grad_sample(const double *__restrict__ u, const int *__restrict__ cells, int npts, int lg, double hi,
                              double *__restrict__ out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npts) return;
    const int n = 1 << lg, c = cells[p];
    const int i = c & (n - 1), j = (c >> lg) & (n - 1), k = c >> (2 * lg);
    const int row = (k * n + j) * n;
    out[3 * p + 0] = (u[row + wrap_p(i, n)] - u[row + wrap_m(i, n)]) * hi * 0.5;
    out[3 * p + 1] = (u[(k * n + wrap_p(j, n)) * n + i] - u[(k * n + wrap_m(j, n)) * n + i]) * hi * 0.5;
    out[3 * p + 2] = (u[(wrap_p(k, n) * n + j) * n + i] - u[(wrap_m(k, n) * n + j) * n + i]) * hi * 0.5;
}

// Transpose of k_grad_sample: out += sum_p sum_d w[3p + d] (delta(c_p + e_d) - delta(c_p - e_d)) / h * 0.5
// (scatter with atomics; out must be cleared by the caller).
static __global__ void k_grad_sample_adjoint(const int *__restrict__ cells, const double *__restrict__ w, int npts, int lg,
                                      double hi, double *__restrict__ out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npts) return;
    const int n = 1 << lg, c = cells[p];
    const int i = c & (n - 1), j = (c >> lg) & (n - 1), k = c >> (2 * lg);
    const int row = (k * n + j) * n;
    const double wx = w[3 * p] * hi * 0.5, wy = w[3 * p + 1] * hi * 0.5, wz = w[3 * p + 2] * hi * 0.5;
    atomicAdd(out + row + wrap_p(i, n), wx);
    atomicAdd(out + row + wrap_m(i, n), -wx);
    atomicAdd(out + (k * n + wrap_p(j, n)) * n + i, wy);
    atomicAdd(out + (k * n + wrap_m(j, n)) * n + i, -wy);
    atomicAdd(out + (wrap_p(k, n) * n + j) * n + i, wz);
    atomicAdd(out + (wrap_m(k, n) * n + j) * n + i, -wz);
}

// Mean removal (the projection P onto zero-mean fields): per-workgroup partial sums over a
// grid-stride share, then every thread of k_sub_mean folds the partials in the same order.
constexpr int kMeanGroups = 256;
static __global__ void k_sum_part(const double *__restrict__ a, long len, double *__restrict__ part)
{
    double s = 0.0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < len; t += (long)gridDim.x * blockDim.x) s += a[t];
    double m = 0.0;
    block_max_sum(m, s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

static __global__ void k_sub_mean(const double *__restrict__ a, double *__restrict__ out, long len, const double *__restrict__ part,
                           int nparts)
{
    double s = 0.0;
    for (int q = 0; q < nparts; ++q) s += part[q];
    const double mean = s / (double)len;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < len; t += (long)gridDim.x * blockDim.x)
        out[t] = a[t] - mean;
}

} // namespace sg
