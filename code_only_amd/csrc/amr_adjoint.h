// Kernels of the AMR adjoint solve (sphere boundary): the transpose of the composite operator as a
// CSR sparse matrix over the dense [block][8^3] index space, the transposes of the boundary
// multipole, of the source mask and of fmg_converge_c2p, and the vector operations of the Krylov
// solve.  FP64 throughout.  Included by amr.hip after amr_kernels.h.
#pragma once

namespace sg {
namespace amr {

// y = M x for a CSR matrix over n rows (one thread per row)
__global__ void k_csr_spmv(double *y, const double *x, const long *rowptr, const int *col, const double *val, long n)
{
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double s = 0.0;
    for (long q = rowptr[r]; q < rowptr[r + 1]; ++q) s += val[q] * x[col[q]];
    y[r] = s;
}

// y = a x + b y (y not read when b is zero)
__global__ void k_axpby(double *y, const double *x, double a, double b, long n)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) y[t] = b == 0.0 ? a * x[t] : a * x[t] + b * y[t];
}

// y = x * w (w per dense entry) or x / w
__global__ void k_scale_by(double *y, const double *x, const double *w, int divide, long n)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) y[t] = divide ? x[t] / w[t] : x[t] * w[t];
}

// one damped Jacobi sweep on A^T z = v: z += w idiag (v - atz), atz = A^T z (idiag zero where A^T
// has no row)
__global__ void k_jacobi(double *z, const double *v, const double *atz, const double *idiag, double w, long n)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) z[t] += w * idiag[t] * (v[t] - atz[t]);
}

// per-workgroup partial sums of x . y (256 threads), then one workgroup folds them in a fixed order
__global__ void __launch_bounds__(256) k_dot_part(const double *x, const double *y, long n, double *part)
{
    __shared__ double sh[256];
    double s = 0.0;
    for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long)gridDim.x * 256) s += x[t] * y[t];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(256) k_dot_final(const double *part, int G, double *out)
{
    __shared__ double sh[256];
    double s = 0.0;
    for (int q = threadIdx.x; q < G; q += 256) s += part[q];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// transpose of fmg_converge_c2p's linear restriction for the covered blocks of one level (block
// b0 .. b0 + nb): each covered cell's adjoint goes in eighths to the 2x2x2 child cells it averages,
// and the covered cell's own entry is cleared (c2p overwrote its value)
__global__ void k_c2p_adjoint(double *a, int b0, int nb, const int *covered, const int *child)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nb * 512) return;
    const int b = b0 + (int)(t >> 9), c = (int)(t & 511);
    if (!covered[b]) return;
    const int i = c & 7, j = (c >> 3) & 7, k = c >> 6;
    const int o = (i >= 4 ? 1 : 0) + (j >= 4 ? 2 : 0) + (k >= 4 ? 4 : 0);
    const int ch = child[8 * b + o];
    const double v = a[t] * 0.125;
    a[t] = 0.0;
    const int i0 = 2 * (i & 3), j0 = 2 * (j & 3), k0 = 2 * (k & 3);
    double *ac = a + (long)ch * 512;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) ac[(i0 + dx) + 8 * (j0 + dy) + 64 * (k0 + dz)] += v;
}

// The multipole boundary of fmg_boundary_phys_sphere (l <= 3, real form): a ghost value is
// sum_q Ysph_q(ghost) kappa_q sum_cells Zsph_q(cell) src(cell) dv over the base blocks.  These are
// the basis values of k_sphere_moments (Zsph) and k_sphere_ghost (Ysph, with its factors of 2 and
// the 2 sin cos pairs folded in), and kappa of k_sphere_coef.
__device__ inline void sphere_trig(double x, double y, double z, double &r, double &cost, double &sint, double &cosp,
                                   double &sinp)
{
    r = sqrt(x * x + y * y + z * z);
    const double rc = sqrt(x * x + y * y);
    cost = z / r;
    sint = rc / r;
    cosp = x / rc;
    sinp = y / rc;
}

__device__ inline void sphere_Z(double x, double y, double z, double Z[16])
{
    double r, ct, st, cp, sp;
    sphere_trig(x, y, z, r, ct, st, cp, sp);
    const double r2 = r * r, r3 = r2 * r, c2 = ct * ct, s2 = st * st, s3 = s2 * st, cp2 = cp * cp, sp2 = sp * sp;
    Z[0] = 1.0;
    Z[1] = r * ct;
    Z[2] = r * st * cp;
    Z[3] = r * st * sp;
    Z[4] = r2 * (3.0 * c2 - 1.0);
    Z[5] = r2 * st * ct * cp;
    Z[6] = r2 * st * ct * sp;
    Z[7] = r2 * s2 * (cp2 - sp2);
    Z[8] = r2 * s2 * 2.0 * sp * cp;
    Z[9] = r3 * ct * (5.0 * c2 - 3.0);
    Z[10] = r3 * st * (5.0 * c2 - 1.0) * cp;
    Z[11] = r3 * st * (5.0 * c2 - 1.0) * sp;
    Z[12] = r3 * ct * s2 * (cp2 - sp2);
    Z[13] = r3 * ct * s2 * 2.0 * sp * cp;
    Z[14] = r3 * s3 * cp * (cp2 - 3.0 * sp2);
    Z[15] = r3 * s3 * sp * (3.0 * cp2 - sp2);
}

__device__ inline void sphere_Y(double x, double y, double z, double Y[16])
{
    double r, ct, st, cp, sp;
    sphere_trig(x, y, z, r, ct, st, cp, sp);
    const double ri = 1.0 / r, r2i = ri * ri, r3i = r2i * ri, r4i = r2i * r2i;
    const double c2 = ct * ct, s2 = st * st, s3 = s2 * st, cp2 = cp * cp, sp2 = sp * sp;
    Y[0] = ri;
    Y[1] = r2i * ct;
    Y[2] = 2.0 * r2i * st * cp;
    Y[3] = 2.0 * r2i * st * sp;
    Y[4] = r3i * (3.0 * c2 - 1.0);
    Y[5] = 2.0 * r3i * st * ct * cp;
    Y[6] = 2.0 * r3i * st * ct * sp;
    Y[7] = 2.0 * r3i * s2 * (cp2 - sp2);
    Y[8] = 4.0 * r3i * s2 * sp * cp;
    Y[9] = r4i * ct * (5.0 * c2 - 3.0);
    Y[10] = 2.0 * r4i * st * (5.0 * c2 - 1.0) * cp;
    Y[11] = 2.0 * r4i * st * (5.0 * c2 - 1.0) * sp;
    Y[12] = 2.0 * r4i * ct * s2 * (cp2 - sp2);
    Y[13] = 4.0 * r4i * ct * s2 * sp * cp;
    Y[14] = 2.0 * r4i * s3 * cp * (cp2 - 3.0 * sp2);
    Y[15] = 2.0 * r4i * s3 * sp * (3.0 * cp2 - sp2);
}

// kappa_q of k_sphere_coef (coef_q = kappa_q A_q)
__device__ inline double sphere_kappa(int q)
{
    const double pi = 3.14159265358979323846;
    const double k[16] = {1.0 / 4, 1.0 / 4, 1.0 / 8, 1.0 / 8, 1.0 / 16, 3.0 / 8, 3.0 / 8, 3.0 / 32,
                          3.0 / 32, 1.0 / 16, 3.0 / 64, 3.0 / 64, 15.0 / 32, 15.0 / 32, 5.0 / 64, 5.0 / 64};
    return -k[q] / pi;
}

// position of ghost item t of the physical-face list (k_sphere_ghost's numbering: face entry
// t / 100, tangential (t % 10 - 1, (t / 10) % 10 - 1))
__device__ inline void ghost_pos(const Geo &G, const int *faces, long t, double &x, double &y, double &z)
{
    const int t1 = (int)(t % 10) - 1, t2 = (int)((t / 10) % 10) - 1, e = faces[t / 100];
    const int b = e / 6, f = e % 6, d = f / 2, ag = (f & 1) ? 8 : -1;
    const int ii = d == 0 ? ag : t1, jj = d == 1 ? ag : (d == 0 ? t1 : t2), kk = d == 2 ? ag : t2;
    x = G.x(b, 0, ii);
    y = G.x(b, 1, jj);
    z = G.x(b, 2, kk);
}

// per workgroup (256 items): partial sums over ghost items of Ysph_q(ghost) gbar(ghost), q = 0..15,
// at part[q * G + workgroup]
__global__ void __launch_bounds__(256) k_sphere_adj_part(const double *gbar, long ng, const int *faces, Geo G, double *part)
{
    __shared__ double sh[16][256];
    double acc[16];
    for (int q = 0; q < 16; ++q) acc[q] = 0.0;
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t < ng && gbar[t] != 0.0) {
        double x, y, z, Y[16];
        ghost_pos(G, faces, t, x, y, z);
        sphere_Y(x, y, z, Y);
        for (int q = 0; q < 16; ++q) acc[q] = Y[q] * gbar[t];
    }
    for (int q = 0; q < 16; ++q) sh[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int q = 0; q < 16; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x < 16) part[(long)threadIdx.x * gridDim.x + blockIdx.x] = sh[threadIdx.x][0];
}

// one workgroup: cbar_q = kappa_q sum of the partials (fixed order)
__global__ void __launch_bounds__(256) k_sphere_adj_coef(const double *part, int G, double *cbar)
{
    __shared__ double sh[256];
    for (int q = 0; q < 16; ++q) {
        double s = 0.0;
        for (int w = threadIdx.x; w < G; w += 256) s += part[(long)q * G + w];
        sh[threadIdx.x] = s;
        __syncthreads();
        for (int st = 128; st > 0; st >>= 1) {
            if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
            __syncthreads();
        }
        if (threadIdx.x == 0) cbar[q] = sphere_kappa(q) * sh[0];
        __syncthreads();
    }
}

// the moments' transpose: a[base cell] += dv sum_q Zsph_q(cell) cbar_q over the base blocks
__global__ void k_sphere_adj_src(double *a, int nbase, Geo G, double dv, const double *cbar)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nbase * 512) return;
    const int b = (int)(t >> 9), c = (int)(t & 511);
    double Z[16];
    sphere_Z(G.x(b, 0, c & 7), G.x(b, 1, (c >> 3) & 7), G.x(b, 2, c >> 6), Z);
    double s = 0.0;
    for (int q = 0; q < 16; ++q) s += Z[q] * cbar[q];
    a[t] += dv * s;
}

// the source mask's transpose (itself): zero outside the sphere of radius^2 rad2
__global__ void k_sphere_mask_dense(double *a, int nblk, Geo G, double rad2)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nblk * 512) return;
    const int b = (int)(t >> 9), c = (int)(t & 511);
    const double x = G.x(b, 0, c & 7), y = G.x(b, 1, (c >> 3) & 7), z = G.x(b, 2, c >> 6);
    if (x * x + y * y + z * z >= rad2) a[t] = 0.0;
}

} // namespace amr
} // namespace sg
