// Exact solve of the gathered base-grid problem (the problem vmg_slvsml hands to mg_cycle,
// vmg.F90:387-432) for the FP32 inner solve of the mixed-precision path, in one workgroup.
//
// The base grid holds n^3 cells (n = 2 x base blocks per side, 32 for a 16^3-block base) with
// either zero ghosts (sphere boundary: mg_boundary_fill0, mg.F90:735-750) or periodic ghosts.  Its
// 7-point operator separates: along each axis the second difference with zero ghosts is
// diagonalised by the DST-I basis sqrt(2/(n+1)) sin(pi (m+1)(k+1)/(n+1)), the periodic one by the
// Hartley basis (cos + sin)(2 pi m k / n) / sqrt(n).  Both bases are symmetric and orthonormal, so
// the same three axis passes serve as forward and inverse transform:
//     u = Q (Lambda^-1 (Q f)),  Lambda = lam_x + lam_y + lam_z,  lam(m) = (2 cos(theta_m) - 2) / h^2,
// with the zero mode of the periodic case set to zero (the mean is removed before the base solve,
// vmg.F90:213-216).  The grid sits in LDS with a row pitch of n + 1 floats so that the axis passes
// are free of bank conflicts; each thread transforms one grid line held in registers.
#pragma once
#include "amr_kernels.h"

namespace sg {
namespace amr {

constexpr int kDirectMaxN = 32;
constexpr int kDirectThreads = kDirectMaxN * kDirectMaxN;

struct BaseDirect {
    int n;              // base grid cells per side
    const float *Q;     // n x n basis, Q[m n + k]
    const float *lam;   // n eigenvalues of the 1-D second difference / h^2
};

inline size_t base_direct_lds_bytes(int n) { return sizeof(float) * ((size_t)n * n * (n + 1) + (size_t)n * n + n); }

// one axis pass: every line along `axis` replaced by Q times it
__device__ inline void direct_pass(float *g, const float *Q, int n, int axis)
{
    const int P = n + 1, t = threadIdx.x;
    if (t < n * n) {
        int base, stride;
        if (axis == 0) { base = t * P; stride = 1; }                                   // line (j, k) = t
        else if (axis == 1) { base = (t / n) * n * P + (t % n); stride = P; }          // line (i, k)
        else { base = (t / n) * P + (t % n); stride = n * P; }                         // line (i, j)
        float x[kDirectMaxN];
#pragma unroll
        for (int k = 0; k < kDirectMaxN; ++k) x[k] = k < n ? g[base + k * stride] : 0.f;
        for (int m = 0; m < n; ++m) {
            const float *q = Q + m * n;
            float acc = 0.f;
#pragma unroll
            for (int k = 0; k < kDirectMaxN; ++k)
                if (k < n) acc += q[k] * x[k];
            g[base + m * stride] = acc;
        }
    }
    __syncthreads();
}

// Au (U of FMG level 2, base blocks 0..nbase-1 first) := exact solution for the source Arhs (RHS).
__global__ void __launch_bounds__(kDirectThreads) k_base_direct(float *Au, const float *Arhs, int nbase,
                                                               const int *ijk, BaseDirect D)
{
    using B = Blk<2>;
    extern __shared__ float sm[];
    const int n = D.n, P = n + 1;
    float *g = sm, *Q = sm + n * n * P, *lam = Q + n * n;
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) Q[t] = D.Q[t];
    for (int t = threadIdx.x; t < n; t += blockDim.x) lam[t] = D.lam[t];
    for (int t = threadIdx.x; t < nbase * 8; t += blockDim.x) {
        const int b = t >> 3, c = t & 7, i = c & 1, j = (c >> 1) & 1, k = c >> 2;
        const int gi = ijk[3 * b] * 2 + i, gj = ijk[3 * b + 1] * 2 + j, gk = ijk[3 * b + 2] * 2 + k;
        g[(gk * n + gj) * P + gi] = Arhs[(long)b * B::LEN + B::gi(i, j, k)];
    }
    __syncthreads();
    for (int a = 0; a < 3; ++a) direct_pass(g, Q, n, a);
    for (int t = threadIdx.x; t < n * n * n; t += blockDim.x) {
        const int i = t % n, j = (t / n) % n, k = t / (n * n);
        const float den = lam[i] + lam[j] + lam[k];
        float &v = g[(k * n + j) * P + i];
        v = den != 0.f ? v / den : 0.f;
    }
    __syncthreads();
    for (int a = 0; a < 3; ++a) direct_pass(g, Q, n, a);
    for (int t = threadIdx.x; t < nbase * 8; t += blockDim.x) {
        const int b = t >> 3, c = t & 7, i = c & 1, j = (c >> 1) & 1, k = c >> 2;
        const int gi = ijk[3 * b] * 2 + i, gj = ijk[3 * b + 1] * 2 + j, gk = ijk[3 * b + 2] * 2 + k;
        Au[(long)b * B::LEN + B::gi(i, j, k)] = g[(gk * n + gj) * P + gi];
    }
}

// ---- the same solve spread over the chip: gather, six axis passes (one thread per output value,
// each accumulated k = 0 .. n-1 in order as above, so the values are those of k_base_direct), the
// division by the eigenvalue sum folded into the third pass and the scatter into the sixth ----
__global__ void k_bd_gather(float *g, const float *Arhs, int nbase, const int *ijk, int n)
{
    using B = Blk<2>;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nbase * 8) return;
    const int b = t >> 3, c = t & 7, i = c & 1, j = (c >> 1) & 1, k = c >> 2;
    const int gi = ijk[3 * b] * 2 + i, gj = ijk[3 * b + 1] * 2 + j, gk = ijk[3 * b + 2] * 2 + k;
    g[(gk * n + gj) * n + gi] = Arhs[(long)b * B::LEN + B::gi(i, j, k)];
}

// out[line, m] = sum_k Q[m][k] in[line, k] along axis AXIS of the dense n^3 grid (k, j, i order).
// MODE 1: divided by lam_i + lam_j + lam_k (0 where that is 0); MODE 2: written to the base blocks.
template <int AXIS, int MODE>
__global__ void k_bd_pass(float *out, const float *in, BaseDirect D, float *Au, const int *inv)
{
    using B = Blk<2>;
    const int n = D.n;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * n * n) return;
    const int i = t % n, j = (t / n) % n, k = t / (n * n);
    const int m = AXIS == 0 ? i : (AXIS == 1 ? j : k);
    const int stride = AXIS == 0 ? 1 : (AXIS == 1 ? n : n * n);
    const int base = t - m * stride;
    const float *q = D.Q + m * n;
    // the line and the basis row loaded whole first (n <= kDirectMaxN), then summed in order
    float x[kDirectMaxN], qq[kDirectMaxN];
#pragma unroll
    for (int kk = 0; kk < kDirectMaxN; ++kk)
        if (kk < n) {
            x[kk] = in[base + kk * stride];
            qq[kk] = q[kk];
        }
    float acc = 0.f;
#pragma unroll
    for (int kk = 0; kk < kDirectMaxN; ++kk)
        if (kk < n) acc += qq[kk] * x[kk];
    if (MODE == 1) {
        const float den = D.lam[i] + D.lam[j] + D.lam[k];
        acc = den != 0.f ? acc / den : 0.f;
    }
    if (MODE == 2) {
        // base block of cell (i, j, k): blocks are 2^3 cells; inv maps block position
        // (i/2, j/2, k/2) to its block number
        const int nb = n / 2;
        const int b = inv[((k >> 1) * nb + (j >> 1)) * nb + (i >> 1)];
        Au[(long)b * B::LEN + B::gi(i & 1, j & 1, k & 1)] = acc;
    } else {
        out[t] = acc;
    }
}

} // namespace amr
} // namespace sg
