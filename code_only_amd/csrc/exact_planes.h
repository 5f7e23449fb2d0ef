// Exact solve of a periodic level of n^3 cells (n <= 64) spread over the chip, for the FP32 inner
// solve of the mixed-precision path: the coarse-grid problem L u = f that the V-cycles below this
// level approximate (mg.F90:142-210) is solved exactly instead.
//
// The 7-point operator separates; the symmetric orthonormal Hartley basis Q of the periodic second
// difference, Q[m n + k] = (cos a + sin a) / sqrt(n), a = 2 pi ((m k) mod n) / n, diagonalises it
// along each axis with eigenvalues lam[m] = (2 cos(2 pi m / n) - 2) / h^2, so
//     u = Q_xyz (Lambda^-1 (Q_xyz f)),  Lambda = (lam_i + lam_j) + lam_k,
// the zero mode set to zero (the zero-mean solution).  Three launches of n workgroups each:
//   k_exact_xy  (plane k):  x pass, then y pass             f -> w
//   k_exact_z   (row j):    z pass, divide by Lambda, z pass w -> w
//   k_exact_xy  (plane k):  x pass, then y pass             w -> u
// Each workgroup stages its n x n plane in LDS with the transformed index k leading ([k][l], rows
// of 68 floats: 16-byte aligned) and Q beside it (Q is symmetric, so its row k is column k).
// Threads t and t + 256 compute the same 4 x 4 block of outputs -- m in [4 mb, 4 mb + 4) of the
// lines l in [4 lb, 4 lb + 4) -- over the first and the second half of k: per k one float4 of Q
// and one float4 of line values feed 16 fused multiply-adds in k order, and the block is
// (first half) + (second half).  It is written back either as [m][l] or transposed as [l][m],
// whichever layout the next pass reads.  The division by Lambda is a product with 1 / Lambda,
// tabulated once per level (formed in double precision, zero for the zero mode).
#pragma once
#include <hip/hip_runtime.h>

namespace sg {

constexpr int kPlaneMaxN = 64;
constexpr int kPlaneThreads = 512;
constexpr int kPlaneBlocks = kPlaneThreads / 2;   // 4 x 4 output blocks per plane, at most
constexpr int kPlanePitch = kPlaneMaxN + 4;
constexpr int kPlaneLoads = kPlaneMaxN * kPlaneMaxN / kPlaneThreads;  // plane values per thread
static_assert((kPlaneMaxN / 4) * (kPlaneMaxN / 4) <= kPlaneBlocks, "one thread pair per 4 x 4 block");

// Q (n x n, row-major, global) into LDS rows of pitch kPlanePitch; every load issued first.
__device__ __forceinline__ void ep_stage_q(float *Qs, const float *Q, int n)
{
    float v[kPlaneLoads];
#pragma unroll
    for (int it = 0; it < kPlaneLoads; ++it) {
        const int c = (int)threadIdx.x + it * kPlaneThreads;
        v[it] = 0.f;
        if (c < n * n) v[it] = Q[c];
    }
#pragma unroll
    for (int it = 0; it < kPlaneLoads; ++it) {
        const int c = (int)threadIdx.x + it * kPlaneThreads;
        if (c < n * n) Qs[(c / n) * kPlanePitch + c % n] = v[it];
    }
}

// S ([k][l], n x n) := Q S, written as [m][l] (transpose == false) or [l][m].  n % 8 == 0 or
// n % 4 == 0 with n / 2 even; part: scratch of kPlaneBlocks * 16 floats for the second halves.
__device__ __forceinline__ void ep_pass(float *S, const float *Qs, float *part, int n, bool transpose)
{
    const int nb = n / 4, t = (int)threadIdx.x % kPlaneBlocks, half = (int)threadIdx.x / kPlaneBlocks;
    const bool act = t < nb * nb;
    const int mb = act ? t / nb : 0, lb = act ? t % nb : 0;
    const int kh = n / 2, kbeg = half * kh;
    float acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.f;
    if (act) {
        // two k at a time: the four loads issued before the 32 multiply-adds (k in order)
        for (int k0 = kbeg; k0 < kbeg + kh; k0 += 2) {
            float4 qv[2], xv[2];
#pragma unroll
            for (int d = 0; d < 2; ++d) {
                qv[d] = *reinterpret_cast<const float4 *>(Qs + (k0 + d) * kPlanePitch + 4 * mb);
                xv[d] = *reinterpret_cast<const float4 *>(S + (k0 + d) * kPlanePitch + 4 * lb);
            }
#pragma unroll
            for (int d = 0; d < 2; ++d) {
                const float q[4] = {qv[d].x, qv[d].y, qv[d].z, qv[d].w}, x[4] = {xv[d].x, xv[d].y, xv[d].z, xv[d].w};
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) acc[a][b] = __builtin_fmaf(q[a], x[b], acc[a][b]);
            }
        }
    }
    if (act && half == 1) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
            *reinterpret_cast<float4 *>(part + (a * kPlaneBlocks + t) * 4) =
                make_float4(acc[a][0], acc[a][1], acc[a][2], acc[a][3]);
    }
    __syncthreads();
    if (act && half == 0) {
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const float4 p = *reinterpret_cast<const float4 *>(part + (a * kPlaneBlocks + t) * 4);
            acc[a][0] += p.x; acc[a][1] += p.y; acc[a][2] += p.z; acc[a][3] += p.w;
        }
        if (!transpose) {
#pragma unroll
            for (int a = 0; a < 4; ++a)
                *reinterpret_cast<float4 *>(S + (4 * mb + a) * kPlanePitch + 4 * lb) =
                    make_float4(acc[a][0], acc[a][1], acc[a][2], acc[a][3]);
        } else {
#pragma unroll
            for (int b = 0; b < 4; ++b)
                *reinterpret_cast<float4 *>(S + (4 * lb + b) * kPlanePitch + 4 * mb) =
                    make_float4(acc[0][b], acc[1][b], acc[2][b], acc[3][b]);
        }
    }
    __syncthreads();
}

// Plane k = `plane` of the dense n^3 array in (i fastest): x pass then y pass, into out.
__device__ __forceinline__ void exact_xy(const float *__restrict__ in, float *__restrict__ out, const float *Q, int n,
                                         int plane)
{
    __shared__ __align__(16) float S[kPlaneMaxN * kPlanePitch];
    __shared__ __align__(16) float Qs[kPlaneMaxN * kPlanePitch];
    __shared__ __align__(16) float part[kPlaneBlocks * 16];
    const long base = (long)plane * n * n;
    float v[kPlaneLoads];
#pragma unroll
    for (int it = 0; it < kPlaneLoads; ++it) {
        const int c = (int)threadIdx.x + it * kPlaneThreads;
        v[it] = 0.f;
        if (c < n * n) v[it] = in[base + c];
    }
    ep_stage_q(Qs, Q, n);
    // x lines (rows j) with the x index i leading: S[i][j]
#pragma unroll
    for (int it = 0; it < kPlaneLoads; ++it) {
        const int c = (int)threadIdx.x + it * kPlaneThreads;
        if (c < n * n) S[(c % n) * kPlanePitch + c / n] = v[it];
    }
    __syncthreads();
    ep_pass(S, Qs, part, n, true);   // S[j][mi]: y index leading, lines mi
    ep_pass(S, Qs, part, n, true);   // S[mi][mj]
    for (int c = (int)threadIdx.x; c < n * n; c += kPlaneThreads) out[base + c] = S[(c % n) * kPlanePitch + c / n];
}

// Row j = `row` of the dense n^3 array w: the (k, i) plane, z pass, product with 1 / Lambda
// (inv, dense n^3 like w), z pass, back into w.
__device__ __forceinline__ void exact_z(float *__restrict__ w, const float *Q, const float *inv, int n, int row)
{
    __shared__ __align__(16) float S[kPlaneMaxN * kPlanePitch];
    __shared__ __align__(16) float Qs[kPlaneMaxN * kPlanePitch];
    __shared__ __align__(16) float part[kPlaneBlocks * 16];
    const int j = row;
    float v[kPlaneLoads], d[kPlaneLoads];
#pragma unroll
    for (int it = 0; it < kPlaneLoads; ++it) {
        const int c = (int)threadIdx.x + it * kPlaneThreads;
        v[it] = d[it] = 0.f;
        if (c < n * n) {
            const long g = ((long)(c / n) * n + j) * n + c % n;
            v[it] = w[g];
            d[it] = inv[g];
        }
    }
    ep_stage_q(Qs, Q, n);
#pragma unroll
    for (int it = 0; it < kPlaneLoads; ++it) {
        const int c = (int)threadIdx.x + it * kPlaneThreads;
        if (c < n * n) S[(c / n) * kPlanePitch + c % n] = v[it];
    }
    __syncthreads();
    ep_pass(S, Qs, part, n, false);  // S[mk][i]
#pragma unroll
    for (int it = 0; it < kPlaneLoads; ++it) {
        const int c = (int)threadIdx.x + it * kPlaneThreads;
        if (c < n * n) S[(c / n) * kPlanePitch + c % n] *= d[it];
    }
    __syncthreads();
    ep_pass(S, Qs, part, n, false);
    for (int c = (int)threadIdx.x; c < n * n; c += kPlaneThreads) {
        const int k = c / n, i = c % n;
        w[((long)k * n + j) * n + i] = S[k * kPlanePitch + i];
    }
}

static __global__ void __launch_bounds__(kPlaneThreads) k_exact_xy(const float *in, float *out, const float *Q, int n)
{
    exact_xy(in, out, Q, n, (int)blockIdx.x);
}

static __global__ void __launch_bounds__(kPlaneThreads) k_exact_z(float *w, const float *Q, const float *inv, int n)
{
    exact_z(w, Q, inv, n, (int)blockIdx.x);
}

// 1 / ((lam_i + lam_j) + lam_k) for the dense n^3 grid (0 where the sum is 0), host, double.
inline void exact_inverse_table(int n, const double *lam, float *inv)
{
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const double den = (lam[i] + lam[j]) + lam[k];
                inv[((size_t)k * n + j) * n + i] = den != 0.0 ? (float)(1.0 / den) : 0.f;
            }
}

} // namespace sg
