// Exact solve of a periodic level of n^3 cells (n <= 64) spread over the chip, for the FP32 inner
// solve of the mixed-precision path: the coarse-grid problem L u = f that the V-cycles below this
// level approximate (mg.F90:142-210) is solved exactly instead.
//
// The 7-point operator separates; the symmetric orthonormal Hartley basis Q of the periodic second
// difference, Q[m n + k] = (cos a + sin a) / sqrt(n), a = 2 pi ((m k) mod n) / n, diagonalises it
// along each axis with eigenvalues lam[m] = (2 cos(2 pi m / n) - 2) / h^2, so
//     u = Q_xyz (Lambda^-1 (Q_xyz f)),  Lambda = (lam_i + lam_j) + lam_k,
// the zero mode set to zero (the zero-mean solution).  Three launches:
//   xy  (plane k):  x pass, then y pass             f -> w
//   z   (row j):    z pass, times 1 / Lambda, z pass w -> w
//   xy  (plane k):  x pass, then y pass             w -> u
// Every transformed value is sum_k Q[m][k] v[k] formed as two chains of fused multiply-adds in k
// order, one over the first and one over the second half of k, added at the end:
// (first half) + (second half).  The division by Lambda is a product with 1 / Lambda, tabulated
// once per level (formed in double precision, zero for the zero mode).
//
// Two launch forms give the same values bit for bit:
// * whole (k_exact_xy, k_exact_z; the batched solver): one 512-thread workgroup per plane (row).
//   The plane sits in LDS with the transformed index leading ([k][l], rows of 68 floats) and Q
//   beside it (Q is symmetric, so its row k is column k).  Threads t and t + 256 form the same
//   4 x 4 block of outputs -- m in [4 mb, 4 mb + 4) of the lines l in [4 lb, 4 lb + 4) -- over the
//   two halves of k: per k one float4 of Q and one float4 of line values feed 16 multiply-adds.
//   Few workgroups of much work each: the form for many planes at once.
// * split (k_exact_xy_split, k_exact_z_split; the single-grid solver): G = plane_groups(n)
//   workgroups of 512 threads per plane (row), split by the range of the first transformed index.
//   Workgroup g of plane k forms the x transform X[mi][j] only for its n / G frequencies mi, then
//   the y transform of those columns; workgroup g of row j transforms its n / G columns i along
//   z, in place.  A task is four outputs (four consecutive frequencies), formed from one float4 of
//   Q per k by two threads, one per half of k.  G times as many workgroups of 1 / G of the work,
//   and two waves per SIMD on the serial chains: the form for one level, whose n planes alone
//   would leave most of the chip idle.
#pragma once
#include <hip/hip_runtime.h>

namespace sg {

constexpr int kPlaneMaxN = 64;
constexpr int kPlaneThreads = 512;
constexpr int kPlaneBlocks = kPlaneThreads / 2;   // 4 x 4 output blocks per plane, at most
constexpr int kPlanePitch = kPlaneMaxN + 4;
constexpr int kPlaneLoads = kPlaneMaxN * kPlaneMaxN / kPlaneThreads;  // plane values per thread
static_assert((kPlaneMaxN / 4) * (kPlaneMaxN / 4) <= kPlaneBlocks, "one thread pair per 4 x 4 block");

// Workgroups per plane (row) of the split form, each owning n / 4 frequencies (a multiple of 4):
// the split form takes n % 16 == 0 (the single-grid levels are powers of two >= 32).
__host__ __device__ constexpr int plane_groups(int) { return 4; }

// Q (n x n, row-major, global) into LDS rows of pitch kPlanePitch; every load issued first.
template <int NT> __device__ __forceinline__ void ep_stage_q(float *Qs, const float *Q, int n)
{
    constexpr int LD = kPlaneMaxN * kPlaneMaxN / NT;
    float v[LD];
#pragma unroll
    for (int it = 0; it < LD; ++it) {
        const int c = (int)threadIdx.x + it * NT;
        v[it] = 0.f;
        if (c < n * n) v[it] = Q[c];
    }
#pragma unroll
    for (int it = 0; it < LD; ++it) {
        const int c = (int)threadIdx.x + it * NT;
        if (c < n * n) Qs[(c / n) * kPlanePitch + c % n] = v[it];
    }
}

// ---- whole form ----------------------------------------------------------------------------------

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
    ep_stage_q<kPlaneThreads>(Qs, Q, n);
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
    ep_stage_q<kPlaneThreads>(Qs, Q, n);
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

// ---- split form ----------------------------------------------------------------------------------

// Chain h (0: k < n/2, 1: k >= n/2) of acc[r] = sum_k Q[m0 + r][k] * V[k * sk + col * sc], r < 4.
__device__ __forceinline__ void ep_dot4_half(const float *Qs, const float *V, int sk, int sc, int n, int m0, int col,
                                             int h, float (&acc)[4])
{
    acc[0] = acc[1] = acc[2] = acc[3] = 0.f;
    const float *v = V + col * sc;
    const int kh = n / 2, k0 = h * kh;
    for (int k = k0; k < k0 + kh; k += 2) {
        const float4 q0 = *reinterpret_cast<const float4 *>(Qs + k * kPlanePitch + m0);
        const float4 q1 = *reinterpret_cast<const float4 *>(Qs + (k + 1) * kPlanePitch + m0);
        const float x0 = v[k * sk], x1 = v[(k + 1) * sk];
        acc[0] = __builtin_fmaf(q0.x, x0, acc[0]);
        acc[1] = __builtin_fmaf(q0.y, x0, acc[1]);
        acc[2] = __builtin_fmaf(q0.z, x0, acc[2]);
        acc[3] = __builtin_fmaf(q0.w, x0, acc[3]);
        acc[0] = __builtin_fmaf(q1.x, x1, acc[0]);
        acc[1] = __builtin_fmaf(q1.y, x1, acc[1]);
        acc[2] = __builtin_fmaf(q1.z, x1, acc[2]);
        acc[3] = __builtin_fmaf(q1.w, x1, acc[3]);
    }
}

// One pass of the split form over `tasks` tasks of four outputs each: thread t < kSplitTasks forms
// the first-half chain of task t (+ trips of kSplitTasks), thread t + kSplitTasks the second-half
// chain, which it hands over through `part`; fin(task, acc) receives (first half) + (second half).
// Every thread runs the same trips (each holds a barrier).
constexpr int kSplitTasks = 256;
constexpr int kSplitThreads = 2 * kSplitTasks;
constexpr int kSplitLoads = kPlaneMaxN * kPlaneMaxN / kSplitThreads;

template <typename Task, typename Fin>
__device__ __forceinline__ void ep_split_pass(const float *Qs, float *part, int n, int tasks, Task task, Fin fin)
{
    const int h = (int)threadIdx.x / kSplitTasks, tl = (int)threadIdx.x % kSplitTasks;
    for (int base = 0; base < tasks; base += kSplitTasks) {
        const int t = base + tl;
        float acc[4] = {0.f, 0.f, 0.f, 0.f};
        if (t < tasks) {
            const float *V;
            int sk, sc, m0, col;
            task(t, V, sk, sc, m0, col);
            ep_dot4_half(Qs, V, sk, sc, n, m0, col, h, acc);
            if (h == 1) *reinterpret_cast<float4 *>(part + 4 * tl) = make_float4(acc[0], acc[1], acc[2], acc[3]);
        }
        __syncthreads();
        if (t < tasks && h == 0) {
            const float4 q = *reinterpret_cast<const float4 *>(part + 4 * tl);
            acc[0] += q.x; acc[1] += q.y; acc[2] += q.z; acc[3] += q.w;
            fin(t, acc);
        }
        __syncthreads();
    }
}

// Workgroup wg = plane * G + g of the dense n^3 array in (i fastest), G = plane_groups(n): for the
// x frequencies mi in [g n/G, (g+1) n/G), X[mi][j] = sum_i Q[mi][i] in[plane][j][i], then
// out[plane][mj][mi] = sum_j Q[mj][j] X[mi][j].
__device__ __forceinline__ void exact_xy_split(const float *__restrict__ in, float *__restrict__ out, const float *Q,
                                               int n, int wg)
{
    constexpr int PS = kPlaneMaxN + 1;                                      // plane rows [j][i], odd pitch
    __shared__ __align__(16) float Qs[kPlaneMaxN * kPlanePitch];
    __shared__ float S[kPlaneMaxN * PS];
    __shared__ __align__(16) float Xs[kPlaneMaxN * kPlanePitch];
    __shared__ __align__(16) float part[4 * kSplitTasks];
    const int G = plane_groups(n), mq = n / G, plane = wg / G, g = wg % G;
    const int px = mq + 4;                                                  // X rows [j][mi local]
    const long base = (long)plane * n * n;
    float v[kSplitLoads];
#pragma unroll
    for (int it = 0; it < kSplitLoads; ++it) {
        const int c = (int)threadIdx.x + it * kSplitThreads;
        v[it] = 0.f;
        if (c < n * n) v[it] = in[base + c];
    }
    ep_stage_q<kSplitThreads>(Qs, Q, n);
#pragma unroll
    for (int it = 0; it < kSplitLoads; ++it) {
        const int c = (int)threadIdx.x + it * kSplitThreads;
        if (c < n * n) S[(c / n) * PS + c % n] = v[it];
    }
    __syncthreads();
    // x pass: task (line j, block b of four local frequencies)
    ep_split_pass(
        Qs, part, n, n * (mq / 4),
        [&](int t, const float *&V, int &sk, int &sc, int &m0, int &col) {
            V = S; sk = 1; sc = PS; m0 = g * mq + 4 * (t / n); col = t % n;
        },
        [&](int t, const float (&acc)[4]) {
            *reinterpret_cast<float4 *>(Xs + (t % n) * px + 4 * (t / n)) = make_float4(acc[0], acc[1], acc[2], acc[3]);
        });
    // y pass: task (local frequency mi, block of four mj)
    ep_split_pass(
        Qs, part, n, mq * (n / 4),
        [&](int t, const float *&V, int &sk, int &sc, int &m0, int &col) {
            V = Xs; sk = px; sc = 1; m0 = 4 * (t / mq); col = t % mq;
        },
        [&](int t, const float (&acc)[4]) {
            const int mi = t % mq, mj0 = 4 * (t / mq);
#pragma unroll
            for (int r = 0; r < 4; ++r) out[base + (long)(mj0 + r) * n + g * mq + mi] = acc[r];
        });
}

// Workgroup wg = row * G + g of the dense n^3 array w: for the columns i in [g n/G, (g+1) n/G)
// of row j, Z[mk][i] = sum_k Q[mk][k] w[k][j][i], times inv[mk][j][i] (1 / Lambda, dense n^3 like
// w), then w[k][j][i] = sum_mk Q[k][mk] Z[mk][i].
__device__ __forceinline__ void exact_z_split(float *__restrict__ w, const float *Q, const float *inv, int n, int wg)
{
    constexpr int PC = kPlaneMaxN;                                          // column block rows [k][i local]
    __shared__ __align__(16) float Qs[kPlaneMaxN * kPlanePitch];
    __shared__ float S[kPlaneMaxN * PC];
    __shared__ float Zs[kPlaneMaxN * PC];
    __shared__ __align__(16) float part[4 * kSplitTasks];
    const int G = plane_groups(n), mq = n / G, j = wg / G, g = wg % G;
    const long nn = (long)n * n, col0 = (long)j * n + g * mq;
    const int tasks = mq * (n / 4);
    float v[kSplitLoads];
#pragma unroll
    for (int it = 0; it < kSplitLoads; ++it) {
        const int c = (int)threadIdx.x + it * kSplitThreads;
        v[it] = 0.f;
        if (c < n * mq) v[it] = w[(c / mq) * nn + col0 + c % mq];
    }
    // 1 / Lambda of the first trip's task of this thread (used by the first-half threads), ahead
    const int t0 = (int)threadIdx.x % kSplitTasks;
    const bool lo = (int)threadIdx.x < kSplitTasks;
    float d0[4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
        d0[r] = lo && t0 < tasks ? inv[(4 * (t0 / mq) + r) * nn + col0 + t0 % mq] : 0.f;
    ep_stage_q<kSplitThreads>(Qs, Q, n);
#pragma unroll
    for (int it = 0; it < kSplitLoads; ++it) {
        const int c = (int)threadIdx.x + it * kSplitThreads;
        if (c < n * mq) S[(c / mq) * PC + c % mq] = v[it];
    }
    __syncthreads();
    auto task_of = [&](const float *V) {
        return [=](int t, const float *&Vo, int &sk, int &sc, int &m0, int &col) {
            Vo = V; sk = PC; sc = 1; m0 = 4 * (t / mq); col = t % mq;
        };
    };
    ep_split_pass(Qs, part, n, tasks, task_of(S), [&](int t, const float (&acc)[4]) {
        const int ii = t % mq, m0 = 4 * (t / mq);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float d = t == t0 ? d0[r] : inv[(m0 + r) * nn + col0 + ii];
            Zs[(m0 + r) * PC + ii] = acc[r] * d;
        }
    });
    ep_split_pass(Qs, part, n, tasks, task_of(Zs), [&](int t, const float (&acc)[4]) {
        const int ii = t % mq, m0 = 4 * (t / mq);
#pragma unroll
        for (int r = 0; r < 4; ++r) w[(m0 + r) * nn + col0 + ii] = acc[r];
    });
}

// XCD-aware numbering: the hardware deals consecutive workgroups round-robin to the 8 XCDs; the
// G workgroups of one plane (which all stage that plane and Q) get consecutive numbers on one XCD.
__device__ __forceinline__ int ep_xcd_block()
{
    const int nb = (int)gridDim.x, b = (int)blockIdx.x;
    return nb % 8 == 0 ? (b % 8) * (nb / 8) + b / 8 : b;
}

static __global__ void __launch_bounds__(kSplitThreads) k_exact_xy_split(const float *in, float *out, const float *Q,
                                                                         int n)
{
    exact_xy_split(in, out, Q, n, ep_xcd_block());
}

static __global__ void __launch_bounds__(kSplitThreads) k_exact_z_split(float *w, const float *Q, const float *inv,
                                                                        int n)
{
    exact_z_split(w, Q, inv, n, ep_xcd_block());
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
