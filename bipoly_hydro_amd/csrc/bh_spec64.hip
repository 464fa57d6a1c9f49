// bh_spec64.hip -- the FP64 spectral solve.
//
// The reference's N Peaceman-Rachford steps (helmadi.F90:195-355) give
// x_N = X0 + delta with A delta = (I - M) r0, r0 = g - A X0 and M = mu_z x
// M_r (the mixed solve's form, bh_mixed.hip).  Along z, A = B_r + B_z with
// B_z = gamma tridiag(-1, 2, -1), diagonalised by the orthonormal sine basis
// S (S = S^T = S^-1); along J, M_r(L) is a dense nri x nri matrix per mode.
// So, per mode plane L (compact interior [k or K][J]):
//   r0      residual in FP64 (res64)
//   T = S r0            FP64 MFMA product (bh_gemm.hip)
//   U = T M_r(L)^T      FP64 MFMA product
//   C = T - mu_k U, (B_r(L) + lambda_k) D = C along J   (thomas64)
//   delta = S D         FP64 MFMA product
//   X = X0 + delta      (update64), then the inverse DFT.
// One pass in FP64 holds the truncated iteration to rounding.
#include "bh_spec64.h"

#include <hip/hip_runtime.h>

namespace bh {

// ---------------------------------------------------------------------------
// r0 on the interior: one thread per point, J fastest.  A x is the 5-point
// form of the ADI operators (potsetup.F90:175-228): brb x + alphar x(J+1) +
// betar x(J-1) (no J-1 term at J = 2) + gamma (2 x - x(K-1) - x(K+1)).  A cold
// X0 is zero on the interior: only its face values enter.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void res64_tile(const Res64Args &a, int bx, int by, int)
{
    const int numr = a.numr, numz = a.numz, nri = numr - 2, nzi = numz - 2;
    const int t = bx * 256 + threadIdx.x, L = by;
    if (t >= nri * nzi) return;
    const int jr = t % nri, kz = t / nri, J = jr + 2, K = kz + 2;
    const long plane = (long)numr * numz;
    const double *X = a.X + L * plane;
    const long c = (long)(K - 1) * numr + (J - 1);
    auto x = [&](long i, bool face) { return (a.cold && !face) ? 0.0 : X[i]; };
    const double xc = x(c, false);
    const double xe = x(c + 1, J + 1 == numr), xw = J > 2 ? x(c - 1, false) : 0.0;
    const double xs = x(c - numr, K - 1 == 1), xn = x(c + numr, K + 1 == numz);
    const double brbJ = a.brb[(long)L * (numr + 1) + J];
    double Ax = brbJ * xc + a.alphar[J] * xe;
    Ax += J > 2 ? a.betar[J] * xw : 0.0;
    Ax += 2.0 * a.gamma * xc - a.gamma * (xs + xn);
    a.r[(long)L * nri * nzi + t] = -a.four_pi * a.rh[L * plane + c] - Ax;
}

Op res64_op(const Res64Args &a, int nplanes)
{
    const int n = (a.numr - 2) * (a.numz - 2);
    return Op::make(OP_RES64, 0, a, (n + 255) / 256, nplanes, 1);
}

// ---------------------------------------------------------------------------
// One thread per line (L, k), the 64 lanes of a wave on 64 consecutive k of
// one plane.  Forward: c_j = T_j - mu_k U_j, and the Thomas elimination of
// tridagr.F90:41-60 with the diagonal shifted by lambda_k (bet_0 = b_0,
// gam_j = c_{j-1} / bet_{j-1}, bet_j = b_j - a_j gam_j, y_j = (c_j - a_j
// y_{j-1}) / bet_j; a_j = betar, c_j = alphar, b_j = brb + lambda_k); y goes
// to U, gam to T.  Back: x_j = y_j - gam_{j+1} x_{j+1}.  Loads come in
// chunks of TS64 elements issued before their use.
// ---------------------------------------------------------------------------
constexpr int TS64 = 16;

__device__ __forceinline__ void thomas64_tile(const Thomas64Args &a, int bx, int by, int)
{
    const int nri = a.nri, nzi = a.nzi, numr = a.numr;
    const int k = bx * 256 + threadIdx.x, L = by;
    if (k >= nzi) return;
    const double *brb = a.brb + (long)L * (numr + 1);
    const double mu = a.mu[k], lam = a.lam[k];
    double *T = a.T + ((long)L * nzi + k) * nri, *U = a.U + ((long)L * nzi + k) * nri;
    double y = 0.0, bet = 1.0;
    for (int j0 = 0; j0 < nri; j0 += TS64) {
        double tv[TS64], uv[TS64];
#pragma unroll
        for (int u = 0; u < TS64; ++u) {
            const int j = j0 + u;
            tv[u] = j < nri ? T[j] : 0.0;
            uv[u] = j < nri ? U[j] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < TS64; ++u) {
            const int j = j0 + u, J = j + 2;
            if (j < nri) {
                const double c = tv[u] - mu * uv[u];
                double g = 0.0;
                if (j == 0) {
                    bet = brb[J] + lam;
                    y = c / bet;
                } else {
                    g = a.alphar[J - 1] / bet;
                    bet = (brb[J] + lam) - a.betar[J] * g;
                    y = (c - a.betar[J] * y) / bet;
                }
                U[j] = y;
                T[j] = g;
            }
        }
    }
    double x = y;
    for (int j1 = nri - 2; j1 >= 0; j1 -= TS64) {
        double yv[TS64], gv[TS64];
#pragma unroll
        for (int u = 0; u < TS64; ++u) {
            const int j = j1 - u;
            yv[u] = j >= 0 ? U[j] : 0.0;
            gv[u] = j >= 0 ? T[j + 1] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < TS64; ++u) {
            const int j = j1 - u;
            if (j >= 0) {
                x = yv[u] - gv[u] * x;
                U[j] = x;
            }
        }
    }
}

Op thomas64_op(const Thomas64Args &a, int nplanes)
{
    return Op::make(OP_THOMAS64, 0, a, (a.nzi + 255) / 256, nplanes, 1);
}

__device__ __forceinline__ void update64_tile(const Update64Args &a, int bx, int by, int)
{
    const int numr = a.numr, numz = a.numz, nri = numr - 2, nzi = numz - 2;
    const int t = bx * 256 + threadIdx.x, L = by;
    if (t >= nri * nzi) return;
    const int jr = t % nri, kz = t / nri;
    const long c = (long)L * numr * numz + (long)(kz + 1) * numr + jr + 1;
    const double d = a.dlt[(long)L * nri * nzi + t];
    a.X[c] = a.cold ? d : a.X[c] + d;
}

Op update64_op(const Update64Args &a, int nplanes)
{
    const int n = (a.numr - 2) * (a.numz - 2);
    return Op::make(OP_UPDATE64, 0, a, (n + 255) / 256, nplanes, 1);
}

#define BH_KERNEL_PAIR(NAME, ARGS_T, NT)                                                          \
    __global__ __launch_bounds__(NT) void NAME##_kernel(ARGS_T a, int nx, int ny, int total)      \
    {                                                                                             \
        BH_SINGLE_LOOP(a, nx, ny, total, (NAME##_tile(a, bx, by, bz)));                           \
    }                                                                                             \
    __global__ __launch_bounds__(NT) void NAME##_grouped(const ARGS_T *__restrict__ args, GroupSched gs) \
    {                                                                                             \
        BH_GROUP_LOOP(ARGS_T, args, gs, (NAME##_tile(ga, bx, by, bz)));                           \
    }
BH_KERNEL_PAIR(res64, Res64Args, 256)
BH_KERNEL_PAIR(thomas64, Thomas64Args, 256)
BH_KERNEL_PAIR(update64, Update64Args, 256)
#undef BH_KERNEL_PAIR

namespace {
bool spec64_ref(int kind, const void **single, const void **grouped)
{
    switch (kind) {
    case OP_RES64: *single = (const void *)&res64_kernel; *grouped = (const void *)&res64_grouped; return true;
    case OP_THOMAS64: *single = (const void *)&thomas64_kernel; *grouped = (const void *)&thomas64_grouped; return true;
    case OP_UPDATE64: *single = (const void *)&update64_kernel; *grouped = (const void *)&update64_grouped; return true;
    default: return false;
    }
}
}  // namespace

bool spec64_launch(const Op &op, hipStream_t s)
{
    const void *k1, *k2;
    if (!spec64_ref(op.kind, &k1, &k2)) return false;
    int nx = op.nx, ny = op.ny, total = (int)op.tiles();
    void *argv[] = {(void *)op.args, (void *)&nx, (void *)&ny, (void *)&total};
    const int nb = persistent_blocks(total, kernel_occupancy(k1, 256, 0));
    (void)hipLaunchKernel(k1, dim3(nb), dim3(256), argv, 0, s);
    return true;
}

bool spec64_launch_grouped(int kind, int variant, const void *d_args, const GroupSched &gs, int nblocks,
                           hipStream_t s)
{
    (void)variant;
    const void *k1, *k2;
    if (!spec64_ref(kind, &k1, &k2)) return false;
    GroupSched g = gs;
    const void *a = d_args;
    void *argv[] = {(void *)&a, (void *)&g};
    (void)hipLaunchKernel(k2, dim3(nblocks), dim3(256), argv, 0, s);
    return true;
}

bool spec64_occupancy(int kind, int variant, int *blocks)
{
    (void)variant;
    const void *k1, *k2;
    if (!spec64_ref(kind, &k1, &k2)) return false;
    *blocks = kernel_occupancy(k2, 256, 0);
    return true;
}

}  // namespace bh
