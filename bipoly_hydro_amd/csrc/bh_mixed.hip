// bh_mixed.hip -- device code of the mixed-precision solve.
//
// The reference's Peaceman-Rachford iteration (helmadi.F90:195-355) uses a
// radial operator B_r that acts along J only and a vertical operator B_z that
// acts along K only, so the two commute and N steps from x0 give exactly
//     x_N = x* + M (x0 - x*),   A x* = g,   A = B_r + B_z,
//     M = prod_s (dt_s - B_r)(dt_s + B_r)^-1 (dt_s - B_z)(dt_s + B_z)^-1,
// i.e. x_N solves A x_N = g - M r0 with r0 = g - A x0.  B_z = gamma
// tridiag(-1, 2, -1) is diagonalised by the sine transform S, so M is
// mu_z(k) x M_r in (sine mode k, J) coordinates, with M_r a dense nri x nri
// matrix per azimuthal mode built once per plan.  The mixed solve evaluates
// this with FP32 MFMA products (sine transforms, the M_r product) and FP64
// residuals of A x = g - w in physical (J, K) coordinates, then refines once.
//
//   residual64      r = g - A x (- w), optionally after x += delta; FP64
//   tridiag_shift   (B_r + lambda_z(k)) y = c for every (mode, k), FP64
//   build_mr        M_r for one pseudo-timestep schedule (plan time), FP64
#include "bh_internal.h"
#include "bh_mixed.h"
#include "bh_residual.h"
#include "bh_ops.h"

#include <hip/hip_runtime.h>

namespace bh {

// the refinement form (delta and w given, Xo written or not) or the first
// residual, cold or warm; the first and last tiles of a plane (face rows
// K = 1, numz) apart
template <int ROWS, bool EDGE>
__device__ __forceinline__ void residual_tile_e(const ResidualArgs &a, int bx, int by, int bz)
{
    if (a.Xo && a.dlt2) {  // a warm start from a cold solve's two corrections (no w), if kept so
        if (!a.lazy_flag || *a.lazy_flag != 0.f) residual_tile_v<ROWS, true, true, EDGE, true, true, false, true>(a, bx, by, bz);
        else residual_tile_v<ROWS, false, false, EDGE, false>(a, bx, by, bz);
        return;
    }
    switch ((a.dlt ? (a.Xo ? 2 : a.dlt2 ? 6 : 4) : 0) + (a.cold ? 1 : 0)) {
    case 0: residual_tile_v<ROWS, false, false, EDGE, false>(a, bx, by, bz); break;
    case 1: residual_tile_v<ROWS, false, true, EDGE, false>(a, bx, by, bz); break;
    case 2: residual_tile_v<ROWS, true, false, EDGE, true>(a, bx, by, bz); break;
    case 3: residual_tile_v<ROWS, true, true, EDGE, true>(a, bx, by, bz); break;
    case 4: residual_tile_v<ROWS, true, false, EDGE, false>(a, bx, by, bz); break;
    case 5: residual_tile_v<ROWS, true, true, EDGE, false>(a, bx, by, bz); break;
    case 6: residual_tile_v<ROWS, true, false, EDGE, false, true>(a, bx, by, bz); break;
    default: residual_tile_v<ROWS, true, true, EDGE, false, true>(a, bx, by, bz); break;
    }
}
template <int ROWS>
__device__ __forceinline__ void residual_tile_r(const ResidualArgs &a, int bx, int by, int bz)
{
    const int nzi = a.numz - 2, K0 = by * ROWS + 2;
    if (by == 0 || K0 + ROWS >= nzi + 2) residual_tile_e<ROWS, true>(a, bx, by, bz);
    else residual_tile_e<ROWS, false>(a, bx, by, bz);
}

int residual_rows(int variant) { return variant == 1 ? 2 : variant == 2 ? 8 : RES_ROWS; }

Op residual_op(const ResidualArgs &a, int nplanes, int rows)
{
    const int v = rows == 2 ? 1 : rows == 8 ? 2 : 0, r = residual_rows(v);
    return Op::make(OP_RESIDUAL, v, a, (a.numr - 2 + 255) / 256, (a.numz - 2 + r - 1) / r, nplanes);
}

// Xo = X + (dlt + dlt2) on the interior (faces of Xo are kept); a cold X is
// zero on the interior and is not read
__device__ __forceinline__ void add_update_tile(const UpdateArgs &a, int bx, int by, int)
{
    const int numr = a.numr, numz = a.numz;
    const long plane = (long)numr * numz;
    const int L = by;
    int t = bx * 256 + threadIdx.x;
    int nri = numr - 2, nzi = numz - 2;
    if (t >= nri * nzi) return;
    int jr = t % nri, kz = t / nri;
    long c = L * plane + (long)(kz + 1) * numr + (jr + 1);
    const double d = a.dlt2 ? (double)a.dlt[c] + (double)a.dlt2[c] : (double)a.dlt[c];
    a.Xo[c] = a.cold ? d : a.X[c] + d;
}

Op add_update_op(const double *X, const float *dlt, double *Xo, int numr, int numz, int nplanes, const float *dlt2,
                 int cold)
{
    UpdateArgs a{X, dlt, Xo, numr, numz, dlt2, cold};
    int n = (numr - 2) * (numz - 2);
    return Op::make(OP_UPDATE, 0, a, (n + 255) / 256, nplanes, 1);
}

// ---------------------------------------------------------------------------
// (B_r(mode L) + lambda_z(k)) y = c along J for every (L, k); data layout
// [L][J'][k] (k fastest, one thread per k: coalesced).  LU factors of the
// shifted operator come precomputed per (L, J', k): ibet and gam (FP32).
// The sequential sweeps work in chunks of TS elements whose loads are all
// issued before the recurrence consumes them.
// ---------------------------------------------------------------------------
constexpr int TS = 16;

__device__ __forceinline__ void tridiag_shift_tile(const TridiagArgs &a, int bx, int by, int)
{
    const float *__restrict__ cin = a.cin;
    float *__restrict__ yout = a.yout;
    const float *__restrict__ ibet = a.ibet;
    const float *__restrict__ gam = a.gam;
    const double *__restrict__ betar = a.betar;
    const int nri = a.nri, nzi = a.nzi;
    const int L = by;
    const int k = bx * 256 + threadIdx.x;
    if (k >= nzi) return;
    const long base = (long)L * nri * nzi + k;
    // forward: y_J = (c_J - a_J y_{J-1}) / bet_J, a_J = betar(J); a_2 y_1 absent
    float y = 0.f;
    for (int j0 = 0; j0 < nri; j0 += TS) {
        float cv[TS], iv[TS];
#pragma unroll
        for (int u = 0; u < TS; ++u) {
            int j = j0 + u;
            long o = base + (long)j * nzi;
            cv[u] = (j < nri) ? cin[o] : 0.f;
            iv[u] = (j < nri) ? ibet[o] : 0.f;
        }
#pragma unroll
        for (int u = 0; u < TS; ++u) {
            int j = j0 + u;
            if (j < nri) {
                float a = (j > 0) ? (float)betar[j + 2] : 0.f;
                y = (cv[u] - a * y) * iv[u];
                yout[base + (long)j * nzi] = y;
            }
        }
    }
    // back: x_J = y_J - gam_{J+1} x_{J+1}
    float x = y;
    for (int j1 = nri - 2; j1 >= 0; j1 -= TS) {
        float yv[TS], gv[TS];
#pragma unroll
        for (int u = 0; u < TS; ++u) {
            int j = j1 - u;
            long o = base + (long)j * nzi;
            yv[u] = (j >= 0) ? yout[o] : 0.f;
            gv[u] = (j >= 0) ? gam[o + nzi] : 0.f;
        }
#pragma unroll
        for (int u = 0; u < TS; ++u) {
            int j = j1 - u;
            if (j >= 0) {
                x = yv[u] - gv[u] * x;
                yout[base + (long)j * nzi] = x;
            }
        }
    }
}

Op tridiag_shift_op(const float *cin, float *yout, const float *ibet, const float *gam, const double *betar, int nri,
                    int nzi, int nplanes)
{
    TridiagArgs a{cin, yout, ibet, gam, betar, nri, nzi};
    return Op::make(OP_TRIDIAG, 0, a, (nzi + 255) / 256, nplanes, 1);
}

// ---------------------------------------------------------------------------
// Plan time: M_r(L) = prod_s (dt_s - B_r)(dt_s + B_r)^-1 column by column,
// one thread per (mode, column); FP64 work vectors in scratch laid out
// [J][thread] for coalescing.  Output Mr[L][J][j'] as FP32.
// ---------------------------------------------------------------------------
// bf16 bits of x, rounded to nearest even (through FP32), and their value
__device__ inline unsigned short bf16_bits(double x)
{
    const unsigned u = __float_as_uint((float)x);
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ inline double bf16_value(unsigned short b) { return (double)__uint_as_float((unsigned)b << 16); }

__global__ void build_mr_kernel(MrArgs a)
{
    const int nri = a.nri;
    const long nthr = (long)a.nmodes * nri;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nthr) return;
    const int Lloc = (int)(t / nri), jc = (int)(t % nri);
    const int L = a.mode0 + Lloc;
    double *v = a.scratch_v + t, *u = a.scratch_u + t, *gm = a.scratch_g + t;
    const long st = nthr;  // stride between J entries
    for (int j = 0; j < nri; ++j) v[j * st] = (j == jc) ? 1.0 : 0.0;
    const double *brb = a.brb + (long)L * (a.numr + 1);
    for (int s = 0; s < a.nsteps; ++s) {
        const double dt = a.dt[s];
        // u = (dt - B_r) v
        for (int j = 0; j < nri; ++j) {
            const int J = j + 2;
            double bv = brb[J] * v[j * st];
            if (j > 0) bv += a.betar[J] * v[(j - 1) * st];
            if (j < nri - 1) bv += a.alphar[J] * v[(j + 1) * st];
            u[j * st] = dt * v[j * st] - bv;
        }
        // (dt + B_r) v = u  (Thomas, tridagr.F90 form)
        double bet = brb[2] + dt;
        double y = u[0] / bet;
        v[0] = y;
        for (int j = 1; j < nri; ++j) {
            const int J = j + 2;
            double g = a.alphar[J - 1] / bet;
            gm[j * st] = g;
            bet = (brb[J] + dt) - a.betar[J] * g;
            y = (u[j * st] - a.betar[J] * y) / bet;
            v[j * st] = y;
        }
        for (int j = nri - 2; j >= 0; --j) v[j * st] = v[j * st] - gm[(j + 1) * st] * v[(j + 1) * st];
    }
    if (a.Mr64) {
        // the transpose in FP64: column jc of M_r is row jc of the output
        double *o64 = a.Mr64 + ((long)L * nri + jc) * nri;
        for (int j = 0; j < nri; ++j) o64[j] = v[j * st];
        return;
    }
    float *out = a.Mr + (long)L * nri * nri;
    if (a.frag) {
        // bf16 pairs x = hi + lo (each rounded to nearest) in the A-fragment
        // order of v_mfma_f32_16x16x32_bf16: [row tile][k chunk of 32][hi, lo]
        // [lane][8]; element (row j, k = jc) at lane (j % 16) + 16 ((jc % 32) / 8),
        // slot jc % 8.  The same bytes as the FP32 matrix.
        unsigned short *ob = (unsigned short *)out;
        const long kc = jc / 32, lk = ((jc % 32) / 8) * 16, e = jc % 8, nkc = nri / 32;
        for (int j = 0; j < nri; ++j) {
            const double x = v[j * st];
            const unsigned short hi = bf16_bits(x), lo = bf16_bits(x - bf16_value(hi));
            const long f = ((long)(j / 16) * nkc + kc) * 2 * 64 + (j % 16) + lk;
            ob[f * 8 + e] = hi;
            ob[(f + 64) * 8 + e] = lo;
        }
    } else {
        for (int j = 0; j < nri; ++j) out[(long)j * nri + jc] = (float)v[j * st];
    }
}

void launch_build_mr(const MrArgs &a, hipStream_t s)
{
    long nthr = (long)a.nmodes * a.nri;
    hipLaunchKernelGGL(build_mr_kernel, dim3((unsigned)((nthr + 127) / 128)), dim3(128), 0, s, a);
}

__global__ void fillf_kernel(float *p, size_t n, float v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

void launch_fillf(float *p, size_t n, float v, hipStream_t s)
{
    if (!n) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fillf_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
}

// single and grouped entry points of the tile functions above
#define BH_KERNEL_PAIR(NAME, ARGS_T, NT)                                                          \
    __global__ __launch_bounds__(NT) void NAME##_kernel(ARGS_T a, int nx, int ny, int total)      \
    {                                                                                             \
        BH_SINGLE_LOOP(a, nx, ny, total, (NAME##_tile(a, bx, by, bz)));                           \
    }                                                                                             \
    __global__ __launch_bounds__(NT) void NAME##_grouped(const ARGS_T *__restrict__ args, GroupSched gs) \
    {                                                                                             \
        BH_GROUP_LOOP(ARGS_T, args, gs, (NAME##_tile(ga, bx, by, bz)));                           \
    }
BH_KERNEL_PAIR(add_update, UpdateArgs, 256)
BH_KERNEL_PAIR(tridiag_shift, TridiagArgs, 256)
#undef BH_KERNEL_PAIR

// the residual, one kernel pair per rows-per-tile variant
template <int ROWS>
__global__ __launch_bounds__(256) void residual_kernel(ResidualArgs a, int nx, int ny, int total)
{
    BH_SINGLE_LOOP(a, nx, ny, total, (residual_tile_r<ROWS>(a, bx, by, bz)));
}
template <int ROWS>
__global__ __launch_bounds__(256) void residual_grouped(const ResidualArgs *__restrict__ args, GroupSched gs)
{
    BH_GROUP_LOOP(ResidualArgs, args, gs, (residual_tile_r<ROWS>(ga, bx, by, bz)));
}

namespace {
bool mixed_ref(int kind, int variant, const void **single, const void **grouped)
{
    switch (kind) {
    case OP_RESIDUAL:
        switch (variant) {
        case 1: *single = (const void *)&residual_kernel<2>; *grouped = (const void *)&residual_grouped<2>; break;
        case 2: *single = (const void *)&residual_kernel<8>; *grouped = (const void *)&residual_grouped<8>; break;
        default: *single = (const void *)&residual_kernel<4>; *grouped = (const void *)&residual_grouped<4>; break;
        }
        return true;
    case OP_UPDATE: *single = (const void *)&add_update_kernel; *grouped = (const void *)&add_update_grouped; return true;
    case OP_TRIDIAG:
        *single = (const void *)&tridiag_shift_kernel;
        *grouped = (const void *)&tridiag_shift_grouped;
        return true;
    default:
        return false;
    }
}
}  // namespace

bool mixed_launch(const Op &op, hipStream_t s)
{
    const void *k1, *k2;
    if (!mixed_ref(op.kind, op.variant, &k1, &k2)) return false;
    int nx = op.nx, ny = op.ny, total = (int)op.tiles();
    void *argv[] = {(void *)op.args, (void *)&nx, (void *)&ny, (void *)&total};
    const int nb = persistent_blocks(total, kernel_occupancy(k1, 256, 0));
    (void)hipLaunchKernel(k1, dim3(nb), dim3(256), argv, 0, s);
    return true;
}

bool mixed_launch_grouped(int kind, int variant, const void *d_args, const GroupSched &gs, int nblocks,
                          hipStream_t s)
{
    const void *k1, *k2;
    if (!mixed_ref(kind, variant, &k1, &k2)) return false;
    GroupSched g = gs;
    const void *a = d_args;
    void *argv[] = {(void *)&a, (void *)&g};
    (void)hipLaunchKernel(k2, dim3(nblocks), dim3(256), argv, 0, s);
    return true;
}

bool mixed_occupancy(int kind, int variant, int *blocks)
{
    const void *k1, *k2;
    if (!mixed_ref(kind, variant, &k1, &k2)) return false;
    *blocks = kernel_occupancy(k2, 256, 0);
    return true;
}

}  // namespace bh
