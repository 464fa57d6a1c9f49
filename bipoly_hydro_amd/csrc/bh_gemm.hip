// bh_gemm.hip -- one MFMA GEMM template for every dense product of the solve.
//
//   C = op(A) * op(B'),  A: R x Kd,  B': Kd x M
//
// FP64 products run on v_mfma_f64_16x16x4f64, FP32 products on
// v_mfma_f32_16x16x4f32.  Operand layouts measured on gfx950 with a one-hot
// probe: A[l%16][l/16] and B[l/16][l%16] for lane l in both; result element v
// of lane l is D[4v + l/16][l%16] for FP64 and D[4(l/16) + v][l%16] for FP32.
//
// B' may be formed while the tile is loaded (BMODE):
//   B_PLAIN      B'[k][c] = B[k][c]
//   B_FOLD_PLUS  B'[r][c] = B[r][c] + B[N-r][c]  (r = 1..N/2-1), B[0], B[N/2]
//   B_FOLD_MINUS B'[r][c] = B[r+1][c] - B[N-r-1][c]  (r = 0..N/2-2)
//   B_ADD_F32    B'[k][c] = B[k][c] + Badd[k][c]  (Badd FP32)
// which splits a real DFT of length N into its even (cosine) and odd (sine)
// halves, each a dense product of half the size (realft.F90 computes the same
// transform by an FFT).  The epilogue (EPI) stores plainly, applies the rhop
// zero mask of potential_solver.F90:70-78, applies the sine-mode filter of
// the mixed solve, or combines the two halves of an inverse real DFT.
#include "bh_gemm.h"

#include <hip/hip_runtime.h>

namespace bh {

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
    typedef double acc_t __attribute__((ext_vector_type(4)));
    static __device__ inline acc_t mma(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int lane, int v) { return 4 * v + (lane >> 4); }
};
template <>
struct Mfma<float> {
    typedef float acc_t __attribute__((ext_vector_type(4)));
    static __device__ inline acc_t mma(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

// Workgroup of 4 waves arranged WR x WC; each wave computes TM x TN MFMA
// tiles of 16 x 16, so the workgroup tile is (WR*TM*16) x (WC*TN*16).  K is
// staged through LDS in chunks of BK, with the next chunk's global loads
// issued before the current chunk is multiplied.
constexpr int BK = 16;

template <typename T, typename BT, int BMODE, int EPI, bool TA, bool TB, int WR, int WC, int TM, int TN>
__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs g)
{
    typedef typename Mfma<T>::acc_t acc_t;
    constexpr int BMt = WR * TM * 16, BNt = WC * TN * 16;
    constexpr int NA = BMt * BK / 256, NB = BK * BNt / 256;  // staged elements per thread
    __shared__ T As[BK][BMt + 2];
    __shared__ T Bs[BK][BNt + 2];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = wave / WC, wc = wave % WC;
    const int row0 = blockIdx.y * BMt, col0 = blockIdx.x * BNt;
    const long z = blockIdx.z;
    const T *A = (const T *)g.A + z * g.sA;
    const BT *B = (const BT *)g.B + z * g.sB;
    const float *Badd = g.Badd ? g.Badd + z * g.sB : nullptr;
    const int R = g.R, M = g.M, Kd = g.Kd, N = g.nfold;

    auto loadA = [&](int k0, T *st) {
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            int t = tid + 256 * i;
            int r, k;
            if (TA) { k = t / BMt; r = t % BMt; }
            else { r = t / BK; k = t % BK; }
            int gr = row0 + r, gk = k0 + k;
            T v = (T)0;
            if (gr < R && gk < Kd) v = TA ? A[(long)gk * g.lda + gr] : A[(long)gr * g.lda + gk];
            st[i] = v;
        }
    };
    auto loadB = [&](int k0, T *st) {
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            int t = tid + 256 * i;
            int k, c;
            if (TB) { c = t / BK; k = t % BK; }
            else { k = t / BNt; c = t % BNt; }
            int gk = k0 + k, gc = col0 + c;
            T v = (T)0;
            if (gk < Kd && gc < M) {
                if (BMODE == B_PLAIN) {
                    v = (T)(TB ? B[(long)gc * g.ldb + gk] : B[(long)gk * g.ldb + gc]);
                } else if (BMODE == B_FOLD_PLUS) {
                    double x = (double)B[(long)gk * g.ldb + gc];
                    if (gk > 0 && gk < N / 2) x += (double)B[(long)(N - gk) * g.ldb + gc];
                    v = (T)x;
                } else if (BMODE == B_FOLD_MINUS) {
                    v = (T)((double)B[(long)(gk + 1) * g.ldb + gc] - (double)B[(long)(N - gk - 1) * g.ldb + gc]);
                } else {  // B_ADD_F32
                    v = (T)((double)B[(long)gk * g.ldb + gc] + (double)Badd[(long)gk * g.ldb + gc]);
                }
            }
            st[i] = v;
        }
    };
    auto storeA = [&](const T *st) {
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            int t = tid + 256 * i;
            int r, k;
            if (TA) { k = t / BMt; r = t % BMt; }
            else { r = t / BK; k = t % BK; }
            As[k][r] = st[i];
        }
    };
    auto storeB = [&](const T *st) {
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            int t = tid + 256 * i;
            int k, c;
            if (TB) { c = t / BK; k = t % BK; }
            else { k = t / BNt; c = t % BNt; }
            Bs[k][c] = st[i];
        }
    };

    acc_t acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = (acc_t){(T)0, (T)0, (T)0, (T)0};

    T sa[NA], sb[NB];
    loadA(0, sa);
    loadB(0, sb);
    for (int k0 = 0; k0 < Kd; k0 += BK) {
        storeA(sa);
        storeB(sb);
        __syncthreads();
        if (k0 + BK < Kd) {  // the next tile's loads fly while this one is multiplied
            loadA(k0 + BK, sa);
            loadB(k0 + BK, sb);
        }
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            T a[TM], b[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) a[i] = As[kk + (lane >> 4)][(wr * TM + i) * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < TN; ++j) b[j] = Bs[kk + (lane >> 4)][(wc * TN + j) * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j) acc[i][j] = Mfma<T>::mma(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }

    T *C = (T *)g.C + z * g.sC;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = row0 + (wr * TM + i) * 16 + Mfma<T>::row(lane, v);
                const int c = col0 + (wc * TN + j) * 16 + (lane & 15);
                if (r >= R || c >= M) continue;
                T val = acc[i][j][v];
                const long o = (long)r * g.ldc + c;
                if (EPI == E_MASK) {
                    // c = (K-1)*numr + (J-1): zero K in {1,2,numz-1,numz}, J in {numr-1,numr}
                    int J = c % g.mask_numr + 1, K = c / g.mask_numr + 1;
                    if (J >= g.mask_numr - 1 || K <= 2 || K >= g.mask_numz - 1) val = (T)0;
                    C[o] = val;
                } else if (EPI == E_FILTER) {
                    T w = (T)g.mu[c] * val;
                    C[o] = w;
                    ((T *)g.C2)[z * g.sC + o] = ((const T *)g.Rin)[z * g.sC + o] - w;
                } else if (EPI == E_INV_COMBINE) {
                    // row r = n - 1 of the odd half Q; C already holds the even half P
                    const int n = r + 1;
                    const long on = (long)n * g.ldc + c, om = (long)(N - n) * g.ldc + c;
                    T p = C[on];
                    C[on] = p - val;
                    C[om] = p + val;
                } else {
                    C[o] = val;
                }
            }
}

// tile shapes (rows x columns): 0: 64 x 256, 1: 128 x 128, 2: 64 x 64,
// 3: 64 x 128, 4: 128 x 64

template <typename T, typename BT, int BMODE, int EPI, bool TA, bool TB>
static void launch(const GemmArgs &g, hipStream_t s)
{
    int shape = g.tile;
    const unsigned nb = g.batch > 0 ? g.batch : 1;
    if (shape == 0) {
        dim3 grid((g.M + 255) / 256, (g.R + 63) / 64, nb);
        hipLaunchKernelGGL((gemm_kernel<T, BT, BMODE, EPI, TA, TB, 1, 4, 4, 4>), grid, dim3(256), 0, s, g);
    } else if (shape == 1) {
        dim3 grid((g.M + 127) / 128, (g.R + 127) / 128, nb);
        hipLaunchKernelGGL((gemm_kernel<T, BT, BMODE, EPI, TA, TB, 2, 2, 4, 4>), grid, dim3(256), 0, s, g);
    } else if (shape == 3) {
        dim3 grid((g.M + 127) / 128, (g.R + 63) / 64, nb);
        hipLaunchKernelGGL((gemm_kernel<T, BT, BMODE, EPI, TA, TB, 2, 2, 2, 4>), grid, dim3(256), 0, s, g);
    } else if (shape == 4) {
        dim3 grid((g.M + 63) / 64, (g.R + 127) / 128, nb);
        hipLaunchKernelGGL((gemm_kernel<T, BT, BMODE, EPI, TA, TB, 2, 2, 4, 2>), grid, dim3(256), 0, s, g);
    } else {
        dim3 grid((g.M + 63) / 64, (g.R + 63) / 64, g.batch > 0 ? g.batch : 1);
        hipLaunchKernelGGL((gemm_kernel<T, BT, BMODE, EPI, TA, TB, 2, 2, 2, 2>), grid, dim3(256), 0, s, g);
    }
}

void gemm64(const GemmArgs &g, int bmode, int epi, hipStream_t s)
{
    if (bmode == B_PLAIN && epi == E_PLAIN) launch<double, double, B_PLAIN, E_PLAIN, false, false>(g, s);
    else if (bmode == B_FOLD_PLUS && epi == E_MASK) launch<double, double, B_FOLD_PLUS, E_MASK, false, false>(g, s);
    else if (bmode == B_FOLD_MINUS && epi == E_MASK) launch<double, double, B_FOLD_MINUS, E_MASK, false, false>(g, s);
    else if (bmode == B_FOLD_PLUS && epi == E_PLAIN) launch<double, double, B_FOLD_PLUS, E_PLAIN, false, false>(g, s);
    else if (bmode == B_FOLD_MINUS && epi == E_PLAIN) launch<double, double, B_FOLD_MINUS, E_PLAIN, false, false>(g, s);
    else if (bmode == B_PLAIN && epi == E_INV_COMBINE)
        launch<double, double, B_PLAIN, E_INV_COMBINE, false, false>(g, s);
    else if (bmode == B_ADD_F32 && epi == E_PLAIN) launch<double, double, B_ADD_F32, E_PLAIN, false, false>(g, s);
    else if (bmode == B_ADD_F32 && epi == E_INV_COMBINE)
        launch<double, double, B_ADD_F32, E_INV_COMBINE, false, false>(g, s);
}

void gemm32(const GemmArgs &g, int bmode, int epi, bool ta, bool tb, bool b_double, hipStream_t s)
{
    if (b_double) {
        if (bmode == B_FOLD_PLUS && epi == E_MASK) launch<float, double, B_FOLD_PLUS, E_MASK, false, false>(g, s);
        else if (bmode == B_FOLD_MINUS && epi == E_MASK) launch<float, double, B_FOLD_MINUS, E_MASK, false, false>(g, s);
        return;
    }
    if (ta && !tb) launch<float, float, B_PLAIN, E_PLAIN, true, false>(g, s);
    else if (!ta && tb) launch<float, float, B_PLAIN, E_PLAIN, false, true>(g, s);
    else if (epi == E_FILTER) launch<float, float, B_PLAIN, E_FILTER, false, false>(g, s);
    else launch<float, float, B_PLAIN, E_PLAIN, false, false>(g, s);
}

}  // namespace bh
