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

constexpr int BM = 64, BN = 64, BK = 16;
constexpr int LDA_T = 4, LDB_T = 4;  // per-thread staged elements of each tile (64*16/256)

template <typename T, typename BT, int BMODE, int EPI, bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs g)
{
    typedef typename Mfma<T>::acc_t acc_t;
    __shared__ T As[BK][BM + 2];
    __shared__ T Bs[BK][BN + 2];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int row0 = blockIdx.y * BM, col0 = blockIdx.x * BN;
    const long z = blockIdx.z;
    const T *A = (const T *)g.A + z * g.sA;
    const BT *B = (const BT *)g.B + z * g.sB;
    const float *Badd = g.Badd ? g.Badd + z * g.sB : nullptr;
    const int R = g.R, M = g.M, Kd = g.Kd, N = g.nfold;

    auto loadA = [&](int k0, T *st) {
#pragma unroll
        for (int i = 0; i < LDA_T; ++i) {
            int t = tid + 256 * i;
            int r, k;
            if (TA) { k = t / BM; r = t % BM; }
            else { r = t / BK; k = t % BK; }
            int gr = row0 + r, gk = k0 + k;
            T v = (T)0;
            if (gr < R && gk < Kd) v = TA ? A[(long)gk * g.lda + gr] : A[(long)gr * g.lda + gk];
            st[i] = v;
        }
    };
    auto loadB = [&](int k0, T *st) {
#pragma unroll
        for (int i = 0; i < LDB_T; ++i) {
            int t = tid + 256 * i;
            int k, c;
            if (TB) { c = t / BK; k = t % BK; }
            else { k = t / BN; c = t % BN; }
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
        for (int i = 0; i < LDA_T; ++i) {
            int t = tid + 256 * i;
            int r, k;
            if (TA) { k = t / BM; r = t % BM; }
            else { r = t / BK; k = t % BK; }
            As[k][r] = st[i];
        }
    };
    auto storeB = [&](const T *st) {
#pragma unroll
        for (int i = 0; i < LDB_T; ++i) {
            int t = tid + 256 * i;
            int k, c;
            if (TB) { c = t / BK; k = t % BK; }
            else { k = t / BN; c = t % BN; }
            Bs[k][c] = st[i];
        }
    };

    acc_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = (acc_t){(T)0, (T)0, (T)0, (T)0};

    T sa[LDA_T], sb[LDB_T];
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
            T a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = As[kk + (lane >> 4)][wr * 32 + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = Bs[kk + (lane >> 4)][wc * 32 + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[i][j] = Mfma<T>::mma(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }

    T *C = (T *)g.C + z * g.sC;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = row0 + wr * 32 + i * 16 + Mfma<T>::row(lane, v);
                const int c = col0 + wc * 32 + j * 16 + (lane & 15);
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

template <typename T, typename BT, int BMODE, int EPI, bool TA, bool TB>
static void launch(const GemmArgs &g, hipStream_t s)
{
    dim3 grid((g.M + BN - 1) / BN, (g.R + BM - 1) / BM, g.batch > 0 ? g.batch : 1);
    hipLaunchKernelGGL((gemm_kernel<T, BT, BMODE, EPI, TA, TB>), grid, dim3(256), 0, s, g);
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
