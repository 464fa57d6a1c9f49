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
#include "bh_ops.h"

#include <hip/hip_runtime.h>

#include <utility>

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

template <typename T, typename BT, int BMODE, int EPI, bool TA, bool TB, int DST, int WR, int WC, int TM, int TN>
__device__ __forceinline__ void gemm_tile(const GemmArgs &g, int rt, int ct, int bz)
{
    typedef typename Mfma<T>::acc_t acc_t;
    constexpr int BMt = WR * TM * 16, BNt = WC * TN * 16;
    constexpr int NA = BMt * BK / 256, NB = BK * BNt / 256;  // staged elements per thread
    __shared__ T As[BK][BMt + 2];
    __shared__ T Bs[BK][BNt + 2];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = wave / WC, wc = wave % WC;
    const long z = bz;
    const T *A = (const T *)g.A + z * g.sA;
    const BT *B = (const BT *)g.B + z * g.sB;
    const float *Badd = g.Badd ? g.Badd + z * g.sB : nullptr;
    T *C = (T *)g.C + z * g.sC;
    const int N = g.nfold;
    int R = g.R, M = g.M, Kd = g.Kd;
    T fsgn = (T)1;
    if constexpr (DST == DST_FWD) {
        // column tiles [0, ncth) cover the even sine modes, the rest the odd
        const int he = (N + 1) / 2, ncth = (he + BNt - 1) / BNt;
        const int half = ct >= ncth ? 1 : 0;
        ct -= half * ncth;
        Kd = M = half ? N / 2 : he;
        B += half * he;
        C += half * he;
        fsgn = half ? (T)-1 : (T)1;
    }
    const int row0 = rt * BMt, col0 = ct * BNt;

    auto loadA = [&](const T *Ap, int k0, int Klen, T *st) {
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            int t = tid + 256 * i;
            int r, k;
            if (TA) { k = t / BMt; r = t % BMt; }
            else { r = t / BK; k = t % BK; }
            int gr = row0 + r, gk = k0 + k;
            T v = (T)0;
            if (gr < R && gk < Klen) {
                if constexpr (DST == DST_FWD) {
                    // planes a and n-1-a folded; the middle plane of an odd n once
                    v = Ap[(long)gk * g.lda + gr];
                    const int mk = N - 1 - gk;
                    if (mk != gk) v += fsgn * Ap[(long)mk * g.lda + gr];
                    else if (fsgn < (T)0) v = (T)0;
                } else {
                    v = TA ? Ap[(long)gk * g.lda + gr] : Ap[(long)gr * g.lda + gk];
                }
            }
            st[i] = v;
        }
    };
    auto loadB = [&](const BT *Bp, int k0, int Klen, T *st) {
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            int t = tid + 256 * i;
            int k, c;
            if (TB) { c = t / BK; k = t % BK; }
            else { k = t / BNt; c = t % BNt; }
            int gk = k0 + k, gc = col0 + c;
            T v = (T)0;
            if (gk < Klen && gc < M) {
                if (BMODE == B_PLAIN) {
                    v = (T)(TB ? Bp[(long)gc * g.ldb + gk] : Bp[(long)gk * g.ldb + gc]);
                } else if (BMODE == B_FOLD_PLUS) {
                    double x = (double)Bp[(long)gk * g.ldb + gc];
                    if (gk > 0 && gk < N / 2) x += (double)Bp[(long)(N - gk) * g.ldb + gc];
                    v = (T)x;
                } else if (BMODE == B_FOLD_MINUS) {
                    v = (T)((double)Bp[(long)(gk + 1) * g.ldb + gc] - (double)Bp[(long)(N - gk - 1) * g.ldb + gc]);
                } else {  // B_ADD_F32
                    v = (T)((double)Bp[(long)gk * g.ldb + gc] + (double)Badd[(long)gk * g.ldb + gc]);
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
    // acc += A[:, 0:Klen] B[0:Klen, :] from the operands at Ap, Bp
    auto kloop = [&](const T *Ap, const BT *Bp, int Klen, acc_t(&acc)[TM][TN]) {
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) acc[i][j] = (acc_t){(T)0, (T)0, (T)0, (T)0};
        T sa[NA], sb[NB];
        loadA(Ap, 0, Klen, sa);
        loadB(Bp, 0, Klen, sb);
        for (int k0 = 0; k0 < Klen; k0 += BK) {
            storeA(sa);
            storeB(sb);
            __syncthreads();
            if (k0 + BK < Klen) {  // the next tile's loads fly while this one is multiplied
                loadA(Ap, k0 + BK, Klen, sa);
                loadB(Bp, k0 + BK, Klen, sb);
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
    };

    acc_t acc[TM][TN];
    if constexpr (DST == DST_INV) {
        // rows z < ceil(n/2): even modes (table columns [0, he)) and odd
        // modes (columns [he, n)); the pair gives rows z and n-1-z
        const int he = (N + 1) / 2;
        acc_t acco[TM][TN];
        kloop(A, B, he, acc);
        kloop(A + he, B + he, N / 2, acco);
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = row0 + (wr * TM + i) * 16 + Mfma<T>::row(lane, v);
                    const int c = col0 + (wc * TN + j) * 16 + (lane & 15);
                    if (r >= he || c >= M) continue;
                    const T e = acc[i][j][v], o = acco[i][j][v];
                    C[(long)r * g.ldc + c] = e + o;
                    const int rm = N - 1 - r;
                    if (rm != r) C[(long)rm * g.ldc + c] = e - o;
                }
        return;
    } else {
        kloop(A, B, Kd, acc);
    }

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
                    // (K in {1,2} kept for isym = 2, 3)
                    int J = c % g.mask_numr + 1, K = c / g.mask_numr + 1;
                    if (J >= g.mask_numr - 1 || (K <= 2 && !g.mask_keep_bottom) || K >= g.mask_numz - 1) val = (T)0;
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

// Template instances: element types, B mode and epilogue per GemmInst, and
// the workgroup arrangement per tile shape.
template <int I>
struct Inst;
#define BH_INST(I, T_, BT_, BM_, EP_, TA_, TB_, DST_)          \
    template <>                                                \
    struct Inst<I> {                                           \
        typedef T_ T;                                          \
        typedef BT_ BT;                                        \
        static constexpr int BMODE = BM_, EPI = EP_, DST = DST_; \
        static constexpr bool TA = TA_, TB = TB_;              \
    };
BH_INST(G64_PLAIN, double, double, B_PLAIN, E_PLAIN, false, false, DST_NONE)
BH_INST(G64_FOLDP_MASK, double, double, B_FOLD_PLUS, E_MASK, false, false, DST_NONE)
BH_INST(G64_FOLDM_MASK, double, double, B_FOLD_MINUS, E_MASK, false, false, DST_NONE)
BH_INST(G64_FOLDP, double, double, B_FOLD_PLUS, E_PLAIN, false, false, DST_NONE)
BH_INST(G64_FOLDM, double, double, B_FOLD_MINUS, E_PLAIN, false, false, DST_NONE)
BH_INST(G64_INV, double, double, B_PLAIN, E_INV_COMBINE, false, false, DST_NONE)
BH_INST(G64_ADD, double, double, B_ADD_F32, E_PLAIN, false, false, DST_NONE)
BH_INST(G64_ADD_INV, double, double, B_ADD_F32, E_INV_COMBINE, false, false, DST_NONE)
BH_INST(G32D_FOLDP_MASK, float, double, B_FOLD_PLUS, E_MASK, false, false, DST_NONE)
BH_INST(G32D_FOLDM_MASK, float, double, B_FOLD_MINUS, E_MASK, false, false, DST_NONE)
BH_INST(G32_TA, float, float, B_PLAIN, E_PLAIN, true, false, DST_NONE)
BH_INST(G32_TB, float, float, B_PLAIN, E_PLAIN, false, true, DST_NONE)
BH_INST(G32_FILTER, float, float, B_PLAIN, E_FILTER, false, false, DST_NONE)
BH_INST(G32_PLAIN, float, float, B_PLAIN, E_PLAIN, false, false, DST_NONE)
BH_INST(G32_DST_FWD, float, float, B_PLAIN, E_PLAIN, true, false, DST_FWD)
BH_INST(G32_DST_INV, float, float, B_PLAIN, E_PLAIN, false, true, DST_INV)
#undef BH_INST

// tile shapes (rows x columns): 0: 64 x 256, 1: 128 x 128, 2: 64 x 64,
// 3: 64 x 128, 4: 128 x 64
template <int S>
struct Shape;
template <> struct Shape<0> { static constexpr int WR = 1, WC = 4, TM = 4, TN = 4; };
template <> struct Shape<1> { static constexpr int WR = 2, WC = 2, TM = 4, TN = 4; };
template <> struct Shape<2> { static constexpr int WR = 2, WC = 2, TM = 2, TN = 2; };
template <> struct Shape<3> { static constexpr int WR = 2, WC = 2, TM = 2, TN = 4; };
template <> struct Shape<4> { static constexpr int WR = 2, WC = 2, TM = 4, TN = 2; };

template <int I, int S>
__device__ __forceinline__ void gemm_tile_is(const GemmArgs &g, int bx, int by, int bz)
{
    typedef Inst<I> In;
    typedef Shape<S> Sh;
    gemm_tile<typename In::T, typename In::BT, In::BMODE, In::EPI, In::TA, In::TB, In::DST, Sh::WR, Sh::WC, Sh::TM,
              Sh::TN>(g, bx, by, bz);
}

// Tiles are numbered row tile fastest (the row tiles that share a column
// panel of B run next to each other, on one XCD) or column tile fastest
// (consecutive tiles share a row block of A), per GemmArgs::col_fastest.
template <int I, int S>
__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs g, int nx, int ny, int total)
{
    BH_SINGLE_LOOP(g, nx, ny, total,
                   (gemm_tile_is<I, S>(g, g.col_fastest ? by : bx, g.col_fastest ? bx : by, bz)));
}

template <int I, int S>
__global__ __launch_bounds__(256) void gemm_grouped_kernel(const GemmArgs *__restrict__ args, GroupSched gs)
{
    BH_GROUP_LOOP(GemmArgs, args, gs,
                  (gemm_tile_is<I, S>(ga, ga.col_fastest ? by : bx, ga.col_fastest ? bx : by, bz)));
}

typedef void (*GemmSingleFn)(GemmArgs, int, int, int);
typedef void (*GemmGroupedFn)(const GemmArgs *, GroupSched);
constexpr int G_NSHAPE = 5;

template <int... V>
struct GemmTable {
    static constexpr GemmSingleFn single[] = {&gemm_kernel<V / G_NSHAPE, V % G_NSHAPE>...};
    static constexpr GemmGroupedFn grouped[] = {&gemm_grouped_kernel<V / G_NSHAPE, V % G_NSHAPE>...};
};
template <int... V>
constexpr GemmSingleFn GemmTable<V...>::single[];
template <int... V>
constexpr GemmGroupedFn GemmTable<V...>::grouped[];

template <int... V>
GemmTable<V...> make_table(std::integer_sequence<int, V...>);
typedef decltype(make_table(std::make_integer_sequence<int, G_NINST * G_NSHAPE>())) Table;

void gemm_tile_rows_cols(int tile, int *rows, int *cols)
{
    static const int r[G_NSHAPE] = {64, 128, 64, 64, 128}, c[G_NSHAPE] = {256, 128, 64, 128, 64};
    if (tile < 0 || tile >= G_NSHAPE) tile = 2;
    *rows = r[tile];
    *cols = c[tile];
}

Op gemm_op(const GemmArgs &g, int inst)
{
    int rows, cols;
    const int tile = (g.tile >= 0 && g.tile < G_NSHAPE) ? g.tile : 2;
    gemm_tile_rows_cols(tile, &rows, &cols);
    GemmArgs a = g;
    a.tile = tile;
    int nr = (g.R + rows - 1) / rows, nc = (g.M + cols - 1) / cols;
    if (inst == G32_DST_FWD) nc = 2 * (((g.nfold + 1) / 2 + cols - 1) / cols);  // even and odd modes
    if (inst == G32_DST_INV) nr = ((g.nfold + 1) / 2 + rows - 1) / rows;        // mirrored row pairs
    return Op::make(OP_GEMM, inst * 8 + tile, a, a.col_fastest ? nc : nr, a.col_fastest ? nr : nc,
                    g.batch > 0 ? g.batch : 1);
}

static int variant_index(int variant)
{
    const int inst = variant / 8, tile = variant % 8;
    if (inst < 0 || inst >= G_NINST || tile >= G_NSHAPE) return -1;
    return inst * G_NSHAPE + tile;
}

bool gemm_launch(const Op &op, hipStream_t s)
{
    if (op.kind != OP_GEMM) return false;
    const int v = variant_index(op.variant);
    if (v < 0) return true;
    const int nb = persistent_blocks(op.tiles(), kernel_occupancy((const void *)Table::single[v], 256, 0));
    hipLaunchKernelGGL(Table::single[v], dim3(nb), dim3(256), 0, s, op.as<GemmArgs>(), op.nx, op.ny,
                       (int)op.tiles());
    return true;
}

bool gemm_launch_grouped(int kind, int variant, const void *d_args, const GroupSched &gs, int nblocks, hipStream_t s)
{
    if (kind != OP_GEMM) return false;
    const int v = variant_index(variant);
    if (v < 0) return true;
    hipLaunchKernelGGL(Table::grouped[v], dim3(nblocks), dim3(256), 0, s, (const GemmArgs *)d_args, gs);
    return true;
}

bool gemm_occupancy(int kind, int variant, int *blocks)
{
    if (kind != OP_GEMM) return false;
    const int v = variant_index(variant);
    *blocks = v >= 0 ? kernel_occupancy((const void *)Table::grouped[v], 256, 0) : 0;
    return true;
}

void gemm64(const GemmArgs &g, int bmode, int epi, hipStream_t s)
{
    launch_op(gemm_op(g, gemm_inst64(bmode, epi)), s);
}

void gemm32(const GemmArgs &g, int bmode, int epi, bool ta, bool tb, bool b_double, hipStream_t s)
{
    launch_op(gemm_op(g, gemm_inst32(bmode, epi, ta, tb, b_double)), s);
}

int gemm_inst64(int bmode, int epi)
{
    if (bmode == B_FOLD_PLUS) return epi == E_MASK ? G64_FOLDP_MASK : G64_FOLDP;
    if (bmode == B_FOLD_MINUS) return epi == E_MASK ? G64_FOLDM_MASK : G64_FOLDM;
    if (bmode == B_ADD_F32) return epi == E_INV_COMBINE ? G64_ADD_INV : G64_ADD;
    return epi == E_INV_COMBINE ? G64_INV : G64_PLAIN;
}

int gemm_inst32(int bmode, int epi, bool ta, bool tb, bool b_double)
{
    if (b_double) return bmode == B_FOLD_MINUS ? G32D_FOLDM_MASK : G32D_FOLDP_MASK;
    if (ta) return G32_TA;
    if (tb) return G32_TB;
    if (epi == E_FILTER) return G32_FILTER;
    (void)bmode;
    return G32_PLAIN;
}

}  // namespace bh
