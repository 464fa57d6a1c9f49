// bh_dft.hip -- the real DFT along phi (realft.F90 with helmadi's reorder,
// helmadi.F90:131-183 and 357-392) as dense MFMA products, one column panel
// per workgroup.
//
// Forward, in helmadi's mode order (rows 0..N/2: a_k; rows N/2+1..N-1: -b_k):
//   u_n = x_n + x_{N-n}, v_n = x_n - x_{N-n}   (n = 1..H, H = N/2 - 1)
//   a_k  = x_0 + (-1)^k x_{N/2} + sum_n cos(2 pi k n / N) u_n
//   -b_k = -sum_n sin(2 pi k n / N) v_n
// Inverse, with the 2/N normalisation:
//   P_n = a_0/N + (-1)^n a_{N/2}/N + sum_k (2/N) cos(2 pi k n / N) a_k
//   Q_n = sum_k (2/N) sin(2 pi k n / N) (-b_k)        (k = 1..H)
//   x_n = P_n - Q_n,  x_{N-n} = P_n + Q_n
//
// A workgroup of 8 waves loads its panel of CB columns once, forms both
// halves (u, v or a, -b) in LDS, and computes every output row: the
// K = H products run on MFMA over row tiles of 16 (rows k or n = 0..N/2-1
// of each half), the terms in x_0, x_{N/2} (a_0, a_{N/2}) are added in the
// epilogue, and the one remaining row (k = N/2, or n = N/2) is a reduction
// over the panel in FP64.  The matrix operand is pre-arranged on the host in
// MFMA fragment order, [row tile][k step][lane], so every fragment load is
// one contiguous 64-lane read from L2.  Row tile t of the forward product
// belongs to wave t % 8; the inverse pairs the P and Q tiles of the same n
// in one wave so that x_n and x_{N-n} are formed in registers.
//
// At N = 512 (the production grids) one radix step comes first (forward:
// radix 4 in the input index; inverse: radix 8 in the output index; the
// derivations are at the kernels), so the products run at size N/4 (N/8)
// on four (eight) virtual columns per column: a quarter (an eighth) of the
// matrix work of the direct form.
#include "bh_dft.h"
#include "bh_ops.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <algorithm>
#include <functional>
#include <vector>

namespace bh {

namespace {

template <typename T>
struct Mma;
template <>
struct Mma<double> {
    typedef double acc_t __attribute__((ext_vector_type(4)));
    static __device__ inline acc_t mma(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int lane, int v) { return 4 * v + (lane >> 4); }
};
template <>
struct Mma<float> {
    typedef float acc_t __attribute__((ext_vector_type(4)));
    static __device__ inline acc_t mma(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

// Products run over K in groups of 16 (four MFMA k-steps).  Within a group
// the K index of (lane l, step g) is 4 (l / 16) + g rather than the MFMA's
// own 4 g + l / 16: any one-to-one assignment gives the same sum when A and
// B agree, and this one makes each lane's four values of a group contiguous,
// so a group costs one 16-byte fragment load of A (from L2) and one of B
// (from LDS, stored column-major with a pitch that spreads the 16 columns
// of a fragment over distinct banks) per tile.
template <typename T>
struct Pad;
template <>
struct Pad<float> {
    static constexpr int v = 4;  // pitch = K + 4 floats: 16 columns x 16 B hit distinct bank quads
};
template <>
struct Pad<double> {
    static constexpr int v = 2;
};

}  // namespace

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------
template <typename T, int CB>
__device__ __forceinline__ void dft_fwd_tile(const DftArgs &a, int bx, unsigned char *lds)
{
    typedef typename Mma<T>::acc_t acc_t;
    typedef T v4 __attribute__((ext_vector_type(4)));
    constexpr int TN = CB / 16;
    const int N = a.N, H = N / 2 - 1, KG = (H + 15) / 16, KP = 16 * KG, KPp = KP + Pad<T>::v;
    const int RTc = (N / 2 + 15) / 16, RTs = (H + 15) / 16, RT = RTc + RTs;
    T *U = (T *)lds;                                 // [CB][KPp]: u_{j+1} at [c][j]
    T *Vv = U + (size_t)CB * KPp;                    // [CB][KPp]: v_{j+1}
    double *x0 = (double *)(Vv + (size_t)CB * KPp);  // [CB] x_0
    double *xh = x0 + CB;                            // [CB] x_{N/2}
    double *part = xh + CB;                          // [16][CB] partial sums of row N/2
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int col0 = bx * CB;
    const double *x = a.x;
    const long ldx = a.ldx;

    // 1. panel: u, v (padded to KP with zeros), x_0, x_{N/2}; each thread
    //    issues all the loads of a batch before it uses any
    constexpr int BATCH = 4;
    for (int e0 = 0; e0 < KP * CB; e0 += 512 * BATCH) {
        double x1[BATCH], x2[BATCH];
#pragma unroll
        for (int b = 0; b < BATCH; ++b) {
            const int e = e0 + 512 * b + tid, j = e / CB, gc = col0 + e % CB;
            const bool ok = e < KP * CB && j < H && gc < a.M;
            x1[b] = ok ? x[(long)(j + 1) * ldx + gc] : 0.0;
            x2[b] = ok ? x[(long)(N - 1 - j) * ldx + gc] : 0.0;
        }
#pragma unroll
        for (int b = 0; b < BATCH; ++b) {
            const int e = e0 + 512 * b + tid, j = e / CB, c = e % CB;
            if (e < KP * CB) {
                U[c * KPp + j] = (T)(x1[b] + x2[b]);
                Vv[c * KPp + j] = (T)(x1[b] - x2[b]);
            }
        }
    }
    if (tid < CB) {
        const int gc = col0 + tid;
        x0[tid] = gc < a.M ? x[gc] : 0.0;
        xh[tid] = gc < a.M ? x[(long)(N / 2) * ldx + gc] : 0.0;
    }
    __syncthreads();

    // 2. row N/2: a_{N/2} = x_0 + (-1)^{N/2} x_{N/2} + sum_n (-1)^n u_n, in FP64
    {
        const int c = tid % CB, q = tid / CB;  // q in [0, 512 / CB)
        constexpr int NQ = 512 / CB;
        double s = 0.0;
        for (int j = q; j < H; j += NQ) {
            const double u = (double)U[c * KPp + j];
            s += ((j + 1) & 1) ? -u : u;
        }
        if (q < 16) part[q * CB + c] = s;
        __syncthreads();
        if (NQ > 16 && q >= 16) part[(q & 15) * CB + c] += s;  // NQ = 32: two passes, fixed order
        __syncthreads();
        if (tid < CB) {
            double t = 0.0;
            for (int i = 0; i < 16; ++i) t += part[i * CB + tid];
            const int gc = col0 + tid;
            if (gc < a.M) {
                double val = x0[tid] + (((N / 2) & 1) ? -xh[tid] : xh[tid]) + t;
                if (a.mask_numr) {
                    const int J = gc % a.mask_numr + 1, K = gc / a.mask_numr + 1;
                    if (J >= a.mask_numr - 1 || (K <= 2 && !a.mask_keep_bottom) || K >= a.mask_numz - 1) val = 0.0;
                }
                ((T *)a.out)[(long)(N / 2) * a.ldo + gc] = (T)val;
            }
        }
    }

    // 3. the MFMA products: wave w owns row tiles w, w + 8, w + 16, w + 24
    int rt[4], ntile = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        rt[i] = wave + 8 * i;
        if (rt[i] < RT) ntile = i + 1;
    }
    bool need_u = false, need_v = false;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (i < ntile) {
            need_u |= rt[i] < RTc;
            need_v |= rt[i] >= RTc;
        }
    acc_t acc[4][TN];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int t = 0; t < TN; ++t) acc[i][t] = (acc_t){(T)0, (T)0, (T)0, (T)0};
    const v4 *Af = (const v4 *)a.A;  // [row tile][group][lane] of 4
    const int jl = 4 * (lane >> 4);
    if (ntile > 0) {
        v4 an[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
            an[i] = (i < ntile) ? Af[((long)rt[i] * KG) * 64 + lane] : (v4){(T)0, (T)0, (T)0, (T)0};
        for (int grp = 0; grp < KG; ++grp) {
            v4 ac[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) ac[i] = an[i];
            if (grp + 1 < KG) {
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (i < ntile) an[i] = Af[((long)rt[i] * KG + grp + 1) * 64 + lane];
            }
            v4 bu[TN], bv[TN];
#pragma unroll
            for (int t = 0; t < TN; ++t) {
                const int off = (16 * t + (lane & 15)) * KPp + 16 * grp + jl;
                bu[t] = need_u ? *(const v4 *)(U + off) : (v4){(T)0, (T)0, (T)0, (T)0};
                bv[t] = need_v ? *(const v4 *)(Vv + off) : (v4){(T)0, (T)0, (T)0, (T)0};
            }
#pragma unroll
            for (int g = 0; g < 4; ++g)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    if (i < ntile) {
                        const bool cosrow = rt[i] < RTc;
#pragma unroll
                        for (int t = 0; t < TN; ++t)
                            acc[i][t] = Mma<T>::mma(ac[i][g], cosrow ? bu[t][g] : bv[t][g], acc[i][t]);
                    }
                }
        }
    }

    // 4. epilogue: a_k (+ x_0, x_{N/2} terms) in rows 0..N/2-1, -b_k in rows N/2+1..N-1
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (i >= ntile) continue;
        const bool cosrow = rt[i] < RTc;
#pragma unroll
        for (int t = 0; t < TN; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = 16 * (cosrow ? rt[i] : rt[i] - RTc) + Mma<T>::row(lane, v);
                const int c = 16 * t + (lane & 15), gc = col0 + c;
                if (gc >= a.M) continue;
                long orow;
                double val = (double)acc[i][t][v];
                if (cosrow) {
                    if (r >= N / 2) continue;
                    orow = r;
                    val = (x0[c] + ((r & 1) ? -xh[c] : xh[c])) + val;
                } else {
                    if (r >= H) continue;
                    orow = N / 2 + 1 + r;
                }
                if (a.mask_numr) {
                    const int J = gc % a.mask_numr + 1, K = gc / a.mask_numr + 1;
                    if (J >= a.mask_numr - 1 || (K <= 2 && !a.mask_keep_bottom) || K >= a.mask_numz - 1) val = 0.0;
                }
                ((T *)a.out)[orow * a.ldo + gc] = (T)val;
            }
    }
}

// ---------------------------------------------------------------------------
// inverse (FP64)
// ---------------------------------------------------------------------------
template <int CB>
__device__ __forceinline__ void dft_inv_tile(const DftArgs &a, int bx, unsigned char *lds)
{
    typedef double T;
    typedef Mma<double>::acc_t acc_t;
    typedef double v4 __attribute__((ext_vector_type(4)));
    constexpr int TN = CB / 16;
    const int N = a.N, H = N / 2 - 1, KG = (H + 15) / 16, KP = 16 * KG, KPp = KP + Pad<T>::v;
    const int NT = (N / 2 + 15) / 16;  // n tiles (n = 0..N/2-1) of P and of Q
    T *Ca = (T *)lds;                    // [CB][KPp] a_k, k = j + 1
    T *Sb = Ca + (size_t)CB * KPp;       // [CB][KPp] -b_k
    double *a0 = Sb + (size_t)CB * KPp;  // [CB]
    double *ah = a0 + CB;                // [CB]
    double *part = ah + CB;              // [16][CB]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int col0 = bx * CB;
    const double *X = a.x;
    const long ldx = a.ldx;
    const double invN = 1.0 / (double)N;

    constexpr int BATCH = 4;
    for (int e0 = 0; e0 < KP * CB; e0 += 512 * BATCH) {
        double p[BATCH], m[BATCH];
#pragma unroll
        for (int b = 0; b < BATCH; ++b) {
            const int e = e0 + 512 * b + tid, j = e / CB, gc = col0 + e % CB;
            const bool ok = e < KP * CB && j < H && gc < a.M;
            p[b] = ok ? X[(long)(j + 1) * ldx + gc] : 0.0;
            m[b] = ok ? X[(long)(N / 2 + 1 + j) * ldx + gc] : 0.0;
        }
#pragma unroll
        for (int b = 0; b < BATCH; ++b) {
            const int e = e0 + 512 * b + tid, j = e / CB, c = e % CB;
            if (e < KP * CB) {
                Ca[c * KPp + j] = p[b];
                Sb[c * KPp + j] = m[b];
            }
        }
    }
    if (tid < CB) {
        const int gc = col0 + tid;
        a0[tid] = gc < a.M ? X[gc] : 0.0;
        ah[tid] = gc < a.M ? X[(long)(N / 2) * ldx + gc] : 0.0;
    }
    __syncthreads();

    // x_{N/2} = P_{N/2} = a_0/N + (-1)^{N/2} a_{N/2}/N + sum_k 2 (-1)^k a_k / N
    {
        const int c = tid % CB, q = tid / CB;
        constexpr int NQ = 512 / CB;
        double s = 0.0;
        for (int j = q; j < H; j += NQ) {
            const double v = Ca[c * KPp + j];
            s += ((j + 1) & 1) ? -v : v;
        }
        if (q < 16) part[q * CB + c] = s;
        __syncthreads();
        if (NQ > 16 && q >= 16) part[(q & 15) * CB + c] += s;
        __syncthreads();
        if (tid < CB) {
            double t = 0.0;
            for (int i = 0; i < 16; ++i) t += part[i * CB + tid];
            const int gc = col0 + tid;
            if (gc < a.M)
                ((double *)a.out)[(long)(N / 2) * a.ldo + gc] =
                    a0[tid] * invN + (((N / 2) & 1) ? -ah[tid] : ah[tid]) * invN + 2.0 * invN * t;
        }
    }

    // wave w: n tiles w, w + 8 of P (A rows 0..NT-1) and of Q (A rows NT..2NT-1)
    const v4 *Af = (const v4 *)a.A;
    acc_t accP[2][TN], accQ[2][TN];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int t = 0; t < TN; ++t) {
            accP[i][t] = (acc_t){0.0, 0.0, 0.0, 0.0};
            accQ[i][t] = (acc_t){0.0, 0.0, 0.0, 0.0};
        }
    int ntile = 0;
#pragma unroll
    for (int i = 0; i < 2; ++i)
        if (wave + 8 * i < NT) ntile = i + 1;
    if (ntile > 0) {
        const v4 zero = {0.0, 0.0, 0.0, 0.0};
        const int jl = 4 * (lane >> 4);
        auto loadA = [&](int grp, v4 (&fp)[2], v4 (&fq)[2]) {
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                fp[i] = (i < ntile) ? Af[((long)(wave + 8 * i) * KG + grp) * 64 + lane] : zero;
                fq[i] = (i < ntile) ? Af[((long)(NT + wave + 8 * i) * KG + grp) * 64 + lane] : zero;
            }
        };
        auto step = [&](int grp, const v4 (&fp)[2], const v4 (&fq)[2]) {
            v4 bc[TN], bs[TN];
#pragma unroll
            for (int t = 0; t < TN; ++t) {
                const int off = (16 * t + (lane & 15)) * KPp + 16 * grp + jl;
                bc[t] = *(const v4 *)(Ca + off);
                bs[t] = *(const v4 *)(Sb + off);
            }
#pragma unroll
            for (int g = 0; g < 4; ++g)
#pragma unroll
                for (int i = 0; i < 2; ++i)
                    if (i < ntile) {
#pragma unroll
                        for (int t = 0; t < TN; ++t) {
                            accP[i][t] = Mma<double>::mma(fp[i][g], bc[t][g], accP[i][t]);
                            accQ[i][t] = Mma<double>::mma(fq[i][g], bs[t][g], accQ[i][t]);
                        }
                    }
        };
        v4 p0[2], q0[2], p1[2], q1[2];
        loadA(0, p0, q0);
        int grp = 0;
        for (; grp + 1 < KG; grp += 2) {
            loadA(grp + 1, p1, q1);
            step(grp, p0, q0);
            if (grp + 2 < KG) loadA(grp + 2, p0, q0);
            step(grp + 1, p1, q1);
        }
        if (grp < KG) step(grp, p0, q0);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        if (i >= ntile) continue;
        const int nt = wave + 8 * i;
#pragma unroll
        for (int t = 0; t < TN; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int n = 16 * nt + Mma<double>::row(lane, v);
                const int c = 16 * t + (lane & 15), gc = col0 + c;
                if (n >= N / 2 || gc >= a.M) continue;
                const double p = (a0[c] * invN + ((n & 1) ? -ah[c] : ah[c]) * invN) + accP[i][t][v];
                const double q = accQ[i][t][v];
                double *out = (double *)a.out;
                out[(long)n * a.ldo + gc] = p - q;
                if (n > 0) out[(long)(N - n) * a.ldo + gc] = p + q;
            }
    }
}

namespace {
constexpr int PN = DFT_MAX_N;  // the radix-4 kernels' transform length
}  // namespace

// ---------------------------------------------------------------------------
// Radix steps at N = 512.  Shared by the forward (radix 4) and the inverse
// (radix 8): panels of R4CB columns, and the persistent panel loop below.
// ---------------------------------------------------------------------------
namespace {
constexpr int R4N = PN / 4, R4KG = 4, R4CB = 16, R4VC = 4 * R4CB;

// Persistent loop over panels t0, t0 + st, ... < t1 (XCD-chunked as in
// BH_TILE_LOOP): the loads of panel t + st are issued into registers
// before the products of panel t run, so HBM and the matrix cores work
// together.  load(args, bx, regs) issues the loads; form(args, bx, regs)
// writes the panel's operands to LDS; work(args, bx) runs the products and
// writes the outputs.
template <class Regs, class Resolve, class Load, class Form, class Work>
__device__ __forceinline__ void panel_loop(int total, Resolve resolve, Load load, Form form, Work work)
{
    const int n_ = gridDim.x, b_ = blockIdx.x;
    int t0 = b_, t1 = total, st = n_;
    if (n_ >= 8 && (n_ & 7) == 0) {
        const int c_ = (total + 7) >> 3, x_ = b_ & 7;
        t0 = x_ * c_ + (b_ >> 3);
        t1 = min(total, (x_ + 1) * c_);
        st = n_ >> 3;
    }
    if (t0 >= t1) return;
    DftArgs a, an;
    int bx, nbx = 0;
    Regs R;
    resolve(t0, a, bx);
    load(a, bx, R);
    for (int t = t0; t < t1; t += st) {
        form(a, bx, R);
        __syncthreads();
        const bool more = t + st < t1;
        if (more) {
            resolve(t + st, an, nbx);
            load(an, nbx, R);
        }
        __builtin_amdgcn_sched_barrier(0);
        work(a, bx);
        __syncthreads();
        if (more) {
            a = an;
            bx = nbx;
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Inverse at N = 512 by one radix-8 step, decimating the output index:
//   x_{8m+r} = (1/N) sum_{k<N} X_k w^{k(8m+r)},   w = e^{2 pi i/N}
//            = IDFT_n(Y^(r))_m,   n = N/8,
//   Y^(r)_k  = (1/8) w^{kr} sum_{q<8} e^{2 pi i qr/8} X_{k+qn}   (X_{N-k} = conj X_k)
// with X_k = a_k + i s_k (s_k the mode rows N/2+1..N-1).  Each Y^(r) is
// Hermitian, so output residue r is the size-n real inverse of (Re, Im) of
// Y^(r)_0..n/2: the P/Q products above at size 64 on eight virtual columns
// per column, an eighth of the matrix work of the direct size-N form.  A
// workgroup takes a panel of R4CB columns; each thread forms the 8-point
// sum of one item (k = 1..n/2-1, or k = 0 together with k = n/2: 16 input
// rows either way; the X + xadd update and its store fused in) into LDS,
// then the products run (wave w: n-tile w % 2 of P and Q, residues
// 2 (w / 2) and 2 (w / 2) + 1) and the rows 8m + r are written.
// ---------------------------------------------------------------------------
namespace {
constexpr int R8N = PN / 8, R8KG = 2, R8KPp = 16 * R8KG + 2, R8VC = 8 * R4CB;
constexpr size_t R8_LDS = (size_t)2 * R8VC * R8KPp * 8 + 2 * R8VC * 8;
constexpr int R8_FRAGS = 4 * R8KG * 64 * 4;  // doubles of the size-n P and Q fragments

// Input modes of the inverse: plain x; x + xadd (the FP32 correction
// loaded beside x); a cold start, where an interior column's register pair
// holds the two FP32 values xint, xadd (bit-packed) and a face column's the
// FP64 x (xadd is zero on the faces); a warm start with two corrections,
// x + (xint + xadd), the pair bit-packed beside x.  Each mode is its own
// kernel, so that the registers held across a panel are only those the
// mode needs.
// INV_COLD_NS: INV_COLD without the store of x (a cold start whose
// warm-start state is left as x's faces and the two corrections).
enum { INV_PLAIN = 0, INV_ADD = 1, INV_COLD = 2, INV_ADD2 = 3, INV_COLD_NS = 4 };
template <int MODE>
struct Inv8Regs {
    double av[8], sv[8];
};
template <>
struct Inv8Regs<INV_ADD> {
    double av[8], sv[8];
    float fa[8], fs[8];
};

typedef float inv_f2 __attribute__((ext_vector_type(2)));
// a cold start's interior column (2 <= J <= numr - 1, 2 <= K <= numz - 1)
__device__ __forceinline__ bool inv8_inner(const DftArgs &a, int gc)
{
    int nr = a.mask_numr;
    asm volatile("" : "+s"(nr));  // the division's constants recomputed per use, not held across panels
    const int J = gc % nr + 1, K = gc / nr + 1;
    return J >= 2 && J <= nr - 1 && K >= 2 && K <= a.mask_numz - 1;
}

// input row of slot (q, s) of item kk (s = 0: av, 1: sv).  kk = 1..n/2-1:
// X_{kk+qn} is a_d + i s_d for d = kk + qn <= N/2 (q < 4) and the conjugate
// of X_{N-d} (q >= 4).  kk = 0: av = a_0, a_n, a_2n, a_3n, a_{n/2},
// a_{3n/2}, a_{5n/2}, a_{7n/2}; sv = a_{N/2}, s_n, s_2n, s_3n, s_{n/2},
// s_{3n/2}, s_{5n/2}, s_{7n/2}.
__device__ __forceinline__ int inv8_slot_row(int kk, int q, int s)
{
    if (kk == 0) {
        const int base = q < 4 ? q * R8N : (q - 4) * R8N + R8N / 2;
        if (s == 0) return base;
        return q == 0 ? PN / 2 : PN / 2 + base;
    }
    const int d = q < 4 ? kk + q * R8N : PN - (kk + q * R8N);
    return s == 0 ? d : PN / 2 + d;
}

template <int MODE>
__device__ __forceinline__ void inv8_load(const DftArgs &a, int bx, Inv8Regs<MODE> &R)
{
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));  // lane-dependent offsets recomputed per panel
    long ldx = a.ldx;
    asm volatile("" : "+s"(ldx));  // (not hoisted into registers held across panels)
    const int kk = tid / R4CB, c = tid - kk * R4CB, gc = min(bx * R4CB + c, a.M - 1);
    // one branch around all the loads of a kind (a branch per load makes
    // the compiler wait where the paths join)
    if ((MODE == INV_COLD || MODE == INV_COLD_NS) && inv8_inner(a, gc)) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const long o0 = (long)inv8_slot_row(kk, q, 0) * ldx + gc, o1 = (long)inv8_slot_row(kk, q, 1) * ldx + gc;
            R.av[q] = __builtin_bit_cast(double, (inv_f2){a.xint[o0], a.xadd[o0]});
            R.sv[q] = __builtin_bit_cast(double, (inv_f2){a.xint[o1], a.xadd[o1]});
        }
    } else {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            R.av[q] = a.x[(long)inv8_slot_row(kk, q, 0) * ldx + gc];
            R.sv[q] = a.x[(long)inv8_slot_row(kk, q, 1) * ldx + gc];
        }
    }
    if constexpr (MODE == INV_ADD) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            R.fa[q] = a.xadd[(long)inv8_slot_row(kk, q, 0) * ldx + gc];
            R.fs[q] = a.xadd[(long)inv8_slot_row(kk, q, 1) * ldx + gc];
        }
    }
}

// z_r = sum_q e^{2 pi i qr/8} x_q, r = 0..7 (even/odd split, two 4-point sums)
__device__ __forceinline__ void dft8_pos(const double *xr, const double *xi, double *zr, double *zi)
{
    double er[4], ei[4], orr[4], oi[4];
    auto four = [](double a0r, double a0i, double a1r, double a1i, double a2r, double a2i, double a3r, double a3i,
                   double *yr, double *yi) {
        const double pr = a0r + a2r, pi = a0i + a2i, mr = a0r - a2r, mi = a0i - a2i;
        const double ur = a1r + a3r, ui = a1i + a3i, dr = a1r - a3r, di = a1i - a3i;
        yr[0] = pr + ur; yi[0] = pi + ui;
        yr[1] = mr - di; yi[1] = mi + dr;
        yr[2] = pr - ur; yi[2] = pi - ui;
        yr[3] = mr + di; yi[3] = mi - dr;
    };
    four(xr[0], xi[0], xr[2], xi[2], xr[4], xi[4], xr[6], xi[6], er, ei);
    four(xr[1], xi[1], xr[3], xi[3], xr[5], xi[5], xr[7], xi[7], orr, oi);
    const double h = 0.70710678118654752440;  // cos(pi/4)
    const double wr[4] = {1.0, h, 0.0, -h}, wi[4] = {0.0, h, 1.0, h};
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const double tr = wr[r] * orr[r] - wi[r] * oi[r], ti = wr[r] * oi[r] + wi[r] * orr[r];
        zr[r] = er[r] + tr;
        zi[r] = ei[r] + ti;
        zr[r + 4] = er[r] - tr;
        zi[r + 4] = ei[r] - ti;
    }
}

// the eight spectra Y^(r)_k into LDS; with xadd, x + xadd is also stored
// (every input row belongs to exactly one item slot)
template <int MODE>
__device__ __forceinline__ void inv8_form(const DftArgs &a, int bx, const Inv8Regs<MODE> &R, unsigned char *lds)
{
    double *Ca = (double *)lds, *Sb = Ca + R8VC * R8KPp, *E0 = Sb + R8VC * R8KPp, *EH = E0 + R8VC;
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    long ldx = a.ldx;
    asm volatile("" : "+s"(ldx));
    const double *tw = (const double *)a.A + R8_FRAGS;  // [k][r] (cos, sin)(2 pi k r / N), k = 0..n/2
    const int kk = tid / R4CB, c = tid - kk * R4CB, gc = bx * R4CB + c;
    double av[8], sv[8];
    const bool inner = (MODE == INV_COLD || MODE == INV_COLD_NS) && inv8_inner(a, min(gc, a.M - 1));
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        av[q] = R.av[q];
        sv[q] = R.sv[q];
        if constexpr (MODE == INV_COLD || MODE == INV_COLD_NS) {
            if (inner) {
                const inv_f2 pa = __builtin_bit_cast(inv_f2, R.av[q]), ps = __builtin_bit_cast(inv_f2, R.sv[q]);
                av[q] = (double)pa.x + (double)pa.y;
                sv[q] = (double)ps.x + (double)ps.y;
            }
        }
    }
    // the warm start's two corrections are loaded here, not with the
    // panel's prefetch: held across the previous panel's products they do
    // not fit in the registers beside x
    float pa[8], pb[8], sa[8], sb[8];
    if constexpr (MODE == INV_ADD2) {
        const int gcl = min(gc, a.M - 1);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const long o0 = (long)inv8_slot_row(kk, q, 0) * ldx + gcl, o1 = (long)inv8_slot_row(kk, q, 1) * ldx + gcl;
            pa[q] = a.xint[o0];
            pb[q] = a.xadd[o0];
            sa[q] = a.xint[o1];
            sb[q] = a.xadd[o1];
        }
    }
    if constexpr (MODE != INV_PLAIN && MODE != INV_COLD_NS) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            if constexpr (MODE == INV_ADD) {
                av[q] += (double)R.fa[q];
                sv[q] += (double)R.fs[q];
            }
            if constexpr (MODE == INV_ADD2) {
                av[q] += (double)pa[q] + (double)pb[q];
                sv[q] += (double)sa[q] + (double)sb[q];
            }
            if (gc < a.M) {
                a.xstore[(long)inv8_slot_row(kk, q, 0) * ldx + gc] = av[q];
                a.xstore[(long)inv8_slot_row(kk, q, 1) * ldx + gc] = sv[q];
            }
        }
    }
    auto eighths = [&](int k, const double *xr, const double *xi, double *yr, double *yi) {
        double zr[8], zi[8];
        dft8_pos(xr, xi, zr, zi);
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const double wc = tw[(k * 8 + r) * 2], ws = tw[(k * 8 + r) * 2 + 1];
            yr[r] = 0.125 * (wc * zr[r] - ws * zi[r]);
            yi[r] = 0.125 * (wc * zi[r] + ws * zr[r]);
        }
    };
    double xr[8], xi[8], yr[8], yi[8];
    if (kk == 0) {
        // k = 0: X_0 = a_0, X_qn = a_qn + i s_qn (q = 1..3), X_{N/2} = a_{N/2}, then conjugates
        xr[0] = av[0]; xi[0] = 0.0;
        xr[4] = sv[0]; xi[4] = 0.0;
#pragma unroll
        for (int q = 1; q < 4; ++q) {
            xr[q] = av[q]; xi[q] = sv[q];
            xr[8 - q] = av[q]; xi[8 - q] = -sv[q];
        }
        eighths(0, xr, xi, yr, yi);
#pragma unroll
        for (int r = 0; r < 8; ++r) E0[r * R4CB + c] = yr[r];
        // k = n/2: X_{n/2 + qn} (q = 0..3) direct, q = 4..7 the conjugates of q = 3..0
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            xr[q] = av[4 + q]; xi[q] = sv[4 + q];
            xr[7 - q] = av[4 + q]; xi[7 - q] = -sv[4 + q];
        }
        eighths(R8N / 2, xr, xi, yr, yi);
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int vc = r * R4CB + c;
            EH[vc] = yr[r];
            Ca[vc * R8KPp + R8N / 2 - 1] = 0.0;  // K padding
            Sb[vc * R8KPp + R8N / 2 - 1] = 0.0;
        }
    } else {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            xr[q] = av[q];
            xi[q] = q < 4 ? sv[q] : -sv[q];
        }
        eighths(kk, xr, xi, yr, yi);
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int vc = r * R4CB + c;
            Ca[vc * R8KPp + kk - 1] = yr[r];
            Sb[vc * R8KPp + kk - 1] = yi[r];
        }
    }
}

// size-n products (wave w: P and Q rows m = 16 (w % 2).., virtual column
// tiles 2 (w / 2) and 2 (w / 2) + 1, i.e. residues r), the rows 8m + r and
// 8(n - m) + r, and the middle rows 8 (n/2) + r
__device__ __forceinline__ void inv8_work(const DftArgs &a, int bx, const unsigned char *lds)
{
    typedef Mma<double>::acc_t acc_t;
    typedef double v4 __attribute__((ext_vector_type(4)));
    const double *Ca = (const double *)lds, *Sb = Ca + R8VC * R8KPp, *E0 = Sb + R8VC * R8KPp, *EH = E0 + R8VC;
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int lane = tid & 63, wave = tid >> 6;
    const int col0 = bx * R4CB;
    const v4 *Af = (const v4 *)a.A;
    const int mt = wave & 1, t0 = 2 * (wave >> 1), jl = 4 * (lane >> 4);
    acc_t accP[2], accQ[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) accP[i] = accQ[i] = (acc_t){0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int grp = 0; grp < R8KG; ++grp) {
        const v4 fp = Af[((long)mt * R8KG + grp) * 64 + lane];
        const v4 fq = Af[((long)(2 + mt) * R8KG + grp) * 64 + lane];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int off = (16 * (t0 + i) + (lane & 15)) * R8KPp + 16 * grp + jl;
            const v4 bc = *(const v4 *)(Ca + off), bs = *(const v4 *)(Sb + off);
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                accP[i] = Mma<double>::mma(fp[g], bc[g], accP[i]);
                accQ[i] = Mma<double>::mma(fq[g], bs[g], accQ[i]);
            }
        }
    }
    const double invn = 1.0 / (double)R8N;
    double *out = (double *)a.out;
    // with the axis ghost: a J = 2 value also goes to J = 1 of the plane
    // rotated by pi; the J = 1 column's own transform is not written
    const int anr = a.axis_numr;
    auto put = [&](int row, int gc, int jm, double v) {
        if (jm != 0) out[(long)row * a.ldo + gc] = v;
        if (jm == 1) out[(long)((row + PN / 2) & (PN - 1)) * a.ldo + gc - 1] = v;
    };
    {
        const int c = lane & 15, gc = col0 + c, jm = anr ? gc % anr : -1;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int r = t0 + i, vc = r * R4CB + c;
            const double e0 = E0[vc] * invn, eh = EH[vc] * invn;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = 16 * mt + Mma<double>::row(lane, v);
                if (gc >= a.M) continue;
                const double p = (e0 + ((m & 1) ? -eh : eh)) + accP[i][v];
                const double q = accQ[i][v];
                put(8 * m + r, gc, jm, p - q);
                if (m > 0) put(8 * (R8N - m) + r, gc, jm, p + q);
            }
        }
    }
    // row 8 (n/2) + r: e0 + eh + sum_k (2/n) (-1)^k a'_k, four lanes per column
    {
        const int vc = tid >> 2, j = tid & 3;
        double sm = 0.0;
        for (int k = 1 + j; k < R8N / 2; k += 4) sm += (k & 1) ? -Ca[vc * R8KPp + k - 1] : Ca[vc * R8KPp + k - 1];
        sm += __shfl_xor(sm, 1, 64);
        sm += __shfl_xor(sm, 2, 64);
        const int r = vc / R4CB, c = vc - r * R4CB, gc = col0 + c;
        if (j == 0 && gc < a.M)
            put(8 * (R8N / 2) + r, gc, anr ? gc % anr : -1, (E0[vc] + EH[vc]) * invn + 2.0 * invn * sm);
    }
}

template <int MODE, class Resolve>
__device__ __forceinline__ void inv8_loop(int total, Resolve resolve, unsigned char *lds)
{
    panel_loop<Inv8Regs<MODE>>(
        total, resolve, [&](const DftArgs &a, int bx, Inv8Regs<MODE> &R) { inv8_load<MODE>(a, bx, R); },
        [&](const DftArgs &a, int bx, const Inv8Regs<MODE> &R) { inv8_form<MODE>(a, bx, R, lds); },
        [&](const DftArgs &a, int bx) { inv8_work(a, bx, lds); });
}

template <int MODE>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void dft_inv8_kernel(DftArgs a, int nx,
                                                                                             int ny, int total)
{
    if constexpr (MODE == INV_COLD_NS)
        if (a.set_flag && blockIdx.x == 0 && threadIdx.x == 0) *a.set_flag = 1.f;
    extern __shared__ __align__(16) unsigned char inv8_lds[];
    inv8_loop<MODE>(total, [&](int t, DftArgs &o, int &bx) { o = a; bx = t; }, inv8_lds);
}

template <int MODE>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void dft_inv8_grouped(
    const DftArgs *__restrict__ args, GroupSched gs)
{
    extern __shared__ __align__(16) unsigned char inv8_lds[];
    inv8_loop<MODE>(
        gs.total,
        [&](int t, DftArgs &o, int &bx) {
            const int g = bh_find_group(gs.grp, gs.ngroups, t);
            o = args[g];
            bx = t - gs.grp[g].tile0;
        },
        inv8_lds);
}
}  // namespace

// ---------------------------------------------------------------------------
// Forward at N = 512 by one radix-4 step, decimating the input index:
//   X_k = sum_{r<4} w^{-kr} F^(r)_{k mod n},  F^(r) = DFT_n(x_{4j+r}),  n = N/4,
// so with G^(r)_k = w^{-kr} F^(r)_k (k <= n/2, F^(r)_{n-k} = conj F^(r)_k),
//   X_k = sum_r G^(r)_k,            X_{n+k}  = sum_r (-i)^r G^(r)_k,
//   X_{n-k} = sum_r G'^(r)_k,       X_{2n-k} = sum_r (-i)^r G'^(r)_k,
//   G'^(r)_k = w^{-(n-k) r} conj F^(r)_k.
// The four length-n real transforms F^(r) run as the folded cosine/sine
// products of the direct form at size n on four virtual columns per
// column (a quarter of its matrix work); a workgroup takes R4Fwd::CB columns,
// folds the quarter sequences into LDS, runs the products (wave w: row
// tile w % 4 of the cosine and of the sine half, the half w / 4 of the
// 16-wide virtual-column tiles), puts F in LDS and combines the rows k, n + k, n - k,
// 2n - k of X in FP64, masked as the direct form is.
// ---------------------------------------------------------------------------
namespace {
template <typename T>
struct R4Fwd {
    // panel columns: an FP32 output row of a panel is then 128 bytes, a
    // whole cache line (64-byte rows, each line written half by one panel
    // and half by the next, ran the stores about 30 % slower)
    static constexpr int CB = sizeof(T) == 4 ? 2 * R4CB : R4CB, VC = 4 * CB;
    static constexpr int KPp = 16 * R4KG + Pad<T>::v;
    static constexpr size_t uv = (size_t)2 * VC * KPp * sizeof(T);          // U, V (then Fa, Fs)
    static constexpr size_t f = (size_t)2 * (R4N / 2 + 1) * VC * sizeof(T);  // Fa, Fs
    static constexpr size_t tw = (size_t)(R4N + 1) * 8 * sizeof(T);  // twiddles (cos, sin)(2 pi k r / N)
    static constexpr size_t bytes = (uv > f ? uv : f) + 3 * VC * sizeof(double) + tw;
    static constexpr int frags = 8 * R4KG * 64 * 4;  // size-n cosine and sine fragments
    static constexpr int JG = 512 / CB;              // thread groups along j in the fold
    static constexpr int NJ = (R4N / 2) / JG;        // fold rows j = jg + JG i per thread
    static_assert(512 % CB == 0 && JG * NJ == R4N / 2, "the fold's thread map covers rows 0..n/2-1");
};

template <typename T>
__device__ __forceinline__ void dft_fwd4_tile(const DftArgs &a, int bx, unsigned char *lds)
{
    typedef typename Mma<T>::acc_t acc_t;
    typedef T v4 __attribute__((ext_vector_type(4)));
    typedef R4Fwd<T> Ly;
    constexpr int KPp = Ly::KPp, CB = Ly::CB, VC = Ly::VC, CT = VC / 16;  // CT: 16-wide column tiles
    T *U = (T *)lds, *V = U + VC * KPp;
    T *Fa = (T *)lds, *Fs = Fa + (R4N / 2 + 1) * VC;  // after the products, over U and V
    double *X0 = (double *)(lds + (Ly::uv > Ly::f ? Ly::uv : Ly::f)), *XH = X0 + VC, *FM = XH + VC;
    T *twl = (T *)(FM + VC);  // [k][r] (cos, sin)(2 pi k r / N), k = 0..n
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));  // lane-dependent offsets recomputed per tile, not held across tiles
    const int lane = tid & 63, wave = tid >> 6;
    const int col0 = bx * CB;
    const double *x = a.x;
    long ldx = a.ldx;
    asm volatile("" : "+s"(ldx));
    // 1. fold the quarter sequences x^(r)_j = x_{4j+r}: u_j, v_j (j = 1..n/2-1), x^(r)_0, x^(r)_{n/2}.
    //    Thread (jg, c) takes rows j = jg + JG i of column c, and group 0 also
    //    j = n/2; every load of the panel (and of the twiddles, from L2) is
    //    issued before the first is used, so the panel costs one memory
    //    latency.  Row j = 0 has no mirror: its second load repeats the first
    //    row (a select on the address, not a branch around the load).
    const T *twg = (const T *)a.A + Ly::frags;
    constexpr int TWN = (R4N + 1) * 8;
    T twv[(TWN + 511) / 512];
#pragma unroll
    for (int i = 0; i < (TWN + 511) / 512; ++i) twv[i] = tid + 512 * i < TWN ? twg[tid + 512 * i] : (T)0;
    {
        constexpr int JG = Ly::JG, NJ = Ly::NJ;
        const int c = tid % CB, jg = tid / CB, gc = col0 + c;
        const int gcc = gc < a.M ? gc : a.M - 1;
        double p[NJ][4], q[NJ][4], ph[4];
#pragma unroll
        for (int i = 0; i < NJ; ++i) {
            const int j = jg + JG * i;
            const int jm = j > 0 ? PN - 4 * j : 0;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                p[i][r] = x[(long)(4 * j + r) * ldx + gcc];
                q[i][r] = x[(long)(jm + r) * ldx + gcc];
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) ph[r] = x[(long)(2 * R4N + r) * ldx + gcc];
#pragma unroll
        for (int i = 0; i < NJ; ++i) {
            const int j = jg + JG * i;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int vc = r * CB + c;
                if (j == 0) {
                    X0[vc] = p[i][r];
                } else {
                    U[vc * KPp + j - 1] = (T)(p[i][r] + q[i][r]);
                    V[vc * KPp + j - 1] = (T)(p[i][r] - q[i][r]);
                }
            }
        }
        if (jg == 0) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int vc = r * CB + c;
                XH[vc] = ph[r];
                U[vc * KPp + R4N / 2 - 1] = (T)0;  // K padding
                V[vc * KPp + R4N / 2 - 1] = (T)0;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < (TWN + 511) / 512; ++i)
        if (tid + 512 * i < TWN) twl[tid + 512 * i] = twv[i];
    __syncthreads();
    // middle row of each quarter: F_{n/2} = x_0 + x_{n/2} + sum_j (-1)^j u_j (FP64 sum)
#pragma unroll
    for (int vc0 = 0; vc0 < VC; vc0 += 64) {
        const int vc = vc0 + (tid >> 3), jj = tid & 7;
        double sm = 0.0;
        for (int j = 1 + jj; j < R4N / 2; j += 8) {
            const double u = (double)U[vc * KPp + j - 1];
            sm += (j & 1) ? -u : u;
        }
        sm += __shfl_xor(sm, 1, 64);
        sm += __shfl_xor(sm, 2, 64);
        sm += __shfl_xor(sm, 4, 64);
        if (jj == 0) FM[vc] = (X0[vc] + XH[vc]) + sm;
    }
    // 2. size-n products: cosine rows k = 16 mt.., sine rows k = 16 mt + 1..
    const v4 *Af = (const v4 *)a.A;
    // wave w: row tile w % 4, column tiles TW (w / 4) .. + TW - 1
    constexpr int TW = CT / 2;
    const int mt = wave & 3, t0 = TW * (wave >> 2), jl = 4 * (lane >> 4);
    acc_t accC[TW], accS[TW];
#pragma unroll
    for (int i = 0; i < TW; ++i) accC[i] = accS[i] = (acc_t){(T)0, (T)0, (T)0, (T)0};
#pragma unroll 1
    for (int grp = 0; grp < R4KG; ++grp) {
        const v4 fc = Af[((long)mt * R4KG + grp) * 64 + lane];
        const v4 fs = Af[((long)(4 + mt) * R4KG + grp) * 64 + lane];
#pragma unroll
        for (int i = 0; i < TW; ++i) {
            const int off = (16 * (t0 + i) + (lane & 15)) * KPp + 16 * grp + jl;
            const v4 bu = *(const v4 *)(U + off), bv = *(const v4 *)(V + off);
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                accC[i] = Mma<T>::mma(fc[g], bu[g], accC[i]);
                accS[i] = Mma<T>::mma(fs[g], bv[g], accS[i]);
            }
        }
    }
    __syncthreads();  // U, V read: F goes over them
#pragma unroll
    for (int i = 0; i < TW; ++i) {
        const int vc = 16 * (t0 + i) + (lane & 15);
        const double e0 = X0[vc], eh = XH[vc];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int k = 16 * mt + Mma<T>::row(lane, v);  // cosine row k, sine row k + 1
            Fa[k * VC + vc] = (T)((e0 + ((k & 1) ? -eh : eh)) + (double)accC[i][v]);
            if (k + 1 < R4N / 2) Fs[(k + 1) * VC + vc] = accS[i][v];
        }
    }
    if (tid < VC) {
        Fa[(R4N / 2) * VC + tid] = (T)FM[tid];
        Fs[tid] = (T)0;                      // F_0, F_{n/2} are real
        Fs[(R4N / 2) * VC + tid] = (T)0;
    }
    __syncthreads();
    // 3. combine: rows k, n + k (always), 2n - k (k < n/2), n - k (0 < k < n/2)
    const T *tw = twl;
#pragma unroll 1
    for (int e = tid; e < (R4N / 2 + 1) * CB; e += 512) {
        const int k = e / CB, c = e - k * CB, gc = col0 + c;
        if (gc >= a.M) continue;
        bool masked = false;
        long ri = -1;  // offset of (J, K) in a plane of rout, if interior
        if (a.mask_numr) {
            const int J = gc % a.mask_numr + 1, K = gc / a.mask_numr + 1;
            masked = J >= a.mask_numr - 1 || (K <= 2 && !a.mask_keep_bottom) || K >= a.mask_numz - 1;
            if (a.rout && J >= 2 && J <= a.mask_numr - 1 && K >= 2 && K <= a.mask_numz - 1)
                ri = (long)(K - 2) * (a.mask_numr - 2) + (J - 2);
        }
        const long rplane = (long)(a.mask_numr - 2) * (a.mask_numz - 2);
        double gr[4], gi[4], hr[4], hi[4];  // G^(r)_k and G'^(r)_k
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int vc = r * CB + c;
            const double fr = (double)Fa[k * VC + vc], fi = (double)Fs[k * VC + vc];
            const double c1 = (double)tw[(k * 4 + r) * 2], s1 = (double)tw[(k * 4 + r) * 2 + 1];
            const double c2 = (double)tw[((R4N - k) * 4 + r) * 2], s2 = (double)tw[((R4N - k) * 4 + r) * 2 + 1];
            // w^{-kr} F = (c1 - i s1)(fr + i fi); w^{-(n-k) r} conj F = (c2 - i s2)(fr - i fi)
            gr[r] = c1 * fr + s1 * fi;
            gi[r] = c1 * fi - s1 * fr;
            hr[r] = c2 * fr - s2 * fi;
            hi[r] = -c2 * fi - s2 * fr;
        }
        // sum_r z_r and sum_r (-i)^r z_r = z0 - i z1 - z2 + i z3
        const double x0r = (gr[0] + gr[2]) + (gr[1] + gr[3]), x0i = (gi[0] + gi[2]) + (gi[1] + gi[3]);
        const double x1r = (gr[0] - gr[2]) + (gi[1] - gi[3]), x1i = (gi[0] - gi[2]) - (gr[1] - gr[3]);
        const double y0r = (hr[0] + hr[2]) + (hr[1] + hr[3]), y0i = (hi[0] + hi[2]) + (hi[1] + hi[3]);
        const double y1r = (hr[0] - hr[2]) + (hi[1] - hi[3]), y1i = (hi[0] - hi[2]) - (hr[1] - hr[3]);
        T *out = (T *)a.out;
        const long ldo = a.ldo;
        auto put = [&](int row, double re, double im) {  // a_row = Re, s-row N/2 + row = Im
            const T vr = (T)(masked ? 0.0 : re), vi = (T)(masked ? 0.0 : im);
            out[(long)row * ldo + gc] = vr;
            if (ri >= 0) a.rout[(long)row * rplane + ri] = (float)(a.rscale * (double)vr);
            if (row > 0 && row < PN / 2) {
                out[(long)(PN / 2 + row) * ldo + gc] = vi;
                if (ri >= 0) a.rout[(long)(PN / 2 + row) * rplane + ri] = (float)(a.rscale * (double)vi);
            }
        };
        put(k, x0r, x0i);
        put(R4N + k, x1r, x1i);
        if (k < R4N / 2) put(2 * R4N - k, y1r, y1i);
        if (k > 0 && k < R4N / 2) put(R4N - k, y0r, y0i);
    }
}

template <typename T>
__global__ __launch_bounds__(512) void dft_fwd4_kernel(DftArgs a, int nx, int ny, int total)
{
    extern __shared__ __align__(16) unsigned char fwd4_lds[];
    BH_SINGLE_LOOP(a, nx, ny, total, (dft_fwd4_tile<T>(a, bx, fwd4_lds)));
}

template <typename T>
__global__ __launch_bounds__(512) void dft_fwd4_grouped(const DftArgs *__restrict__ args, GroupSched gs)
{
    extern __shared__ __align__(16) unsigned char fwd4_lds[];
    BH_GROUP_LOOP(DftArgs, args, gs, (dft_fwd4_tile<T>(ga, bx, fwd4_lds)));
}
}  // namespace

// ---------------------------------------------------------------------------
// entry points.  variant = kind + 4 log2(N); kind 0: forward FP32 (32
// columns per panel), 1: forward FP64 (16 columns), 2: inverse FP64 (16)
// ---------------------------------------------------------------------------
static int dft_cols(int kind) { return kind == DFT_FWD_F32 ? 32 : 16; }

size_t dft_lds_bytes(int kind, int N)
{
    const int H = N / 2 - 1, KP = 16 * ((H + 15) / 16), CB = dft_cols(kind);
    const size_t es = kind == DFT_FWD_F32 ? 4 : 8;
    const int KPp = KP + (kind == DFT_FWD_F32 ? 4 : 2);
    return 2 * (size_t)KPp * CB * es + (2 + 16) * (size_t)CB * 8;
}

static int log2i(int n)
{
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}

// N < 512: V = the kind (0 forward FP32, 1 forward FP64, 2 inverse FP64)
template <int V>
__device__ __forceinline__ void dft_tile_v(const DftArgs &a, int bx, unsigned char *lds)
{
    if (V == 0) dft_fwd_tile<float, 32>(a, bx, lds);
    else if (V == 1) dft_fwd_tile<double, 16>(a, bx, lds);
    else dft_inv_tile<16>(a, bx, lds);
}

template <int V>
__global__ __launch_bounds__(512, 4) void dft_kernel(DftArgs a, int nx, int ny, int total)
{
    extern __shared__ __align__(16) unsigned char dft_lds[];
    BH_SINGLE_LOOP(a, nx, ny, total, (dft_tile_v<V>(a, bx, dft_lds)));
}

template <int V>
__global__ __launch_bounds__(512, 4) void dft_grouped(const DftArgs *__restrict__ args, GroupSched gs)
{
    extern __shared__ __align__(16) unsigned char dft_lds[];
    BH_GROUP_LOOP(DftArgs, args, gs, (dft_tile_v<V>(ga, bx, dft_lds)));
}

Op dft_op(const DftArgs &a, int kind)
{
    // N = 512: the radix kernels' panels (the FP32 forward's are twice as wide)
    const int CB = a.N == DFT_MAX_N ? (kind == DFT_FWD_F32 ? R4Fwd<float>::CB : R4CB) : dft_cols(kind);
    // variant: kind + 4 log2 N + 64 (input mode of the inverse at N = 512)
    const int mode = kind == DFT_INV_F64 && a.N == DFT_MAX_N
                         ? (a.xint ? (a.warm ? INV_ADD2 : a.xstore ? INV_COLD : INV_COLD_NS)
                                   : a.xadd ? INV_ADD : INV_PLAIN)
                         : 0;
    return Op::make(OP_DFT, kind + 4 * log2i(a.N) + 64 * mode, a, (a.M + CB - 1) / CB, 1, 1);
}

namespace {
const void *dft_fn(int variant, bool grouped)
{
    switch (variant % 4) {
    case DFT_FWD_F32: return grouped ? (const void *)&dft_grouped<0> : (const void *)&dft_kernel<0>;
    case DFT_FWD_F64: return grouped ? (const void *)&dft_grouped<1> : (const void *)&dft_kernel<1>;
    default: return grouped ? (const void *)&dft_grouped<2> : (const void *)&dft_kernel<2>;
    }
}
}  // namespace

// the radix-4 kernels of N = 512: function and dynamic LDS per variant
static const void *r4_fn(int variant, bool grouped, int *lds)
{
    switch (variant % 4) {
    case DFT_INV_F64:
        *lds = (int)R8_LDS;
        switch (variant / 64) {
        case INV_PLAIN: return grouped ? (const void *)&dft_inv8_grouped<0> : (const void *)&dft_inv8_kernel<0>;
        case INV_ADD: return grouped ? (const void *)&dft_inv8_grouped<1> : (const void *)&dft_inv8_kernel<1>;
        case INV_COLD: return grouped ? (const void *)&dft_inv8_grouped<2> : (const void *)&dft_inv8_kernel<2>;
        case INV_COLD_NS: return grouped ? (const void *)&dft_inv8_grouped<4> : (const void *)&dft_inv8_kernel<4>;
        default: return grouped ? (const void *)&dft_inv8_grouped<3> : (const void *)&dft_inv8_kernel<3>;
        }
    case DFT_FWD_F32:
        *lds = (int)R4Fwd<float>::bytes;
        return grouped ? (const void *)&dft_fwd4_grouped<float> : (const void *)&dft_fwd4_kernel<float>;
    default:
        *lds = (int)R4Fwd<double>::bytes;
        return grouped ? (const void *)&dft_fwd4_grouped<double> : (const void *)&dft_fwd4_kernel<double>;
    }
}

bool dft_launch(const Op &op, hipStream_t s)
{
    if (op.kind != OP_DFT) return false;
    int nx = op.nx, ny = op.ny, total = (int)op.tiles();
    void *argv[] = {(void *)op.args, (void *)&nx, (void *)&ny, (void *)&total};
    int nb = persistent_blocks(total, 0);
    if ((1 << (op.variant / 4 % 16)) == DFT_MAX_N) {
        int lds;
        const void *fn = r4_fn(op.variant, false, &lds);
        const int cap = op.as<DftArgs>().max_blocks;  // side stream: leave room for the main stream
        if (cap > 0 && nb > cap) nb = cap >= 8 ? cap / 8 * 8 : cap;
        (void)hipLaunchKernel(fn, dim3(nb), dim3(512), argv, lds, s);
        return true;
    }
    const int lds = (int)dft_lds_bytes(op.variant % 4, 1 << (op.variant / 4 % 16));
    (void)hipLaunchKernel(dft_fn(op.variant, false), dim3(nb), dim3(512), argv, lds, s);
    return true;
}

bool dft_launch_grouped(int kind, int variant, const void *d_args, const GroupSched &gs, int nblocks, hipStream_t s)
{
    if (kind != OP_DFT) return false;
    GroupSched g = gs;
    const void *ap = d_args;
    void *argv[] = {(void *)&ap, (void *)&g};
    int lds;
    const void *fn;
    if ((1 << (variant / 4 % 16)) == DFT_MAX_N) {
        fn = r4_fn(variant, true, &lds);
    } else {
        lds = (int)dft_lds_bytes(variant % 4, 1 << (variant / 4 % 16));
        fn = dft_fn(variant, true);
    }
    (void)hipLaunchKernel(fn, dim3(nblocks), dim3(512), argv, lds, s);
    return true;
}

bool dft_occupancy(int kind, int variant, int *blocks)
{
    if (kind != OP_DFT) return false;
    int lds;
    const void *fn;
    if ((1 << (variant / 4 % 16)) == DFT_MAX_N) {
        fn = r4_fn(variant, true, &lds);
    } else {
        lds = (int)dft_lds_bytes(variant % 4, 1 << (variant / 4 % 16));
        fn = dft_fn(variant, true);
    }
    *blocks = kernel_occupancy(fn, 512, lds);
    return true;
}

// MFMA fragment order of the matrix operands (host): [row tile][group of
// 16 k][lane][4], element g of lane l holding row 16 rt + l % 16 and
// k = 16 group + 4 (l / 16) + g (the order the kernels read B in).
template <typename T>
static void frag_fill(std::vector<T> *out, int RT, int KG, const std::function<double(int, int)> &A)
{
    out->assign((size_t)RT * KG * 64 * 4, (T)0);
    for (int rt = 0; rt < RT; ++rt)
        for (int grp = 0; grp < KG; ++grp)
            for (int l = 0; l < 64; ++l)
                for (int g = 0; g < 4; ++g)
                    (*out)[(((size_t)rt * KG + grp) * 64 + l) * 4 + g] =
                        (T)A(rt * 16 + (l & 15), 16 * grp + 4 * (l >> 4) + g);
}

void dft_forward_fragments(int N, std::vector<float> *f32, std::vector<double> *f64)
{
    if (N == DFT_MAX_N) {
        // radix-4 form: the size-N/4 operand, then the twiddles (cos, sin)(2 pi k r / N)
        // for k = 0..N/4, r = 0..3, in the operand's precision
        dft_forward_fragments(N / 4, f32, f64);
        const double pi = std::acos(-1.0);
        for (int k = 0; k <= N / 4; ++k)
            for (int r = 0; r < 4; ++r) {
                const double ang = 2.0 * pi * (double)((long)k * r % N) / (double)N;
                if (f32) {
                    f32->push_back((float)std::cos(ang));
                    f32->push_back((float)std::sin(ang));
                }
                if (f64) {
                    f64->push_back(std::cos(ang));
                    f64->push_back(std::sin(ang));
                }
            }
        return;
    }
    const int H = N / 2 - 1, KS = (H + 15) / 16;
    const int RTc = (N / 2 + 15) / 16, RTs = (H + 15) / 16;
    const double pi = std::acos(-1.0);
    auto A = [&](int row, int j) -> double {
        if (j >= H) return 0.0;
        const long n = j + 1;
        if (row < 16 * RTc) {  // cosine rows k = row (k < N/2)
            const long k = row;
            if (k >= N / 2) return 0.0;
            return std::cos(2.0 * pi * (double)((k * n) % N) / (double)N);
        }
        const long k = row - 16 * RTc + 1;  // sine rows k = 1..H
        if (k > H) return 0.0;
        return -std::sin(2.0 * pi * (double)((k * n) % N) / (double)N);
    };
    if (f32) frag_fill<float>(f32, RTc + RTs, KS, A);
    if (f64) frag_fill<double>(f64, RTc + RTs, KS, A);
}

void dft_inverse_fragments(int N, std::vector<double> *f64)
{
    if (N == DFT_MAX_N) {
        // radix-8 form: the size-N/8 operand, then the twiddles w^{kr}
        // (cos, sin) for k = 0..N/16, r = 0..7
        dft_inverse_fragments(N / 8, f64);
        const double pi = std::acos(-1.0);
        for (int k = 0; k <= N / 16; ++k)
            for (int r = 0; r < 8; ++r) {
                const double ang = 2.0 * pi * (double)((long)k * r % N) / (double)N;
                f64->push_back(std::cos(ang));
                f64->push_back(std::sin(ang));
            }
        return;
    }
    const int H = N / 2 - 1, KS = (H + 15) / 16;
    const int NT = (N / 2 + 15) / 16;
    const double pi = std::acos(-1.0);
    auto A = [&](int row, int j) -> double {
        if (j >= H) return 0.0;
        const long k = j + 1;
        const bool q = row >= 16 * NT;
        const long n = q ? row - 16 * NT : row;
        if (n >= N / 2) return 0.0;
        const double ang = 2.0 * pi * (double)((k * n) % N) / (double)N;
        return 2.0 / (double)N * (q ? std::sin(ang) : std::cos(ang));
    };
    frag_fill<double>(f64, 2 * NT, KS, A);
}

}  // namespace bh
