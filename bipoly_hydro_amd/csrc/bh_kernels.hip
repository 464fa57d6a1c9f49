// bh_kernels.hip -- CDNA4 (gfx950) device code of the self-gravity solve, FP64.
//
//   line_sweep    one ADI half-step: RHS stencil + Thomas solve with one
//                 wavefront per tridiagonal line (helmadi.F90:201-343,
//                 tridagr.F90, tridagz.F90), written transposed so the next
//                 half-step reads its lines contiguously (the on-device form
//                 of swap_r_z / swap_z_r)
//   gen_*_table   Green's-function tables (tm.F90 / sm.F90 / bm.F90)
//   proj          azimuthal mode projections of rho (bessel.F90:176-235)
//   contract_*    boundary contraction (bessel.F90:186-262)
//   face_modes    boundary faces in Fourier mode order, placed in the potential (bessel.F90:396-488)
#include "bh_internal.h"
#include "bh_ops.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

namespace bh {

// ---------------------------------------------------------------------------
__global__ void setup_fill_kernel(double *p, size_t n, double v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

void launch_fill(double *p, size_t n, double v, hipStream_t s)
{
    if (!n) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(setup_fill_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
}

// Tiled fill for solve sequences: tile bx covers FILL_TILE elements.
constexpr int FILL_TILE = 2048;
__device__ __forceinline__ void fill_tile(const FillArgs &a, int bx, int, int)
{
    const long i0 = (long)bx * FILL_TILE;
#pragma unroll
    for (int u = 0; u < FILL_TILE / 256; ++u) {
        const long i = i0 + u * 256 + threadIdx.x;
        if (i < a.n) a.p[i] = a.v;
    }
}

// [plane][K][J] -> [plane][J][K]
// J layout -> the compact K-layout interior of LineArgs (rows J' = J - 2 of
// compact_pitch(numz - 2) doubles, element K' = K - 2): 32 x 32 tiles of
// interior points through LDS, each output row segment 256 aligned bytes
__device__ __forceinline__ void transpose_tile(const TransposeArgs &a, int bx, int by, int bz)
{
    __shared__ double tile[32][33];
    const int numr = a.numr, numz = a.numz, nri = numr - 2, nzi = numz - 2;
    const long pk = compact_pitch(nzi);
    const double *s = a.src + (long)bz * numr * numz;
    double *d = a.dst + (long)bz * tiled_stride(nri, nzi);
    int j0 = bx * 32, k0 = by * 32;
    int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int r = ty; r < 32; r += 8) {
        int k = k0 + r, j = j0 + tx;
        if (k < nzi && j < nri) tile[r][tx] = s[(long)(k + 1) * numr + j + 1];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        int j = j0 + r, k = k0 + tx;
        if (k < nzi && j < nri) d[(long)j * pk + k] = tile[tx][r];
    }
}

Op transpose_op(const double *src, double *dst, int numr, int numz, int nplanes)
{
    TransposeArgs a{src, dst, numr, numz};
    return Op::make(OP_TRANSPOSE, 0, a, (numr - 2 + 31) / 32, (numz - 2 + 31) / 32, nplanes);
}

Op fill_op(double *p, size_t n, size_t capacity, double v)
{
    FillArgs a{p, (long)n, v};
    return Op::make(OP_FILL, 0, a, (int)((capacity + FILL_TILE - 1) / FILL_TILE), 1, 1);
}

// ---------------------------------------------------------------------------
// ADI half-step, one wavefront (or half of one) per tridiagonal line.
// The Thomas forward sweep y_e = alpha_e y_{e-1} + ibet_e rhs_e and back
// substitution x_e = y_e + gback_e x_{e+1} are first-order linear
// recurrences: each lane composes the affine maps of its own consecutive
// elements, an inclusive scan of those maps over the line's lanes (DPP row
// shifts inside rows of 16 lanes, row broadcasts upwards, readlane
// downwards) joins the lanes, and each lane then finishes its elements from
// the value entering it (DPP wave shifts).  A wave owns consecutive lines
// and rolls the neighbour rows in registers, loading the next line's rows
// while it solves.
// ---------------------------------------------------------------------------
template <int CTRL>
__device__ inline double dpp_d(double old, double v)
{
    long long o = __double_as_longlong(old), x = __double_as_longlong(v);
    int lo = __builtin_amdgcn_update_dpp((int)o, (int)x, CTRL, 0xf, 0xf, false);
    int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(x >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ inline double readlane_d(double v, int l)
{
    long long x = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)x, l);
    int hi = __builtin_amdgcn_readlane((int)(x >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// compose (A,B) <- (A,B) o (Ap,Bp):  y = A (Ap y0 + Bp) + B
__device__ inline void compose(double &A, double &B, double Ap, double Bp)
{
    B = A * Bp + B;
    A = A * Ap;
}

// dpp_d with a row mask: rows outside ROWS keep `old`
template <int CTRL, int ROWS>
__device__ inline double dppm_d(double old, double v)
{
    long long o = __double_as_longlong(old), x = __double_as_longlong(v);
    int lo = __builtin_amdgcn_update_dpp((int)o, (int)x, CTRL, ROWS, 0xf, false);
    int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(x >> 32), CTRL, ROWS, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Scans over the lanes of a line: the whole wave (H = 1) or each half of 32
// lanes (H = 2, two lines side by side).
// Inclusive scan upwards: row shifts inside the rows of 16 lanes, then lane
// 15 broadcast to row 1 and lane 47 to row 3 (row_bcast:15), and for H = 1
// lane 31 to rows 2 and 3 (row_bcast:31); rows left out compose with the
// identity.
template <int H>
__device__ inline void scan_up(double &A, double &B)
{
    compose(A, B, dpp_d<0x111>(1.0, A), dpp_d<0x111>(0.0, B));  // row_shr:1
    compose(A, B, dpp_d<0x112>(1.0, A), dpp_d<0x112>(0.0, B));  // row_shr:2
    compose(A, B, dpp_d<0x114>(1.0, A), dpp_d<0x114>(0.0, B));  // row_shr:4
    compose(A, B, dpp_d<0x118>(1.0, A), dpp_d<0x118>(0.0, B));  // row_shr:8
    compose(A, B, dppm_d<0x142, 0xa>(1.0, A), dppm_d<0x142, 0xa>(0.0, B));  // row_bcast:15
    if (H == 1) compose(A, B, dppm_d<0x143, 0xc>(1.0, A), dppm_d<0x143, 0xc>(0.0, B));  // row_bcast:31
}

// inclusive scan downwards: row shifts, then the first lane of row 1 (of
// row 3) into row 0 (row 2), and for H = 1 lane 32 into rows 0 and 1
template <int H>
__device__ inline void scan_down(double &A, double &B, int lane)
{
    compose(A, B, dpp_d<0x101>(1.0, A), dpp_d<0x101>(0.0, B));  // row_shl:1
    compose(A, B, dpp_d<0x102>(1.0, A), dpp_d<0x102>(0.0, B));  // row_shl:2
    compose(A, B, dpp_d<0x104>(1.0, A), dpp_d<0x104>(0.0, B));  // row_shl:4
    compose(A, B, dpp_d<0x108>(1.0, A), dpp_d<0x108>(0.0, B));  // row_shl:8
    const int row = lane >> 4;
    double A16 = readlane_d(A, 16), B16 = readlane_d(B, 16);
    double A48 = readlane_d(A, 48), B48 = readlane_d(B, 48);
    if (row == 0) compose(A, B, A16, B16);
    if (row == 2) compose(A, B, A48, B48);
    if (H == 1) {
        double A32 = readlane_d(A, 32), B32 = readlane_d(B, 32);
        if (row <= 1) compose(A, B, A32, B32);
    }
}

// the value of lane - 1 (wave_shr:1) and of lane + 1 (wave_shl:1), 0 for
// the first (last) lane of a line
template <int H>
__device__ inline double from_lane_below(double v, int ll)
{
    const double r = dpp_d<0x138>(0.0, v);
    return (H == 2 && ll == 0) ? 0.0 : r;
}
template <int H>
__device__ inline double from_lane_above(double v, int ll)
{
    const double r = dpp_d<0x130>(0.0, v);
    return (H == 2 && ll == 31) ? 0.0 : r;
}

// LS_LINES lines per workgroup (4 waves, LS_LINES/4 consecutive lines
// each), a line on all 64 lanes of a wave (H = 1) or two lines side by side
// on its halves (H = 2, each half LS_LINES/8 consecutive lines); LDS tile
// [LS_LINES][(64/H)*V + 1] doubles for the transposed write.  Blocked
// layout: lane ll of a line owns the V consecutive elements e = ll*V ..
// ll*V+V-1, composes their affine maps locally, and one scan per direction
// joins the lanes.
template <int V, int LS_LINES, int DIR, int H>
__device__ __forceinline__ void line_sweep_tile(const LineArgs &a, int bx, int by, double *sres)
{
    constexpr int LPW = LS_LINES / 4, LPH = LPW / H, W = 64 / H;
    static_assert(LPH * H == LPW, "lines per wave must split over the halves");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int hl = lane / W, ll = lane % W;  // half of the wave, lane within the line
    const int L = by;  // mode plane
    const int line0 = bx * LS_LINES;
    const int lenp = W * V + 1;
    const double *xin = a.xin + (long)L * a.SA;
    const double *rh = a.rh + (long)L * a.SA;
    const int len = a.len;
    const int e0 = ll * V;  // first element of this lane
    const double dtt_mtwogamma = a.dt - 2.0 * a.gamma;

    // Thomas factors: the same for every line of this mode plane
    double fa[V], fi[V], fg[V];
    {
        const double *alpha = a.alpha + (long)L * a.coef_stride;
        const double *ibet = a.ibet + (long)L * a.coef_stride;
        const double *gback = a.gback + (long)L * a.coef_stride;
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const int e = e0 + v;
            const bool ok = e < len;
            fa[v] = ok ? alpha[e] : 0.0;
            fi[v] = ok ? ibet[e] : 0.0;
            fg[v] = ok ? gback[e] : 0.0;
        }
    }
    // this half's lines first .. first + nmine - 1; lines past the end are
    // solved on clamped rows (every lane takes part in the scans) and not
    // stored
    const int first = line0 + wave * LPW + hl * LPH;
    const int nmine = min(LPH, a.nlines - first);  // may be <= 0
    auto load_row = [&](const double *row, double *dst) {
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const int e = e0 + v;
            dst[v] = (e < len) ? row[e + 1] : 0.0;
        }
    };
    // row q of the potential: q = 0 and nlines + 1 are the faces, the
    // interior rows come from the compact input or (zero when cold) from xin
    const double *xt = a.xt + (long)L * a.STi;
    const double *xf = a.xout + (long)L * a.SB;
    const long pin = compact_pitch(len);
    auto load_x = [&](int q, double *dst) {
        q = min(q, a.nlines + 1);
        if (q >= 1 && q <= a.nlines && (a.in_tiled || a.cold)) {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const int e = e0 + v;
                dst[v] = (e < len && a.in_tiled) ? xt[(q - 1) * pin + e] : 0.0;
            }
        } else {
            load_row(xin + (long)q * a.PA, dst);
        }
    };
    // rho-hat row q (an interior line): the input layout, or the compact
    // interior (rh_compact; rows of compact_pitch(len), plane stride STr)
    auto load_rh = [&](int q, double *dst) {
        q = min(q, a.nlines);
        if (a.rh_compact) {
            const double *row = a.rh + (long)L * a.STr + (long)(q - 1) * pin;
#pragma unroll
            for (int v = 0; v < V; ++v) dst[v] = (e0 + v < len) ? row[e0 + v] : 0.0;
        } else {
            load_row(rh + (long)q * a.PA, dst);
        }
    };
    double xm[V], xc[V], xn[V], rc[V], rn[V];
    load_x(first, xm);
    load_x(first + 1, xc);
    load_x(first + 2, xn);
    load_rh(first + 1, rc);
    // the right-hand side per element: t = -4 pi rc + cc xc + cn xn + cm xm,
    // plus wl * eb1 at the line's last element and wf * eb0 at its first
    // (lane weights 0 or 1, formed once; padding elements load as 0)
    double wl[V], wf[DIR == 1 ? V : 1];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        wl[v] = (e0 + v == len - 1) ? 1.0 : 0.0;
        if (DIR == 1) wf[v] = (e0 + v == 0 && !a.eqsym) ? 1.0 : 0.0;
    }
    for (int k = 0; k < LPH; ++k) {
        const int i = min(first + k, a.nlines - 1);  // interior line, row i+1 of the input layout
        // the boundary values at the ends of this line, from the face rows
        // of the other layout (xout), where the lines' ends are contiguous
        const double b0 = DIR == 1 ? xf[i + 1] : 0.0, b1 = xf[(long)(len + 1) * a.PB + i + 1];
        double xnn[V];
        const bool more = k + 1 < LPH;
        if (more) {
            load_x(first + k + 3, xnn);
            load_rh(first + k + 2, rn);
        }
        double cc, cm, cn, eb0 = 0.0, eb1;
        if (DIR == 0) {
            // knownr (helmadi.F90:240-257); the row next to the equator
            // (isym = 2, 3) leaves out its mirror K = 1 and weights K = 2
            // by dt - gamma (helmadi.F90:228-236)
            const bool eqrow = a.eqsym && i == 0;  // K = 2
            cc = eqrow ? a.dt - a.gamma : dtt_mtwogamma;
            cm = eqrow ? 0.0 : a.gamma;
            cn = a.gamma;
            eb1 = -a.alphar[a.numr - 1] * b1;
        } else {
            // knownz (helmadi.F90:284-340); the bottom face enters for
            // isym = 1 only
            const int J = i + 2;
            cc = a.elzb[(long)L * (a.numr + 1) + J] + a.dt;
            cn = -a.alphar[J];
            cm = (J > 2) ? -a.betar[J] : 0.0;
            eb0 = a.gamma * b0;
            eb1 = a.gamma * b1;
        }
        double val[V];
#pragma unroll
        for (int v = 0; v < V; ++v) {
            double t = -a.four_pi * rc[v] + cc * xc[v] + cn * xn[v];
            t = fma(cm, xm[v], t);
            t = fma(wl[v], eb1, t);
            if (DIR == 1) t = fma(wf[v], eb0, t);
            val[v] = t;
        }
        // forward sweep y_e = fa_e y_{e-1} + fi_e rhs_e: local composition,
        // wavefront scan, then the local values from the incoming y
        {
            double A = 1.0, B = 0.0;
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const double b = fi[v] * val[v];
                B = fa[v] * B + b;
                A = fa[v] * A;
            }
            scan_up<H>(A, B);
            double yin = from_lane_below<H>(B, ll);  // y of the element before this lane's first
#pragma unroll
            for (int v = 0; v < V; ++v) {
                yin = fa[v] * yin + fi[v] * val[v];
                val[v] = yin;
            }
        }
        // back substitution x_e = y_e + fg_e x_{e+1}
        {
            double A = 1.0, B = 0.0;
#pragma unroll
            for (int v = V - 1; v >= 0; --v) {
                B = fg[v] * B + val[v];
                A = fg[v] * A;
            }
            scan_down<H>(A, B, lane);
            double xin_ = from_lane_above<H>(B, ll);  // x of the element after this lane's last
#pragma unroll
            for (int v = V - 1; v >= 0; --v) {
                xin_ = fg[v] * xin_ + val[v];
                val[v] = xin_;
            }
        }
        const int jl = first - line0 + k;
        if (k < nmine)
#pragma unroll
            for (int v = 0; v < V; ++v) sres[jl * lenp + e0 + v] = val[v];
        // roll the neighbour rows
#pragma unroll
        for (int v = 0; v < V; ++v) {
            xm[v] = xc[v];
            xc[v] = xn[v];
            if (more) {
                xn[v] = xnn[v];
                rc[v] = rn[v];
            }
        }
    }
    __syncthreads();
    const int nl = min(LS_LINES, a.nlines - line0);
    if (a.out_tiled) {
        // element e of line line0 + j: 16 lines fill one aligned 128-byte
        // line of each output row
        const long pout = compact_pitch(a.nlines);
        double *yt = a.yt + (long)L * a.STo + line0;
        for (int t = threadIdx.x; t < LS_LINES * len; t += 256) {
            int j = t % LS_LINES, e = t / LS_LINES;
            if (j < nl) yt[e * pout + j] = sres[j * lenp + e];
        }
        return;
    }
    // transposed, coalesced write: element e of line i -> out[(e+1)*PB + i+1]
    double *xout = a.xout + (long)L * a.SB;
    for (int t = threadIdx.x; t < LS_LINES * len; t += 256) {
        int j = t % LS_LINES, e = t / LS_LINES;
        if (j < nl) xout[(long)(e + 1) * a.PB + line0 + j + 1] = sres[j * lenp + e];
    }
}

template <int V, int NL, int DIR, int H>
__global__ __launch_bounds__(256) void line_sweep_kernel(LineArgs a, int nx, int ny, int total)
{
    extern __shared__ double sres[];
    BH_SINGLE_LOOP(a, nx, ny, total, (line_sweep_tile<V, NL, DIR, H>(a, bx, by, sres)));
}

template <int V, int NL, int DIR, int H>
__global__ __launch_bounds__(256) void line_sweep_grouped(const LineArgs *__restrict__ args, GroupSched gs)
{
    extern __shared__ double sres[];
    BH_GROUP_LOOP(LineArgs, args, gs, (line_sweep_tile<V, NL, DIR, H>(ga, bx, by, sres)));
}

// variant = V * 100 + NL: V elements per lane (line length <= 64 V), NL
// lines per workgroup
// variant = DIR * 10000 + H * 1000 + V * 100 + NL
static int sweep_lds(int variant)
{
    const int nl = variant % 100, v = (variant / 100) % 10, h = (variant / 1000) % 10;
    return nl * ((64 / h) * v + 1) * (int)sizeof(double);
}

Op sweep_op(const LineArgs &a, int nplanes, int nl)
{
    // lines of up to 128 elements two to a wave, longer ones one to a wave
    const int H = a.len <= 128 ? 2 : 1;
    const int c = (a.len + 64 / H - 1) / (64 / H);
    const int V = c <= 1 ? 1 : c <= 2 ? 2 : c <= 4 ? 4 : 8;
    if (nl != 8 && nl != 16) nl = (V <= 4) ? 16 : 8;
    // the compact layouts take whole 128-byte segments of 16 lines: 16 lines
    // per workgroup, 32 for lines two to a wave
    if (a.in_tiled || a.out_tiled) nl = H == 2 ? 32 : 16;
    return Op::make(OP_SWEEP, (a.dir ? 10000 : 0) + H * 1000 + V * 100 + nl, a, (a.nlines + nl - 1) / nl, nplanes,
                    1);
}

void launch_line_sweep(const LineArgs &a, int nplanes, hipStream_t s) { launch_op(sweep_op(a, nplanes, 0), s); }

// axis ghost by a pi rotation: pot(1,K,L) = pot(2,K,L +- numphi/2), or by
// a mirror for isym = 3: pot(1,K,L) = pot(2,K,L) (helmadi.F90:396-405);
// the equator for isym = 2, 3: pot(J,1,L) = pot(J,2,L) (helmadi.F90:407-411)
__device__ __forceinline__ void axis_tile(const AxisArgs &a, int bx, int by, int)
{
    const int numr = a.numr, numz = a.numz, numphi = a.numphi;
    double *pot = a.pot;
    const int L = by;
    const long plane = (long)numr * numz;
    if (a.mode == EQUATOR) {
        const int J = bx * 128 + threadIdx.x;
        if (J < numr) pot[(long)L * plane + J] = pot[(long)L * plane + numr + J];
        return;
    }
    int K = bx * 128 + threadIdx.x;
    if (K >= numz) return;
    int h = numphi / 2;
    int Ls = a.mode == AXIS_MIRROR ? L : (L < h) ? L + h : L - h;
    pot[(long)L * plane + (long)K * numr] = pot[(long)Ls * plane + (long)K * numr + 1];
}

Op axis_op(double *pot, int numr, int numz, int numphi, int mode)
{
    AxisArgs a{pot, numr, numz, numphi, mode};
    return Op::make(OP_AXIS, 0, a, ((mode == EQUATOR ? numr : numz) + 127) / 128, numphi, 1);
}

// ---------------------------------------------------------------------------
// Green's functions.  G_m = sqrt(R'/R) / pi * Q_{m-1/2}(chi) with
// chi = ((z_B - z')^2 + R^2 + R'^2) / (2 R R')  (tm.F90:162-167, sm.F90:178-182)
// Q evaluated as the reference does (tm.F90:168-194): below hypr_cutoff from
// K, E and the upward Legendre recurrence; from the cutoff on by the
// hypergeometric series truncated after nser terms, here summed by Horner's
// rule in 1/chi^2 with coefficients exp(dcoefh) prepared on the host.
// ---------------------------------------------------------------------------
// Complete elliptic integrals of the first and second kind from the
// arithmetic-geometric mean (Gauss; Abramowitz & Stegun 17.6.1-17.6.4):
// a0 = 1, b0 = k' = sqrt(1 - k^2), c0 = k; a_{n+1} = (a_n + b_n) / 2,
// b_{n+1} = sqrt(a_n b_n), c_{n+1} = (a_n - b_n) / 2; K = pi / (2 a_inf),
// E = K (1 - sum_n 2^(n-1) c_n^2).  kp2 = 1 - k^2 is given separately so
// that it keeps its digits as k -> 1.  Convergence is quadratic: once
// c <= 1e-8 a, a is within ~c^2 / (4a) <= 3e-17 a of the limit and the
// terms left of the sum are below 1e-32.
__device__ void d_ellke(double k2, double kp2, double *K, double *E)
{
    double a = 1.0, b = sqrt(kp2), s = 0.5 * k2, w = 0.5;
    for (int n = 0; n < 16; ++n) {
        const double c = 0.5 * (a - b), an = 0.5 * (a + b);
        b = sqrt(a * b);
        a = an;
        w += w;
        s += w * c * c;
        if (c <= 1e-8 * a) break;
    }
    *K = 1.5707963267948966192 / a;
    *E = *K * (1.0 - s);
}

// Q_{mu-1/2}(xm) for the mode slots m = 0..mmax-1 (mu = m; isym = 3: the
// even mu = 2m, from Q_{-1/2}, Q_{3/2} and the two-step recurrence of
// tm.F90:292-330), stored to out[m * mstride] by STORE: Q_PLAIN coef * Q,
// Q_RAW Q alone, Q_IMAGE coef * (Q + the raw value already there) -- the
// image sum coef (Q(chi+) + Q(chi-)) of tm.F90:282-285 in two passes
// (rounded to the table's type).
enum { Q_PLAIN = 0, Q_RAW = 1, Q_IMAGE = 2 };
template <int STORE, class TT>
__device__ __forceinline__ void d_qstore(TT *o, double coef, double q)
{
    if (STORE == Q_PLAIN) *o = (TT)(coef * q);
    else if (STORE == Q_RAW) *o = (TT)q;
    else *o = (TT)(coef * (q + (double)*o));
}
// Mode slots m >= QREF (plans with mmax above the reference's 30,
// pot.h:3).  The upward recurrence multiplies its rounding error by about
// e^{2 mu eta} (chi = cosh eta) and the 111-term series does not converge
// for high orders near chi = 1, so these slots take Q from the ratios
// R_mu = Q_{mu-1/2} / Q_{mu-3/2} of the recurrence's minimal solution, run
// downwards (its stable direction; Gautschi, SIAM Rev. 9 (1967)):
//   R_mu = (mu - 1/2) / (2 mu chi - (mu + 1/2) R_{mu+1}),
// started at mu = M + 1 with the limit ratio e^{-eta} and M >= mu_top + 16
// + 20 / eta (the start's error is damped by e^{-2 eta} per step), and
// normalised by Q_{-1/2} = k K(k), k'^2 = (chi - 1) / (chi + 1) formed
// directly so that it keeps its digits as chi -> 1.  Slots go in chunks of
// QCHUNK, each rerunning the downward pass, so that only QCHUNK ratios are
// held.  isym = 3 (slot m: mu = 2m) multiplies two ratios per slot.
constexpr int QREF = 30, QCHUNK = 16;
template <int STORE, class TT>
__device__ void d_qhigh_store(const TableArgs &a, double xm, double coef, TT *out, long mstride)
{
    const int mmax = a.mmax, fs = a.isym == 3 ? 2 : 1;
    const double e1 = xm - 1.0, sh = sqrt(e1 * (xm + 1.0));  // sinh eta
    const double eta = log1p(e1 + sh), rinf = 1.0 / (xm + sh);
    const int mu_top = fs * (mmax - 1);
    const int M = mu_top + 16 + (int)fmin(ceil(20.0 / eta), 1.0e6);
    double kk, ee;
    const double k2 = 2.0 / (1.0 + xm);
    d_ellke(k2, e1 / (xm + 1.0), &kk, &ee);
    double qprev = sqrt(k2) * kk;  // Q_{-1/2}, then the slot below the chunk
    auto step = [&](int mu, double r) {
        return ((double)mu - 0.5) / (2.0 * (double)mu * xm - ((double)mu + 0.5) * r);
    };
    for (int slo = QREF; slo < mmax; slo += QCHUNK) {
        const int shi = min(slo + QCHUNK, mmax);
        double sr[QCHUNK];
        double r = rinf;
        int mu = M;
        for (; mu > fs * (shi - 1); --mu) r = step(mu, r);
#pragma unroll
        for (int i = QCHUNK - 1; i >= 0; --i) {
            sr[i] = 1.0;
            if (slo + i < shi) {
                for (int k = 0; k < fs; ++k, --mu) {
                    r = step(mu, r);
                    sr[i] *= r;
                }
            }
        }
        if (slo == QREF) {  // slot QREF - 1: Q_{-1/2} R_1 ... R_{fs (QREF - 1)}
            double prod = 1.0;
            for (; mu >= 1; --mu) {
                r = step(mu, r);
                prod *= r;
            }
            qprev *= prod;
        }
#pragma unroll
        for (int i = 0; i < QCHUNK; ++i) {
            if (slo + i < shi) {
                qprev *= sr[i];
                d_qstore<STORE>(out + (long)(slo + i) * mstride, coef, qprev);
            }
        }
    }
}

template <int STORE, class TT>
__device__ void d_qhalf_store(const TableArgs &a, double xm, double coef, TT *out, long mstride)
{
    // slots below the reference's mmax as the reference evaluates them
    const int mmax = min(a.mmax, QREF);
    if (a.mmax > QREF) d_qhigh_store<STORE>(a, xm, coef, out, mstride);
    if (xm < a.hypr_cutoff) {
        double mum = sqrt(2.0 / (1.0 + xm));
        double lam = sqrt(2.0 * (1.0 + xm));
        // K(k), E(k) with k = mum (tm.F90:170-171 takes them from ellf, elle)
        double kk, ee;
        d_ellke(mum * mum, (1.0 - mum) * (1.0 + mum), &kk, &ee);
        double q0 = kk * mum;
        d_qstore<STORE>(out, coef, q0);
        if (mmax < 2) return;
        if (a.isym == 3) {
            const double four_thirds = 4.0 / 3.0, one_third = 1.0 / 3.0;
            double q1 = (four_thirds * xm * xm - one_third) * mum * kk - four_thirds * xm * lam * ee;
            d_qstore<STORE>(out + mstride, coef, q1);
            for (int m = 3; m <= mmax; ++m) {
                const double nu = a.nu[m], den = (2.0 * nu - 1.0) * (nu + 2.0) * (nu + 1.0);
                const double q2 = q1 * ((2.0 * nu + 3.0) * (2.0 * nu + 1.0) * xm * xm / ((nu + 2.0) * (nu + 1.0)) -
                                        (2.0 * nu + 3.0) * nu * nu / den - (nu + 1.0) / (nu + 2.0)) -
                                  q0 * (2.0 * nu + 3.0) * (nu - 1.0) * nu / den;
                d_qstore<STORE>(out + (long)(m - 1) * mstride, coef, q2);
                q0 = q1;
                q1 = q2;
            }
            return;
        }
        double q1 = xm * mum * kk - lam * ee;
        d_qstore<STORE>(out + mstride, coef, q1);
        for (int m = 3; m <= mmax; ++m) {
            double nu = a.nu[m];
            double q2 = (2.0 * nu + 1.0) / (nu + 1.0) * xm * q1 - nu / (nu + 1.0) * q0;
            d_qstore<STORE>(out + (long)(m - 1) * mstride, coef, q2);
            q0 = q1;
            q1 = q2;
        }
    } else {
        double l2x = log(2.0 * xm);
        double y = 1.0 / (xm * xm);
        const double mstep = a.isym == 3 ? 2.0 : 1.0;
        for (int m = 0; m < mmax; ++m) {
            const double *c = a.cser + (long)m * a.nser;
            double s = c[a.nser - 1];
            for (int ir = a.nser - 2; ir >= 0; --ir) s = s * y + c[ir];
            d_qstore<STORE>(out + (long)m * mstride, coef, exp(a.coefh[m] - (mstep * (double)m + 0.5) * l2x) * s);
        }
    }
}

// Top table in blocked symmetric form.  Q_{m-1/2}(chi) depends on the
// source radius R' and the target radius R only through chi, which is
// symmetric in them, so only the blocks (I, J), J <= I, of 16 x 16 targets
// x sources are stored: TS[m][pair][kz][i][j], pair = I (I + 1) / 2 + J,
// target jo = 16 I + i, source jr = 16 J + j (block plane order tb_off),
// value Q / pi (the sqrt(R'/R) factor of tm.F90:162 is applied to the
// projections and the result).
template <class TT>
__global__ __launch_bounds__(128) void gen_topsym_table_kernel(TableArgs a, TT *T)
{
    const int nri = a.numr - 2, nzi = a.numz - 2, NB = (nri + 15) / 16, NP = NB * (NB + 1) / 2;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)NP * nzi * 256;
    if (t >= total) return;
    const int e = (int)(t & 255), i = e >> 4, j = e & 15;
    const int kz = (int)((t >> 8) % nzi);
    const int pair = (int)((t >> 8) / nzi);
    int I = 0;
    while ((I + 1) * (I + 2) / 2 <= pair) ++I;
    const int J = pair - I * (I + 1) / 2;
    const int jo = 16 * I + i, jr = 16 * J + j;
    const long mstride = (long)NP * nzi * 256;
    TT *out = T + ((long)pair * nzi + kz) * 256 + tb_off<TT>(i, j);
    if (jo >= nri || jr >= nri) {
        for (int m = 0; m < a.mmax; ++m) out[(long)m * mstride] = (TT)0;
        return;
    }
    const int jl = jr + 2, kl = kz + 2, jg = jo + 2;
    const double rhf = a.rhf_g[jl], rhfinv = a.rhfinv_g[jl];
    const double zT = a.zhf_g[a.numz], zhf = a.zhf_g[kl];
    if (a.isym == 1) {
        const double xm = 0.5 * rhfinv * a.rhfinv_g[jg] *
                          ((zT - zhf) * (zT - zhf) + a.rhf_g[jg] * a.rhf_g[jg] + rhf * rhf);
        d_qhalf_store<Q_PLAIN>(a, xm, a.piinv, out, mstride);
        return;
    }
    // isym = 2, 3: the source and its image below the equator (tm.F90:205-214)
    const double xm = 0.5 * rhfinv * a.rhfinv_g[jg] * ((zT - zhf) * (zT - zhf) + rhf * rhf + a.rhf_g[jg] * a.rhf_g[jg]);
    const double xp = 0.5 * rhfinv * a.rhfinv_g[jg] * ((zT + zhf) * (zT + zhf) + rhf * rhf + a.rhf_g[jg] * a.rhf_g[jg]);
    d_qhalf_store<Q_RAW>(a, xm, a.piinv, out, mstride);
    d_qhalf_store<Q_IMAGE>(a, xp, a.piinv, out, mstride);
}

__global__ __launch_bounds__(128) void gen_side_table_kernel(TableArgs a, double *S)
{
    const int nri = a.numr - 2, nzi = a.numz - 2;
    const int nd = 2 * nzi - 1;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nd * nri) return;
    int jr = (int)(t % nri);
    int d = (int)(t / nri);
    int jl = jr + 2;
    double dz = (double)(d - (nzi - 1)) * a.dz;
    double rB = a.rhf_g[a.numr], rBinv = 1.0 / rB;
    double rhf = a.rhf_g[jl], rhfinv = a.rhfinv_g[jl];
    double coef = sqrt(rhf * rBinv) * a.piinv;
    double xm = a.isym == 1 ? 0.5 * rhfinv * rBinv * (dz * dz + rB * rB + rhf * rhf)
                            : 0.5 * rhfinv * rBinv * (dz * dz + rhf * rhf + rB * rB);
    d_qhalf_store<Q_PLAIN>(a, xm, coef, S + (long)d * nri + jr, (long)nd * nri);
}

// isym = 2, 3: the side table's image part (sm.F90:222-226), Hankel in z:
// H[m][s][jr], s = kg + kl - 4, from z_target + z_source = (s + 1) dz
__global__ __launch_bounds__(128) void gen_hankel_table_kernel(TableArgs a, double *H)
{
    const int nri = a.numr - 2, nzi = a.numz - 2;
    const int nd = 2 * nzi - 1;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nd * nri) return;
    int jr = (int)(t % nri);
    int sidx = (int)(t / nri);
    int jl = jr + 2;
    double sz = (double)(sidx + 1) * a.dz;
    double rB = a.rhf_g[a.numr], rBinv = 1.0 / rB;
    double rhf = a.rhf_g[jl], rhfinv = a.rhfinv_g[jl];
    double coef = sqrt(rhf * rBinv) * a.piinv;
    double xp = 0.5 * rhfinv * rBinv * (sz * sz + rhf * rhf + rB * rB);
    d_qhalf_store<Q_PLAIN>(a, xp, coef, H + (long)sidx * nri + jr, (long)nd * nri);
}

void launch_gen_topsym_table(const TableArgs &a, double *T, float *T32, hipStream_t s)
{
    const long NB = (a.numr - 2 + 15) / 16;
    const long total = NB * (NB + 1) / 2 * (a.numz - 2) * 256;
    const dim3 grid((unsigned)((total + 127) / 128));
    if (T) hipLaunchKernelGGL(gen_topsym_table_kernel<double>, grid, dim3(128), 0, s, a, T);
    else hipLaunchKernelGGL(gen_topsym_table_kernel<float>, grid, dim3(128), 0, s, a, T32);
}

void launch_gen_side_table(const TableArgs &a, double *S, double *H, hipStream_t s)
{
    long total = (long)(2 * (a.numz - 2) - 1) * (a.numr - 2);
    hipLaunchKernelGGL(gen_side_table_kernel, dim3((unsigned)((total + 127) / 128)), dim3(128), 0, s, a, S);
    if (H) hipLaunchKernelGGL(gen_hankel_table_kernel, dim3((unsigned)((total + 127) / 128)), dim3(128), 0, s, a, H);
}

// ---------------------------------------------------------------------------
// Azimuthal projections of rho (bessel.F90:176-184, 226-235): per column
// (J, K), proj[0] = sum_L rho_L, and for m = 1..mmax-1 proj[m] = sum_L
// cos(m theta_L) rho_L and proj[mmax-1+m] = sum_L sin(m theta_L) rho_L, with
// theta_L = (L - 1/2) dphi (potsetup.F90:131-139).
//
// The planes fold twice.  theta -> 2 pi - theta pairs plane l (0-based) with
// N-1-l: cosines take u = rho_l + rho_{N-1-l}, sines v = rho_l - rho_{N-1-l}.
// theta -> pi - theta pairs l with N/2-1-l and multiplies cos(m theta) by
// (-1)^m and sin(m theta) by -(-1)^m.  So, with the quad of planes a,
// N-1-a, N/2-1-a, N/2+a (a < N/4) giving u_a, v_a and u_b, v_b,
//   even-m cosines  sum_a cos(m theta_a) (u_a + u_b)
//   odd-m cosines   sum_a cos(m theta_a) (u_a - u_b)
//   even-m sines    sum_a sin(m theta_a) (v_a - v_b)
//   odd-m sines     sum_a sin(m theta_a) (v_a + v_b)
// a quarter of the multiply-adds of the plain sums.  Each is a product
// (16 modes x N/4 quads) x (N/4 quads x columns) on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64, k = 4 quads per step, exact FP64 products and
// sums): a wavefront owns 16 columns and the four 16-mode row tiles of one
// PROJ_MODES chunk (grid x), a workgroup 64 columns (grid y).  The quad
// planes stream in through a ring of PROJ_RING k-steps of loads, so rho is
// read once from HBM with the matrix work hidden under it.
// ---------------------------------------------------------------------------
constexpr int PROJ_RING = 4;

__host__ __device__ inline int proj_quads(int numphi) { return numphi / 4 < 4 ? 4 : numphi / 4; }

// coef[c][a][i][rt], a < max(N/4, 4): rt 0 cos(m theta_a) with m = c
// PROJ_MODES + 2i, rt 1 the same with m + 1, rt 2 and 3 the sines; theta_a =
// (a + 1/2) dphi.  1 and 0 for m = 0; zeros for m >= mmax and for quads
// past N/4 (N < 16).  The angle is formed as the bessel rows form it (L =
// a + 1); for isym = 3 slot m has frequency 2m (potsetup.F90:120-129), all
// even, and every row tile then folds as the even ones do (ProjArgs::even).
int proj_coef_elems(int numphi, int mmax) { return (mmax + PROJ_MODES - 1) / PROJ_MODES * proj_quads(numphi) * 64; }

void proj_coefficients(int numphi, int mmax, double dphi, std::vector<double> *coef, int isym)
{
    const int Q = numphi / 4, Qt = proj_quads(numphi), nc = (mmax + PROJ_MODES - 1) / PROJ_MODES;
    coef->assign((size_t)proj_coef_elems(numphi, mmax), 0.0);
    for (int c = 0; c < nc; ++c)
        for (int q = 0; q < Q; ++q)
            for (int i = 0; i < 16; ++i)
                for (int rt = 0; rt < 4; ++rt) {
                    const int m = c * PROJ_MODES + 2 * i + (rt & 1);
                    if (m >= mmax) continue;
                    const double ang = (isym == 3 ? dphi : 0.5 * dphi) * (double)m * (2.0 * (double)(q + 1) - 1.0);
                    const double v = rt < 2 ? (m == 0 ? 1.0 : std::cos(ang)) : (m == 0 ? 0.0 : std::sin(ang));
                    (*coef)[(((size_t)c * Qt + q) * 16 + i) * 4 + rt] = v;
                }
}

__device__ __forceinline__ void proj_tile(const ProjArgs &a, int chunk, int cb, int)
{
    typedef double d2 __attribute__((ext_vector_type(2)));
    typedef double d4 __attribute__((ext_vector_type(4)));
    const int lane = threadIdx.x & 63, r16 = lane & 15, k4 = lane >> 4;
    const int wave = threadIdx.x >> 6;
    const int M = a.M, N = a.numphi, Q = N / 4, Qt = proj_quads(N), nks = Qt / 4;
    const int col = cb * 64 + wave * 16 + r16;
    const int cc = col < M ? col : M - 1;
    // this lane's quad in a k-step is 4 ks + k4; past N/4 (N < 16) its
    // coefficients are zero and its loads repeat the last quad
    const int kq = k4 < Q ? k4 : Q - 1;
    const double *pa = a.rho + (long)kq * M + cc;                // plane a
    const double *pd = a.rho + (long)(N - 1 - kq) * M + cc;      // plane N-1-a
    const double *pb = a.rho + (long)(N / 2 - 1 - kq) * M + cc;  // plane N/2-1-a
    const double *pc = a.rho + (long)(N / 2 + kq) * M + cc;      // plane N/2+a
    const double *tq = a.coef + ((long)chunk * Qt + k4) * 64 + r16 * 4;
    d4 acc[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r] = (d4){0.0, 0.0, 0.0, 0.0};
    struct Step {
        double xa, xd, xb, xc;
        d2 c01, c23;
    };
    auto load = [&](int ks, Step &s) {
        ks = min(ks, nks - 1);  // past the end: the last step again, unused
        const long o = (long)ks * 4 * M;
        s.xa = pa[o];
        s.xd = pd[-o];
        s.xb = pb[-o];
        s.xc = pc[o];
        const d2 *t = (const d2 *)(tq + (long)ks * 4 * 64);
        s.c01 = t[0];
        s.c23 = t[1];
    };
    // the odd-slot row tiles fold as odd frequencies, or (isym = 3: slot m
    // has frequency 2m) as even ones: the sign of pair b in their sums
    const double ob = a.even ? 1.0 : -1.0;
    auto step = [&](const Step &s) {
        const double ua = s.xa + s.xd, va = s.xa - s.xd;  // pair a
        const double ub = s.xb + s.xc, vb = s.xb - s.xc;  // pair N/2-1-a
        acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(s.c01.x, ua + ub, acc[0], 0, 0, 0);
        acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(s.c01.y, fma(ob, ub, ua), acc[1], 0, 0, 0);
        acc[2] = __builtin_amdgcn_mfma_f64_16x16x4f64(s.c23.x, va - vb, acc[2], 0, 0, 0);
        acc[3] = __builtin_amdgcn_mfma_f64_16x16x4f64(s.c23.y, fma(-ob, vb, va), acc[3], 0, 0, 0);
    };
    Step ring[PROJ_RING];
#pragma unroll
    for (int u = 0; u < PROJ_RING; ++u) load(u, ring[u]);
    __builtin_amdgcn_sched_barrier(0);
    for (int j = 0; j < nks; j += PROJ_RING) {
#pragma unroll
        for (int u = 0; u < PROJ_RING; ++u) {
            if (j + u < nks) step(ring[u]);
            load(j + PROJ_RING + u, ring[u]);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    // result D[4v + k4][r16] of row tile rt: mode m = chunk PROJ_MODES +
    // 2 (4v + k4) + (rt & 1), column r16
    const int mmax = a.mmax;
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int m = chunk * PROJ_MODES + 2 * (4 * v + k4) + (rt & 1);
            const bool sine = rt >= 2;
            const int row = sine ? mmax - 1 + m : m;
            if (col < M && m < mmax && (!sine || m > 0)) a.proj[(long)row * M + col] = acc[rt][v];
        }
}

Op proj_op(const double *rho, const double *coef, double *proj, int M, int numphi, int mmax, int isym)
{
    ProjArgs a{rho, coef, proj, M, numphi, mmax, isym == 3 ? 1 : 0};
    return Op::make(OP_PROJ, 0, a, (mmax + PROJ_MODES - 1) / PROJ_MODES, (M + 63) / 64, 1);
}

// ---------------------------------------------------------------------------
// Boundary contraction.  Top and bottom share one table: the bottom Green's
// function at source height z' equals the top one at -z' on the symmetric
// isym = 1 grid (bm.F90 vs tm.F90), so each table entry is contracted with
// the projections and with their z-mirror.  Deterministic reductions.
// ---------------------------------------------------------------------------
constexpr int CT_TGT = 8;  // contraction targets per workgroup
// The source points of each contraction are split into CT_SPLIT contiguous
// ranges handled by separate workgroups; the face rebuild adds the partial
// sums in a fixed order, so results are reproducible bit for bit.

// Top and bottom with the symmetric table.  The projections are first
// scaled by sqrt(R'): Pt[m][kz][jr][q] = sqrt(rhf(jr + 2)) proj, q = top
// cos, top sin (source row kz), bottom cos, bottom sin (mirrored row
// nzi - 1 - kz); rows jr >= nri are zero.
__device__ __forceinline__ void topsym_scale_tile(const ContractArgs &a, int bx, int by, int)
{
    const int numr = a.numr, numz = a.numz, mmax = a.mmax, nri = numr - 2, nzi = numz - 2;
    const int NB16 = 16 * ((nri + 15) / 16);
    const int m = by;
    const int t = bx * 256 + threadIdx.x;  // kz * NB16 + jr
    if (t >= nzi * NB16) return;
    const int kz = t / NB16, jr = t % NB16;
    const long plane = (long)numr * numz;
    double4 v = {0.0, 0.0, 0.0, 0.0};
    if (jr < nri) {
        const double *pc = a.proj + (long)m * plane;
        const double *ps = a.proj + (long)(m > 0 ? (mmax - 1 + m) : 0) * plane;
        const long pa = (long)(kz + 1) * numr + jr + 1, pm = (long)(nzi - kz) * numr + jr + 1;
        const double sq = a.aux[jr + 2];
        v.x = sq * pc[pa];
        v.y = m > 0 ? sq * ps[pa] : 0.0;
        v.z = sq * pc[pm];
        v.w = m > 0 ? sq * ps[pm] : 0.0;
    }
    ((double4 *)a.out)[((long)m * nzi + kz) * NB16 + jr] = v;
}

Op topsym_scale_op(const double *proj, const double *sqrt_rhf, double *Pt, int numr, int numz, int mmax)
{
    ContractArgs a{nullptr, proj, Pt, numr, numz, mmax, sqrt_rhf, nullptr};
    const int NB16 = 16 * ((numr - 2 + 15) / 16);
    return Op::make(OP_TOPSYM_SCALE, 0, a, ((numz - 2) * NB16 + 255) / 256, mmax, 1);
}

// Workgroup (w, m, split): row blocks I = w and NB-1-w (equal work: NB+1
// blocks in all), source planes kz of its split.  Thread (i, j) of 256
// takes element (i, j) of every block: the row part adds T_ij Pt(jr) to
// target jo, the transposed part (off the diagonal) T_ij Pt(jo) to target
// jr.  Per-target partial sums of the workgroup go to
// Ptb[m][split][w][jo][q], reduced in a fixed order by topsym_reduce.
// TT: the table's element type (FP32 tables serve the mixed solve; the
// products and sums are FP64 either way).
template <class TT>
__device__ __forceinline__ void contract_topsym_t(const ContractArgs &a, const TT *table, int bx, int by, int bz)
{
    const int numr = a.numr, numz = a.numz, nri = numr - 2, nzi = numz - 2;
    const int NB = (nri + 15) / 16, NP = NB * (NB + 1) / 2, NB16 = 16 * NB, W = (NB + 1) / 2;
    const int w = bx, m = by, sp = bz;
    constexpr int KZC = TS_PLANES;  // source planes per workgroup
    const int kz0 = sp * KZC, nkz = max(0, min(nzi, kz0 + KZC) - kz0);
    __shared__ double4 y[512];       // per-target sums (NB16 <= 512)
    __shared__ double4 red[4][16];   // column partials of the 4 waves
    __shared__ double4 sI[KZC][16];  // scaled projections of the row block, per plane
    __shared__ double4 sJ[KZC][16];  // ... of the column block
    const int tid = threadIdx.x, i = tid >> 4, j = tid & 15, lane = tid & 63, wave = tid >> 6;
    for (int k = tid; k < NB16; k += 256) y[k] = (double4){0.0, 0.0, 0.0, 0.0};
    const double4 *Pt = (const double4 *)a.proj + ((long)m * nzi + kz0) * NB16;
    const double4 zero4 = {0.0, 0.0, 0.0, 0.0};
    for (int pass = 0; pass < 2; ++pass) {
        const int I = pass == 0 ? w : NB - 1 - w;
        if (pass == 1 && I == w) break;
        __syncthreads();
        sI[i][j] = i < nkz ? Pt[(long)i * NB16 + 16 * I + j] : zero4;  // 256 threads = KZC x 16
        double4 row = zero4;
        for (int J = 0; J <= I; ++J) {
            const bool off = J != I;
            __syncthreads();  // sJ free (and sI written)
            sJ[i][j] = i < nkz ? Pt[(long)i * NB16 + 16 * J + j] : zero4;
            const TT *blk = table + (((long)m * NP + I * (I + 1) / 2 + J) * nzi + kz0) * 256 + tb_off<TT>(i, j);
            double q[KZC];
#pragma unroll
            for (int u = 0; u < KZC; ++u) q[u] = u < nkz ? (double)blk[(long)u * 256] : 0.0;
            __syncthreads();  // sJ written
            double4 col = zero4;
#pragma unroll
            for (int u = 0; u < KZC; ++u) {
                const double4 pj = sJ[u][j];
                row.x += q[u] * pj.x; row.y += q[u] * pj.y; row.z += q[u] * pj.z; row.w += q[u] * pj.w;
                const double4 pi = sI[u][i];
                col.x += q[u] * pi.x; col.y += q[u] * pi.y; col.z += q[u] * pi.z; col.w += q[u] * pi.w;
            }
            if (off) {
                // sum over i: lanes j, j+16, j+32, j+48 of each wave, then the 4 waves
#pragma unroll
                for (int d = 16; d < 64; d <<= 1) {
                    col.x += __shfl_xor(col.x, d, 64); col.y += __shfl_xor(col.y, d, 64);
                    col.z += __shfl_xor(col.z, d, 64); col.w += __shfl_xor(col.w, d, 64);
                }
                if (lane < 16) red[wave][lane] = col;
                __syncthreads();
                if (tid < 16) {
                    const double4 s0 = red[0][tid], s1 = red[1][tid], s2 = red[2][tid], s3 = red[3][tid];
                    double4 &t = y[16 * J + tid];
                    t.x += (s0.x + s1.x) + (s2.x + s3.x); t.y += (s0.y + s1.y) + (s2.y + s3.y);
                    t.z += (s0.z + s1.z) + (s2.z + s3.z); t.w += (s0.w + s1.w) + (s2.w + s3.w);
                }
            }
        }
        // sum over j: 16 consecutive lanes
#pragma unroll
        for (int d = 1; d < 16; d <<= 1) {
            row.x += __shfl_xor(row.x, d, 64); row.y += __shfl_xor(row.y, d, 64);
            row.z += __shfl_xor(row.z, d, 64); row.w += __shfl_xor(row.w, d, 64);
        }
        __syncthreads();
        if (j == 0) {
            double4 &t = y[16 * I + i];
            t.x += row.x; t.y += row.y; t.z += row.z; t.w += row.w;
        }
    }
    __syncthreads();
    const int NS = (nzi + KZC - 1) / KZC;
    double4 *out = (double4 *)a.out + (((long)m * NS + sp) * W + w) * NB16;
    for (int k = tid; k < NB16; k += 256) out[k] = y[k];
}

// ---------------------------------------------------------------------------
// Top/bottom contraction over the symmetric table, register-blocked (grids
// with nri <= 256).  A workgroup takes one mode m and TB_PLANES source
// planes, whose scaled projections (TB_PLANES x NB16 double4) sit in LDS;
// its two halves of TB_GROUPS 16-lane groups take half of the planes each.  Lane (ig, jg) of a group owns rows 4ig..4ig+3 and
// columns 4jg..4jg+3 of the group's 16 x 16 block (I, J), reads them as
// 16-byte pieces (the table's block plane order), and sums over the planes,
// in registers, the rows (targets 16I + 4ig + a, from the column-block
// projections) and the columns (targets 16J + 4jg + b, from the row-block
// ones).  The groups take the blocks in the rounds of a round-robin
// tournament among the NB row blocks (every round pairs each block with
// one other, so the 2 x TB_GROUPS targets of a round are distinct), then
// the diagonal blocks (rows only), and after each round every group adds
// its sums into the LDS totals, one half after the other: no two groups
// touch one total at once and the order is fixed, so the sums are
// reproducible bit for bit.  Table loads run a ring of plane steps ahead,
// across rounds.
// ---------------------------------------------------------------------------
constexpr int TB_GROUPS = 8;

// round r (0..n-2) of the circle-method round robin among n (even) teams:
// match k (0..n/2-1)
__device__ inline void tb_match(int n, int r, int k, int *x, int *y)
{
    if (k == 0) {
        *x = n - 1;
        *y = r;
    } else {
        *x = (r + k) % (n - 1);
        *y = (r - k + n - 1) % (n - 1);
    }
}

template <class TT>
__device__ __forceinline__ void contract_tb_t(const ContractArgs &a, const TT *table, int chunk, int m, int)
{
    typedef typename std::conditional<sizeof(TT) == 4, float4, double2>::type Piece;
    constexpr int V = 16 / sizeof(TT);        // values per 16-byte piece
    constexpr int NL = 16 / V;                // pieces per lane and plane
    constexpr int PH = TB_PLANES / 2;         // planes per half
    constexpr int RING = sizeof(TT) == 4 ? 4 : 2;
    static_assert(PH % RING == 0, "ring must divide the planes of a half");
    const int nri = a.numr - 2, nzi = a.numz - 2;
    const int NB = (nri + 15) / 16, NP = NB * (NB + 1) / 2, NB16 = 16 * NB;
    const int kz0 = chunk * TB_PLANES, nkz = min(TB_PLANES, nzi - kz0);
    const int tid = threadIdx.x, g = (tid >> 4) & (TB_GROUPS - 1), hf = tid / (16 * TB_GROUPS), l16 = tid & 15;
    const int ig = l16 >> 2, jg = l16 & 3;
    typedef double d4 __attribute__((ext_vector_type(4)));
    __shared__ d4 Ps[TB_PLANES][256];
    __shared__ d4 y[256];
    const d4 zero4 = {0.0, 0.0, 0.0, 0.0};
    // the scaled projections of the workgroup's planes straight from the
    // projection rows: sqrt(rhf) x (cos, sin) of source plane kz and of its
    // mirror nzi - 1 - kz (what topsym_scale_tile stores for the unblocked
    // contraction, computed here instead of read back)
    {
        const long plane = (long)a.numr * a.numz;
        const double *pc = a.proj + (long)m * plane, *ps = a.proj + (long)(m > 0 ? (a.mmax - 1 + m) : 0) * plane;
        for (int e = tid; e < TB_PLANES * NB16; e += 32 * TB_GROUPS) {
            const int u = e / NB16, p = e - u * NB16;
            d4 v = zero4;
            if (u < nkz && p < nri) {
                const int kz = kz0 + u;
                const long pa = (long)(kz + 1) * a.numr + p + 1, pm = (long)(nzi - kz) * a.numr + p + 1;
                const double sq = a.aux[p + 2];
                v.x = sq * pc[pa];
                v.y = m > 0 ? sq * ps[pa] : 0.0;
                v.z = sq * pc[pm];
                v.w = m > 0 ? sq * ps[pm] : 0.0;
            }
            Ps[u][p] = v;
        }
    }
    for (int e = tid; e < NB16; e += 32 * TB_GROUPS) y[e] = zero4;
    __syncthreads();

    // schedule: Roff rounds of off-diagonal blocks (sub-rounds when the
    // tournament has more matches than groups), then the diagonal blocks
    const int n = NB + (NB & 1), half = n / 2, sub = (half + TB_GROUPS - 1) / TB_GROUPS;
    const int Roff = (n - 1) * sub, Rtot = Roff + (NB + TB_GROUPS - 1) / TB_GROUPS;
    struct Blk {
        int I, J;
        bool ok;
    };
    auto block_of = [&](int R) {
        Blk b;
        R = min(R, Rtot - 1);
        if (R < Roff) {
            const int k = (R % sub) * TB_GROUPS + g;
            int x, yy;
            tb_match(n, R / sub, min(k, half - 1), &x, &yy);
            b.ok = k < half && x < NB && yy < NB;
            b.I = b.ok ? max(x, yy) : 0;
            b.J = b.ok ? min(x, yy) : 0;
        } else {
            const int d = (R - Roff) * TB_GROUPS + g;
            b.ok = d < NB;
            b.I = b.J = b.ok ? d : 0;
        }
        return b;
    };
    const int u0 = hf * PH;  // this half's first plane
    const TT *tb = table + ((long)m * NP * nzi + kz0) * 256 + l16 * V;
    // the table is streamed once per solve: nontemporal loads (12 % faster
    // than cached ones for this pattern), which also leave the caches to
    // the projections and the other streams of the solve
    typedef TT Raw __attribute__((ext_vector_type(V)));
    auto load = [&](Piece *dst, const Blk &b, int u) {
        const TT *src = tb + ((long)(b.I * (b.I + 1) / 2 + b.J) * nzi + min(u0 + u, nkz - 1)) * 256;
#pragma unroll
        for (int h = 0; h < NL; ++h) {
            const Raw v = __builtin_nontemporal_load((const Raw *)(src + h * 16 * V));
            __builtin_memcpy(&dst[h], &v, 16);
        }
    };
    Piece ring[RING][NL];
    Blk cur = block_of(0), nxt = block_of(1);
#pragma unroll
    for (int s = 0; s < RING; ++s) load(ring[s], cur, s);
    __builtin_amdgcn_sched_barrier(0);

    auto tv = [&](const Piece *t, int r, int c) -> double {
        if constexpr (sizeof(TT) == 4) {
            const float4 &q = t[r];
            return (double)(c == 0 ? q.x : c == 1 ? q.y : c == 2 ? q.z : q.w);
        } else {
            const double2 &q = t[2 * r + (c >> 1)];
            return (c & 1) ? q.y : q.x;
        }
    };
    auto fma4 = [](d4 &acc, double s, const d4 &v) {
        acc.x = fma(s, v.x, acc.x);
        acc.y = fma(s, v.y, acc.y);
        acc.z = fma(s, v.z, acc.z);
        acc.w = fma(s, v.w, acc.w);
    };
    auto sel = [](bool c, const d4 &p, const d4 &q) { return c ? p : q; };
    auto shx = [](const d4 &v, int mask) {
        return (d4){__shfl_xor(v.x, mask, 64), __shfl_xor(v.y, mask, 64), __shfl_xor(v.z, mask, 64),
                    __shfl_xor(v.w, mask, 64)};
    };
    auto add4 = [](const d4 &p, const d4 &q) { return p + q; };
    // sum of v[0..3] over the lanes of bits m1, m2 of the lane index; lane
    // with bits (b2 b1) ends with the total of v[2 b2 + b1]
    auto reduce_scatter = [&](const d4 *v, int m1, int m2) {
        const bool b1 = (l16 & m1) != 0, b2 = (l16 & m2) != 0;
        const d4 k0 = add4(sel(b1, v[1], v[0]), shx(sel(b1, v[0], v[1]), m1));
        const d4 k1 = add4(sel(b1, v[3], v[2]), shx(sel(b1, v[2], v[3]), m1));
        return add4(sel(b2, k1, k0), shx(sel(b2, k0, k1), m2));
    };

    auto round = [&](auto off_tag) {
        constexpr bool OFF = decltype(off_tag)::value;
        d4 rowp[4], colp[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) rowp[k] = colp[k] = zero4;
#pragma unroll
        for (int u = 0; u < PH; ++u) {
            const int slot = u % RING;
            Piece t[NL];
#pragma unroll
            for (int h = 0; h < NL; ++h) t[h] = ring[slot][h];
            if (u + RING < PH) load(ring[slot], cur, u + RING);
            else load(ring[slot], nxt, u + RING - PH);
            __builtin_amdgcn_sched_barrier(0);
            const d4 *pp = Ps[u0 + u];
            d4 pj[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) pj[c] = pp[16 * cur.J + 4 * jg + c];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const d4 pi = pp[16 * cur.I + 4 * ig + r];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const double q = tv(t, r, c);
                    fma4(rowp[r], q, pj[c]);
                    if (OFF) fma4(colp[c], q, pi);
                }
            }
        }
        // rows over the 4 lanes jg, columns over the 4 lanes ig
        const d4 rs = reduce_scatter(rowp, 1, 2);
        d4 cs = zero4;
        if (OFF) cs = reduce_scatter(colp, 4, 8);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (hf == h && cur.ok) {
                d4 &yr = y[16 * cur.I + 4 * ig + jg];
                yr = add4(yr, rs);
                if (OFF) {
                    d4 &yc = y[16 * cur.J + 4 * jg + ig];
                    yc = add4(yc, cs);
                }
            }
            __syncthreads();
        }
    };
    int R = 0;
    for (; R < Roff; ++R) {
        round(std::integral_constant<bool, true>());
        cur = nxt;
        nxt = block_of(R + 2);
    }
    for (; R < Rtot; ++R) {
        round(std::integral_constant<bool, false>());
        cur = nxt;
        nxt = block_of(R + 2);
    }
    d4 *out = (d4 *)a.out + ((long)m * tb_chunks(a.numz) + chunk) * NB16;
    for (int e = tid; e < NB16; e += 32 * TB_GROUPS) out[e] = y[e];
}

__device__ __forceinline__ void contract_tb64_tile(const ContractArgs &a, int bx, int by, int bz)
{
    contract_tb_t<double>(a, a.table, bx, by, bz);
}
__device__ __forceinline__ void contract_tb32_tile(const ContractArgs &a, int bx, int by, int bz)
{
    contract_tb_t<float>(a, a.table32, bx, by, bz);
}

__device__ __forceinline__ void contract_topsym_tile(const ContractArgs &a, int bx, int by, int bz)
{
    contract_topsym_t<double>(a, a.table, bx, by, bz);
}
__device__ __forceinline__ void contract_topsym32_tile(const ContractArgs &a, int bx, int by, int bz)
{
    contract_topsym_t<float>(a, a.table32, bx, by, bz);
}

Op contract_topsym_op(const double *TS, const float *TS32, const double *proj, const double *sqrt_rhf,
                      const double *Pt, double *Ptb, int numr, int numz, int mmax)
{
    if (tb_blocked(numr)) {
        ContractArgs a{TS, proj, Ptb, numr, numz, mmax, sqrt_rhf, TS32, 0};
        return Op::make(OP_CONTRACT_TOPSYM, TS ? 2 : 3, a, tb_chunks(numz), mmax, 1);
    }
    ContractArgs a{TS, Pt, Ptb, numr, numz, mmax, nullptr, TS32, 0};
    const int NB = (numr - 2 + 15) / 16;
    return Op::make(OP_CONTRACT_TOPSYM, TS ? 0 : 1, a, (NB + 1) / 2, mmax, topsym_splits(numz));
}

// Stb[m][q][jo] = sum over (split, w) of Ptb / sqrt(rhf(jo + 2)), in a fixed order
__device__ __forceinline__ void topsym_reduce_tile(const ContractArgs &a, int bx, int by, int)
{
    const int nri = a.numr - 2, NB = (nri + 15) / 16, NB16 = 16 * NB, W = (NB + 1) / 2;
    const int m = by, e = bx * 256 + threadIdx.x, q = e / nri, jo = e % nri;
    if (q >= 4) return;
    const int nk = a.nparts ? a.nparts : topsym_splits(a.numz) * W;
    const double *P = a.proj + ((long)m * nk * NB16 + jo) * 4 + q;
    double s = 0.0;
    for (int k0 = 0; k0 < nk; k0 += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = k0 + u < nk ? P[(long)(k0 + u) * NB16 * 4] : 0.0;
#pragma unroll
        for (int u = 0; u < 8; ++u) s += v[u];
    }
    a.out[((long)m * 4 + q) * nri + jo] = s / a.aux[jo + 2];
}

Op topsym_reduce_op(const double *Ptb, const double *sqrt_rhf, double *Stb, int numr, int numz, int mmax)
{
    ContractArgs a{nullptr, Ptb, Stb, numr, numz, mmax, sqrt_rhf, nullptr, tb_blocked(numr) ? tb_chunks(numz) : 0};
    return Op::make(OP_TOPSYM_REDUCE, 0, a, (4 * (numr - 2) + 255) / 256, mmax, 1);
}

// Side: Ss[split][m][q][ko], q = cos, sin; Toeplitz table S[m][ko - kz + nzi - 1][jr]
// (isym = 2, 3: plus the image table H[m][ko + kz][jr]).  A workgroup takes
// one mode m and the SD_PLANES source planes of one split, for every
// target ko.  Thread jr reads its column's projections of those planes
// once, into registers; the targets follow in tiles of SD_TGT, and the
// table values of a tile (its SD_TGT x SD_PLANES (ko, kz) pairs lie on
// SD_TGT + SD_PLANES - 1 diagonals) sit in a register window that slides
// by SD_TGT from tile to tile, so each table value of the split's range of
// diagonals is loaded once per column.  After each tile the workgroup adds
// its columns in a fixed order (wavefront butterflies, then the 4
// wavefronts) into LDS totals, column chunks of 256 in order, and the
// totals are written at the end: the sums are reproducible bit for bit.
// SD_GROUPS > 1 would split the targets between that many workgroups (grid
// z), reading the projections once per group; one group measured fastest.
constexpr int SD_TGT = 16, SD_GROUPS = 1;
static_assert(SD_PLANES == SD_TGT, "the window holds 2 SD_TGT - 1 diagonals");

template <bool IMG>
__device__ __forceinline__ void contract_side_t(const ContractArgs &a, int bx, int by, int bz)
{
    const double *__restrict__ S = a.table;
    const double *__restrict__ proj = a.proj;
    const int numr = a.numr, numz = a.numz, mmax = a.mmax;
    const int nri = numr - 2, nzi = numz - 2, nd = 2 * nzi - 1;
    const long plane = (long)numr * numz;
    const int sp = bx, m = by;
    const int kz0 = sp * SD_PLANES, nk = min(SD_PLANES, nzi - kz0);
    const int ntall = (nzi + SD_TGT - 1) / SD_TGT, tpg = (ntall + SD_GROUPS - 1) / SD_GROUPS;
    const int T0 = bz * tpg, ntile = min(ntall, T0 + tpg);  // this workgroup's tiles [T0, ntile)
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const double *sm = S + (long)m * nd * nri;
    const double *hm = IMG ? a.hankel + (long)m * nd * nri : nullptr;
    const double *pc = proj + (long)m * plane;
    const double *ps = proj + (long)(m > 0 ? mmax - 1 + m : 0) * plane;
    __shared__ double tot[2][512];  // per-target totals (nzi <= 512)
    __shared__ double red[4][2 * SD_TGT];
    for (int k = tid; k < 2 * nzi; k += 256) tot[k / nzi][k % nzi] = 0.0;
    for (int jr0 = 0; jr0 < nri; jr0 += 256) {
        const int jr = jr0 + tid;
        const bool jok = jr < nri;
        const int jc = jok ? jr : nri - 1;
        double xc[SD_PLANES], xs[SD_PLANES];
#pragma unroll
        for (int k = 0; k < SD_PLANES; ++k) {
            const bool ok = jok && k < nk;
            const long o = (long)(min(kz0 + k, nzi - 1) + 1) * numr + jc + 1;
            const double c = pc[o], sv = ps[o];
            xc[k] = ok ? c : 0.0;
            xs[k] = ok && m > 0 ? sv : 0.0;
        }
        // tile T: d = ko - kz + nzi - 1 = dlo + i, i = t - k + SD_TGT - 1; the
        // image index s = ko + kz = slo + t + k
        auto ldw = [&](int d) { return sm[(long)min(max(d, 0), nd - 1) * nri + jc]; };
        auto ldh = [&](int sx) { return hm[(long)min(sx, nd - 1) * nri + jc]; };
        const int dlo0 = nzi - 1 - kz0 - (SD_TGT - 1);
        double w[2 * SD_TGT - 1], wh[IMG ? 2 * SD_TGT - 1 : 1];
#pragma unroll
        for (int i = 0; i < 2 * SD_TGT - 1; ++i) {
            w[i] = ldw(dlo0 + SD_TGT * T0 + i);
            if constexpr (IMG) wh[i] = ldh(kz0 + SD_TGT * T0 + i);
        }
#pragma unroll 1
        for (int T = T0; T < ntile; ++T) {
            double acc[2][SD_TGT];
#pragma unroll
            for (int t = 0; t < SD_TGT; ++t) acc[0][t] = acc[1][t] = 0.0;
#pragma unroll
            for (int k = 0; k < SD_PLANES; ++k)
#pragma unroll
                for (int t = 0; t < SD_TGT; ++t) {
                    const double g = IMG ? w[SD_TGT - 1 + t - k] + wh[t + k] : w[SD_TGT - 1 + t - k];
                    acc[0][t] = fma(g, xc[k], acc[0][t]);
                    acc[1][t] = fma(g, xs[k], acc[1][t]);
                }
            // the next tile's window: SD_TGT - 1 values move down, SD_TGT load
            if (T + 1 < ntile) {
#pragma unroll
                for (int i = 0; i < SD_TGT - 1; ++i) {
                    w[i] = w[i + SD_TGT];
                    if constexpr (IMG) wh[i] = wh[i + SD_TGT];
                }
#pragma unroll
                for (int i = SD_TGT - 1; i < 2 * SD_TGT - 1; ++i) {
                    w[i] = ldw(dlo0 + SD_TGT * (T + 1) + i);
                    if constexpr (IMG) wh[i] = ldh(kz0 + SD_TGT * (T + 1) + i);
                }
            }
            // sum over the 256 threads: reduce-scatter over lane bits 5..1
            // (after the step on bit b the lane keeps the half of its values
            // selected by that bit), the pair, then the 4 wavefronts in order
            double v[2 * SD_TGT];
#pragma unroll
            for (int t = 0; t < SD_TGT; ++t) {
                v[t] = acc[0][t];
                v[SD_TGT + t] = acc[1][t];
            }
            int n = 2 * SD_TGT;
#pragma unroll
            for (int b = 32; b >= 2; b >>= 1) {
                const bool hi = (lane & b) != 0;
                n >>= 1;
#pragma unroll
                for (int i = 0; i < SD_TGT; ++i) {
                    if (i < n) {
                        const double keep = hi ? v[n + i] : v[i], give = hi ? v[i] : v[n + i];
                        v[i] = keep + __shfl_xor(give, b, 64);
                    }
                }
            }
            // lane holds value lane >> 1 summed over the 32 lanes that share
            // lane bit 0; add the pair
            const double pr = v[0] + __shfl_xor(v[0], 1, 64);
            if ((lane & 1) == 0) red[wave][lane >> 1] = pr;
            __syncthreads();
            if (tid < 2 * SD_TGT) {
                const int q = tid / SD_TGT, ko = SD_TGT * T + tid % SD_TGT;
                const double r = (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]);
                if (ko < nzi) tot[q][ko] += r;
            }
            __syncthreads();  // red is free for the next tile
        }
    }
    const int k0 = SD_TGT * T0, nt = min(nzi, SD_TGT * ntile) - k0;  // this workgroup's targets
    for (int k = tid; k < 2 * nt; k += 256) {
        const int q = k / nt, ko = k0 + k % nt;
        a.out[(((long)sp * mmax + m) * 2 + q) * nzi + ko] = tot[q][ko];
    }
}

__device__ __forceinline__ void contract_side_tile(const ContractArgs &a, int bx, int by, int bz)
{
    contract_side_t<false>(a, bx, by, bz);
}
__device__ __forceinline__ void contract_side_img_tile(const ContractArgs &a, int bx, int by, int bz)
{
    contract_side_t<true>(a, bx, by, bz);
}

Op contract_side_op(const double *S, const double *proj, double *Ss, int numr, int numz, int mmax, const double *H)
{
    ContractArgs a{S, proj, Ss, numr, numz, mmax, nullptr, nullptr, 0, H};
    return Op::make(OP_CONTRACT_SIDE, H ? 1 : 0, a, side_splits(numz), mmax, SD_GROUPS);
}

// Boundary faces in helmadi's Fourier mode order.  bessel rebuilds each
// physical face value as a trigonometric sum over the azimuthal modes
// (bessel.F90:396-488),
//   F(theta) = f (S_C(0) + sum_{m>=1} 2 S_C(m) cos(m theta) + 2 S_S(m) sin(m theta)),
//   theta_L = (L - 1/2) dphi,  f = -dr dz dphi,
// and helmadi transforms it along phi (helmadi.F90:131-156).  The discrete
// Fourier transform of such a sum is exact in closed form: with
// X_k = sum_L F_L e^{-2 pi i k (L-1)/N} and alpha = m dphi / 2,
//   X_k = f N [ S_C(0) [k = 0 mod N]
//             + sum_{m = k mod N} ((S_C c + S_S s) + i (S_C s - S_S c))
//             + sum_{m = -k mod N} ((S_C c + S_S s) - i (S_C s - S_S c)) ],
// c = cos alpha, s = sin alpha, and the mode rows hold a_k = Re X_k (rows
// 0..N/2) and -b_k = Im X_k (rows N/2+1..N-1, k = 1..N/2-1).  So the faces
// come straight from the contraction sums, with no phi-space face array and
// no transform (aliasing included when mmax > N/2).  With isym = 3 slot m
// carries frequency p = 2m (its cos, sin rows have angle 2m theta) and
// counts where p = +-k (mod N); f then doubles (bessel.F90:170-174).
// the face value fi (0 <= fi < 2 nri + nzi) of mode row `row`
__device__ double face_mode_value(const RebuildArgs &a, int fi, int row)
{
    const double *__restrict__ Stb = a.Stb;
    const double *__restrict__ Ss = a.Ss;
    const int numr = a.numr, numz = a.numz, N = a.numphi, mmax = a.mmax;
    const int nri = numr - 2, nzi = numz - 2;
    const double *sc, *ss;
    long st, sst;  // per-m stride, per-split stride
    int idx;
    int nsp;  // partial sums stored per value
    if (fi < nri) { idx = fi; sc = Stb; ss = Stb + nri; st = 4 * (long)nri; nsp = a.nsplit_tb; }
    else if (fi < 2 * nri) { idx = fi - nri; sc = Stb + 2 * nri; ss = Stb + 3 * nri; st = 4 * (long)nri; nsp = a.nsplit_tb; }
    else { idx = fi - 2 * nri; sc = Ss; ss = Ss + nzi; st = 2 * (long)nzi; nsp = side_splits(numz); }
    sst = st * mmax;
    // partial sums of the source ranges of the contraction, in a fixed order
    // top (q = 0, 1) and bottom (q = 2, 3) straight from the contraction's
    // partials when Ptb is given: the reduction's sum, in its order
    const bool fromP = a.Ptb && fi < 2 * nri;
    const int q0 = fi < nri ? 0 : 2;
    const int NB16 = 16 * ((nri + 15) / 16);
    auto S_p = [&](int m, int q) {  // eight loads in flight at a time
        const double *P = a.Ptb + ((long)m * a.nk * NB16 + idx) * 4 + q;
        double s = 0.0;
        for (int k0 = 0; k0 < a.nk; k0 += 8) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = k0 + u < a.nk ? P[(long)(k0 + u) * NB16 * 4] : 0.0;
#pragma unroll
            for (int u = 0; u < 8; ++u) s += v[u];
        }
        return s / a.sqrt_rhf[idx + 2];
    };
    auto S_c = [&](int m) {
        if (fromP) return S_p(m, q0);
        double v = sc[m * st + idx];
        for (int p = 1; p < nsp; ++p) v += sc[p * sst + m * st + idx];
        return v;
    };
    auto S_s = [&](int m) {
        if (fromP) return S_p(m, q0 + 1);
        double v = ss[m * st + idx];
        for (int p = 1; p < nsp; ++p) v += ss[p * sst + m * st + idx];
        return v;
    };
    const bool imag = row > N / 2;
    const int k = imag ? row - N / 2 : row;
    double re = 0.0, im = 0.0;
    if (k % N == 0) re = S_c(0);
    // the m = 1..mmax-1 with fstep m = k or -k (mod N), in increasing order
    // (k = 0..N/2: fstep m = k + jN and N - k + jN, m >= 1; none when
    // fstep = 2 and k is odd)
    const int fs = a.fstep, stp = N / fs;
    int mp = (k == 0 ? N : k) / fs, mq = (k == 0 ? N : N - k) / fs;
    if (k % fs) mp = mq = mmax;
    for (;;) {
        const int m = min(mp, mq);
        if (m >= mmax) break;
        const bool plus = m == mp, minus = m == mq;
        const double c = a.besc[(long)m * N], s = a.bess[(long)m * N];
        const double C = S_c(m), S = S_s(m);
        const double r = C * c + S * s, q = C * s - S * c;
        if (plus) { re += r; im += q; mp += stp; }
        if (minus) { re += r; im -= q; mq += stp; }
    }
    return a.factor * (double)N * (imag ? im : re);
}

__device__ __forceinline__ void face_modes_tile(const RebuildArgs &a, int bx, int by, int)
{
    const int numr = a.numr, numz = a.numz;
    const int nri = numr - 2, nzi = numz - 2;
    const int nf = 2 * nri + nzi;
    const int fi = bx * 128 + threadIdx.x;
    const int row = a.kmap ? (by < a.kmap ? by : a.numphi / 2 + 1 + (by - a.kmap)) : by;
    if (a.flag && bx == 0 && by == 0 && threadIdx.x == 0) {
        a.flag[1] = a.flag[0];
        a.flag[0] = 0.f;
    }
    if (a.X1 && fi == nf) {  // the outer corners of the side face are zero (bessel.F90:428, 456, 487)
        const long plane = (long)numr * numz;
        double *x1 = a.X1 + row * plane;
        x1[(long)(numz - 1) * numr + (numr - 1)] = 0.0;
        x1[numr - 1] = 0.0;
        if (a.X2) {
            double *x2 = a.X2 + row * plane;
            x2[a.x2_klayout ? (long)(numr - 1) * numz + (numz - 1) : (long)(numz - 1) * numr + (numr - 1)] = 0.0;
            x2[a.x2_klayout ? (long)(numr - 1) * numz : numr - 1] = 0.0;
        }
        return;
    }
    if (fi >= nf) return;
    if (!a.bottom && fi >= nri && fi < 2 * nri) return;  // isym = 2, 3: no bottom face
    const double v = face_mode_value(a, fi, row);
    a.F[(long)row * nf + fi] = v;
    if (a.r) {
        // the interior point next to this face point, when this thread owns
        // it: the top face owns row K = numz - 1, the bottom face row K = 2,
        // the side face J = numr - 1 for K = 3..numz - 2.  A x0 there has
        // only face terms: alphar(J) x(J+1) (east face) and -gamma (x(K-1) +
        // x(K+1)) (bottom and top faces); other faces' values at corners
        // come from face_mode_value.
        int J, K;
        double xe = 0.0, xkm = 0.0, xkp = 0.0;
        bool own = true;
        if (fi < nri) {
            J = fi + 2; K = numz - 1;
            xkp = v;
            if (K - 1 == 1) xkm = face_mode_value(a, nri + fi, row);
        } else if (fi < 2 * nri) {
            J = fi - nri + 2; K = 2;
            own = K != numz - 1;
            xkm = v;
        } else {
            J = numr - 1; K = fi - 2 * nri + 2;
            own = K >= 3 && K <= numz - 2;
            xe = v;
        }
        if (own) {
            if (fi < 2 * nri && J + 1 == numr) xe = face_mode_value(a, 2 * nri + (K - 2), row);
            const double Ax = a.alphar[J] * xe - a.gamma * (xkm + xkp);
            const long ci = (long)(K - 1) * numr + (J - 1);
            const long rp = a.rh_pitch ? a.rh_pitch : (long)numr * numz;
            const double rr = -a.four_pi * (double)a.rh[row * rp + ci] - Ax;
            a.r[(long)row * nri * nzi + (long)(K - 2) * nri + (J - 2)] = (float)rr;
        }
    }
    if (a.X1) {  // top J = fi + 2, K = numz; bottom K = 1; side J = numr, K = fi - 2 nri + 2
        int J, K;
        if (fi < nri) { J = fi + 2; K = numz; }
        else if (fi < 2 * nri) { J = fi - nri + 2; K = 1; }
        else { J = numr; K = fi - 2 * nri + 2; }
        const long plane = (long)numr * numz;
        a.X1[row * plane + (long)(K - 1) * numr + (J - 1)] = v;
        if (a.X2)
            a.X2[row * plane + (a.x2_klayout ? (long)(J - 1) * numz + (K - 1) : (long)(K - 1) * numr + (J - 1))] = v;
    }
}

void face_modes_add_residual(Op *op, float *r, const float *rh, long rh_pitch, const double *alphar, double gamma,
                             double four_pi)
{
    RebuildArgs a = op->as<RebuildArgs>();
    a.r = r;
    a.rh = rh;
    a.rh_pitch = rh_pitch;
    a.alphar = alphar;
    a.gamma = gamma;
    a.four_pi = four_pi;
    std::memcpy(op->args, &a, sizeof a);
}

Op face_modes_op(const double *Stb, const double *Ss, const double *besc, const double *bess, double *FF, int numr,
                 int numz, int numphi, int mmax, double factor, int nsplit_tb, double *X1, double *X2, int x2_klayout,
                 int isym, bool reach_rows)
{
    RebuildArgs a{Stb, Ss, besc, bess, FF, numr, numz, numphi, mmax, factor, nsplit_tb, X1, X2, x2_klayout,
                  isym == 3 ? 2 : 1, isym == 1 ? 1 : 0};
    const int nf = 2 * (numr - 2) + (numz - 2) + (X1 ? 1 : 0);  // + the corner thread
    // fstep = 1: slot m lands on cosine row m and sine row numphi/2 + m
    // (m >= 1) alone while m < numphi / 2
    if (reach_rows && isym != 3 && 2 * mmax <= numphi) a.kmap = mmax;
    return Op::make(OP_FACE_MODES, 0, a, (nf + 127) / 128, a.kmap ? 2 * mmax - 1 : numphi, 1);
}

// ---------------------------------------------------------------------------
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
// two waves per SIMD: the register budget is held to 256 per lane (the
// register-blocked top contraction, whose LDS allows two workgroups per CU)
#define BH_KERNEL_PAIR2(NAME, ARGS_T, NT)                                                         \
    __global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2))) void NAME##_kernel(   \
        ARGS_T a, int nx, int ny, int total)                                                      \
    {                                                                                             \
        BH_SINGLE_LOOP(a, nx, ny, total, (NAME##_tile(a, bx, by, bz)));                           \
    }                                                                                             \
    __global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2))) void NAME##_grouped(  \
        const ARGS_T *__restrict__ args, GroupSched gs)                                           \
    {                                                                                             \
        BH_GROUP_LOOP(ARGS_T, args, gs, (NAME##_tile(ga, bx, by, bz)));                           \
    }
BH_KERNEL_PAIR(fill, FillArgs, 256)
BH_KERNEL_PAIR(transpose, TransposeArgs, 256)
BH_KERNEL_PAIR(axis, AxisArgs, 128)
BH_KERNEL_PAIR(proj, ProjArgs, 256)
BH_KERNEL_PAIR(contract_side, ContractArgs, 256)
BH_KERNEL_PAIR(contract_side_img, ContractArgs, 256)
BH_KERNEL_PAIR(topsym_scale, ContractArgs, 256)
BH_KERNEL_PAIR(contract_topsym, ContractArgs, 256)
BH_KERNEL_PAIR(contract_topsym32, ContractArgs, 256)
BH_KERNEL_PAIR2(contract_tb64, ContractArgs, 256)
BH_KERNEL_PAIR(contract_tb32, ContractArgs, 256)
#undef BH_KERNEL_PAIR2
BH_KERNEL_PAIR(topsym_reduce, ContractArgs, 256)
BH_KERNEL_PAIR(face_modes, RebuildArgs, 128)
#undef BH_KERNEL_PAIR

namespace {
struct KernelRef {
    const void *single, *grouped;
    int threads;
};

bool kernel_ref(int kind, int variant, KernelRef *k, int *lds)
{
    *lds = 0;
#define BH_REF(K_, NAME, NT)                                                                      \
    case K_:                                                                                      \
        *k = {(const void *)&NAME##_kernel, (const void *)&NAME##_grouped, NT};                   \
        return true;
    switch (kind) {
        BH_REF(OP_FILL, fill, 256)
        BH_REF(OP_TRANSPOSE, transpose, 256)
        BH_REF(OP_AXIS, axis, 128)
        BH_REF(OP_PROJ, proj, 256)
    case OP_CONTRACT_SIDE:
        if (variant == 1) *k = {(const void *)&contract_side_img_kernel, (const void *)&contract_side_img_grouped, 256};
        else *k = {(const void *)&contract_side_kernel, (const void *)&contract_side_grouped, 256};
        return true;
        BH_REF(OP_TOPSYM_SCALE, topsym_scale, 256)
    case OP_CONTRACT_TOPSYM:
        switch (variant) {
        case 0: *k = {(const void *)&contract_topsym_kernel, (const void *)&contract_topsym_grouped, 256}; break;
        case 1: *k = {(const void *)&contract_topsym32_kernel, (const void *)&contract_topsym32_grouped, 256}; break;
        case 2: *k = {(const void *)&contract_tb64_kernel, (const void *)&contract_tb64_grouped, 256}; break;
        default: *k = {(const void *)&contract_tb32_kernel, (const void *)&contract_tb32_grouped, 256}; break;
        }
        return true;
        BH_REF(OP_TOPSYM_REDUCE, topsym_reduce, 256)
        BH_REF(OP_FACE_MODES, face_modes, 128)
    case OP_SWEEP:
        *lds = sweep_lds(variant);
        switch (variant) {
#define BH_SW(V, NL, D, H)                                                                       \
    case D * 10000 + H * 1000 + V * 100 + NL:                                                    \
        *k = {(const void *)&line_sweep_kernel<V, NL, D, H>, (const void *)&line_sweep_grouped<V, NL, D, H>, 256}; \
        return true;
            BH_SW(4, 8, 0, 1) BH_SW(4, 16, 0, 1) BH_SW(8, 8, 0, 1) BH_SW(8, 16, 0, 1)
            BH_SW(1, 8, 0, 2) BH_SW(1, 16, 0, 2) BH_SW(2, 8, 0, 2) BH_SW(2, 16, 0, 2) BH_SW(4, 8, 0, 2) BH_SW(4, 16, 0, 2)
            BH_SW(4, 8, 1, 1) BH_SW(4, 16, 1, 1) BH_SW(8, 8, 1, 1) BH_SW(8, 16, 1, 1)
            BH_SW(1, 8, 1, 2) BH_SW(1, 16, 1, 2) BH_SW(2, 8, 1, 2) BH_SW(2, 16, 1, 2) BH_SW(4, 8, 1, 2) BH_SW(4, 16, 1, 2)
            BH_SW(1, 32, 0, 2) BH_SW(2, 32, 0, 2) BH_SW(4, 32, 0, 2) BH_SW(1, 32, 1, 2) BH_SW(2, 32, 1, 2) BH_SW(4, 32, 1, 2)
#undef BH_SW
        default:
            return false;
        }
    default:
        return false;
    }
#undef BH_REF
}
}  // namespace

bool kernels_launch(const Op &op, hipStream_t s)
{
    KernelRef k;
    int lds;
    if (!kernel_ref(op.kind, op.variant, &k, &lds)) return false;
    int nx = op.nx, ny = op.ny, total = (int)op.tiles();
    void *argv[] = {(void *)op.args, (void *)&nx, (void *)&ny, (void *)&total};
    const int nb = persistent_blocks(total, kernel_occupancy(k.single, k.threads, lds));
    (void)hipLaunchKernel(k.single, dim3(nb), dim3(k.threads), argv, lds, s);
    return true;
}

bool kernels_launch_grouped(int kind, int variant, const void *d_args, const GroupSched &gs, int nblocks, hipStream_t s)
{
    KernelRef k;
    int lds;
    if (!kernel_ref(kind, variant, &k, &lds)) return false;
    GroupSched g = gs;
    const void *a = d_args;
    void *argv[] = {(void *)&a, (void *)&g};
    (void)hipLaunchKernel(k.grouped, dim3(nblocks), dim3(k.threads), argv, lds, s);
    return true;
}

bool kernels_occupancy(int kind, int variant, int *blocks)
{
    KernelRef k;
    int lds;
    if (!kernel_ref(kind, variant, &k, &lds)) return false;
    *blocks = kernel_occupancy(k.grouped, k.threads, lds);
    return true;
}

}  // namespace bh
