// bh_internal.h -- declarations shared by the plan code and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <vector>

#include "bh_ops.h"

namespace bh {

// Geometry on the device side.  J-layout: [mode/phi][K][J], J fastest, pitch
// numr (identical to the reference's (numr, numz, numphi) column-major array).
// K-layout: [mode][J][K], K fastest, pitch numz.
struct Geo {
    int numr, numz, numphi, mmax;
    int nri, nzi;  // interior counts numr-2, numz-2
    double dr, dz, dphi, gamma, four_pi, piinv;
    int isym;      // 1 no symmetry, 2 equatorial, 3 equatorial + pi (setup_frac.F90:313)
};

// Line-sweep arguments for one ADI half-step (helmadi.F90:195-353).
struct LineArgs {
    const double *xin;   // potential, input layout
    double *xout;        // potential, output (transposed) layout; its face rows give the lines' end values
    const double *rh;    // -> rho-hat, input layout
    long SA, PA;         // input: plane stride, row pitch
    long SB, PB;         // output: plane stride, row pitch
    int nlines, len;     // interior lines / elements per line
    const double *alpha, *ibet, *gback;  // Thomas factors, per element
    long coef_stride;    // per-mode stride of the factors (0: shared by all modes)
    int dir;             // 0 radial (lines = K, elements = J), 1 vertical
    double dt, gamma, four_pi;
    const double *alphar, *betar;  // [numr + 1], 1-based
    const double *elzb;            // [numphi][numr + 1], 1-based J
    int numr;
    // isym = 2, 3: K = 2 borders the equator, whose mirror cell K = 1 takes
    // K = 2's value (helmadi.F90:211-238, 334-340)
    int eqsym;
    // Compact interiors between the half-steps: element e of interior line
    // i read at xt[i * P(len) + e] (plane stride STi) and written at
    // yt[e * P(nlines) + i] (plane stride STo), P = compact_pitch, so that a
    // workgroup's 16 lines write whole aligned 128-byte lines.  The faces (and, unless in_tiled, the interior) stay
    // in xin's layout; out_tiled = 0 writes xout.  cold: xin's interior is
    // zero (a cold start) and is not read.
    const double *xt;
    double *yt;
    long STi, STo;
    int in_tiled, out_tiled, cold;
    // rh in the same compact form as xt (plane stride STr)
    int rh_compact;
    long STr;
};
// Row pitch (doubles) of a compact interior with rows of n elements: n
// rounded up to 16 (aligned 128-byte lines) plus 16, so that the rows of a
// plane do not sit a power of two apart (rows 1 and 2 KB apart, as the 128
// and 256 of the 258 x 130 grid would be, ran its half-steps 4-7 % slower).
__host__ __device__ inline long compact_pitch(int n) { return (long)(n + 15) / 16 * 16 + 16; }
// Plane stride of a compact interior read as nlines lines of len elements.
__host__ __device__ inline long tiled_stride(int nlines, int len) { return (long)nlines * compact_pitch(len); }

struct FillArgs {
    double *p;
    long n;
    double v;
};
struct TransposeArgs {
    const double *src;
    double *dst;
    int numr, numz;
};
struct AxisArgs {
    double *pot;
    int numr, numz, numphi;
    int mode;  // AXIS_ROTATE, AXIS_MIRROR or EQUATOR
};
enum { AXIS_ROTATE = 0, AXIS_MIRROR = 1, EQUATOR = 2 };
struct ContractArgs {
    const double *table, *proj;
    double *out;
    int numr, numz, mmax;
    const double *aux;     // sqrt(rhf) [numr + 1], 1-based J (symmetric top table)
    const float *table32;  // FP32 symmetric top table (in place of table)
    int nparts;            // partial sums per value (top/bottom reduction)
    const double *hankel;  // side, isym = 2, 3: image table H[m][kg + kl - 4][jr] (null: none)
};
struct ProjArgs {
    const double *rho;   // [numphi][M]
    const double *coef;  // [chunk][numphi/2][64] folded mode coefficients
    double *proj;        // [2*mmax-1][M]
    int M, numphi, mmax;
    int even;            // isym = 3: every slot's frequency 2m is even
};
struct RebuildArgs {
    const double *Stb, *Ss, *besc, *bess;
    double *F;  // faces in Fourier mode order
    int numr, numz, numphi, mmax;
    double factor;
    int nsplit_tb;  // partial sums per top/bottom value in Stb
    double *X1, *X2;  // when X1 is given: each face value also placed at its (J, K) in X1 and X2
    int x2_klayout;
    int fstep;   // azimuthal frequency of mode slot m is fstep * m (2 for isym = 3)
    int bottom;  // the bottom face exists (isym = 1)
    // with r (a cold mixed start, isym = 1): also the first residual at the
    // interior points next to a face (rows K = 2 and numz - 1, column J =
    // numr - 1, the only ones where A x0 is not zero), r = -4 pi rh - A x0
    // with x0 zero inside, compact r[L][K-2][J-2]; the forward DFT has
    // written -4 pi rh at every other interior point
    float *r;
    const float *rh;
    long rh_pitch;
    const double *alphar;
    double gamma, four_pi;
    // kmap > 0: the launch covers only the mode rows that boundary modes
    // reach (fstep = 1, mmax <= numphi / 2): cosine rows 0..kmap-1 and sine
    // rows numphi/2 + 1 .. numphi/2 + kmap - 1, grid row by -> mode row; the
    // faces of every other row are zero in every solve and already zero in
    // the potential buffers
    int kmap;
    // flag (bh_plan::lazy_flag): one thread moves flag[0] to flag[1] and
    // clears flag[0]
    float *flag;
    // Ptb (with sqrt_rhf): the top and bottom sums are taken straight from
    // the contraction's nk partials [m][k][16 NB][4] (summed in k order and
    // divided by sqrt(rhf), as the reduction kernel would), Stb unused
    const double *Ptb, *sqrt_rhf;
    int nk;
};

// grid-stride fill for plan set-up
void launch_fill(double *p, size_t n, double v, hipStream_t s);
// Ops of the solve sequence.  fill_op writes n values but its launch grid
// covers `capacity` (n may be 0: same grid, no work).
Op fill_op(double *p, size_t n, size_t capacity, double v);
// [plane][K][J] -> [plane][J][K]
// J layout -> compact K-layout interior (rows of compact_pitch(numz - 2))
Op transpose_op(const double *src, double *dst, int numr, int numz, int nplanes);
// one ADI half-step; nl = lines per workgroup (8 or 16, 0: default)
Op sweep_op(const LineArgs &a, int nplanes, int nl);
void launch_line_sweep(const LineArgs &a, int nplanes, hipStream_t s);
// axis ghost by a pi rotation (helmadi.F90:396-405; mode AXIS_ROTATE), or
// a mirror (isym = 3, AXIS_MIRROR); EQUATOR: pot(J, 1, L) = pot(J, 2, L)
// (helmadi.F90:407-411)
Op axis_op(double *pot, int numr, int numz, int numphi, int mode = AXIS_ROTATE);

// Green's tables ------------------------------------------------------------
struct TableArgs {
    int numr, numz, mmax, nser;
    int isym;               // 2, 3: add the image below the equator; 3: even modes
    double hypr_cutoff, piinv, dz;
    const double *rhf_g;    // [numr + 1] 1-based
    const double *rhfinv_g; // [numr + 1]
    const double *zhf_g;    // [numz + 1]
    const double *coefh;    // [mmax]
    const double *cser;     // [mmax][nser] = exp(dcoefh)
    const double *nu;       // [mmax + 1] 1-based
};
// blocked symmetric top table TS[m][pair][kz][16][16] (see bh_kernels.hip)
// into T (FP64) or, when T is null, T32 (FP32)
void launch_gen_topsym_table(const TableArgs &a, double *T, float *T32, hipStream_t s);
inline size_t topsym_table_elems(int numr, int numz, int mmax)
{
    const size_t NB = (numr - 2 + 15) / 16;
    return (size_t)mmax * NB * (NB + 1) / 2 * (numz - 2) * 256;
}
// side table in Toeplitz form S[m][d][jr], d = kg - kl + nzi - 1; with
// isym = 2, 3 also the image part H[m][s][jr], s = kg + kl - 4 (H may be null)
void launch_gen_side_table(const TableArgs &a, double *S, double *H, hipStream_t s);

// boundary contraction and face rebuild (bessel.F90:176-296, 396-488);
// the side contraction writes one partial sum per split of SD_PLANES source
// planes, which the face kernel adds in order
constexpr int SD_PLANES = 16;
__host__ __device__ inline int side_splits(int numz) { return (numz - 2 + SD_PLANES - 1) / SD_PLANES; }
// the symmetric top/bottom contraction takes 16 source planes per workgroup
constexpr int TS_PLANES = 16;
__host__ __device__ inline int topsym_splits(int numz) { return (numz - 2 + TS_PLANES - 1) / TS_PLANES; }
// proj: [2*mmax-1][numz*numr] projections (row 0 = m0, 1..mmax-1 cos, mmax.. sin)
// Stb: [mmax][4][nri] (top cos, top sin, bottom cos, bottom sin)
// Ss:  [side_splits][mmax][2][nzi]
// proj from rho: the azimuthal mode sums of every (J, K) column, the planes
// folded in quads (FP64 matrix cores); coef from proj_coefficients
constexpr int PROJ_MODES = 32;  // modes per chunk (grid x)
int proj_coef_elems(int numphi, int mmax);
Op proj_op(const double *rho, const double *coef, double *proj, int M, int numphi, int mmax, int isym = 1);
void proj_coefficients(int numphi, int mmax, double dphi, std::vector<double> *coef, int isym = 1);
// Order of the 256 values of one 16 x 16 block plane of the symmetric top
// table: sixteen 4 x 4 sub-blocks (rows 4ig.., columns 4jg..), stored so
// that 16-byte piece h of sub-block g = 4ig + jg sits at (h * 16 + g) * 16 B
// (FP32: piece h = row 4ig + h; FP64: row 4ig + h/2, columns 4jg + 2(h%2)..)
template <class TT>
__host__ __device__ inline int tb_off(int i, int j)
{
    if (sizeof(TT) == 4) return (((i & 3) * 16 + (i >> 2) * 4 + (j >> 2)) << 2) + (j & 3);
    return ((((i & 3) * 2 + ((j & 3) >> 1)) * 16 + (i >> 2) * 4 + (j >> 2)) << 1) + (j & 1);
}
// the register-blocked contraction takes TB_PLANES source planes per
// workgroup and grids with nri <= 256 (partials Ptb[m][chunk][NB16][4])
constexpr int TB_PLANES = 8;
__host__ __device__ inline int tb_chunks(int numz) { return (numz - 2 + TB_PLANES - 1) / TB_PLANES; }
__host__ __device__ inline bool tb_blocked(int numr) { return numr - 2 <= 256; }
// symmetric-table contraction: scale the projections (Pt, mmax*nzi*NB16*4;
// the unblocked form only), contract (Ptb), reduce into Stb (one slot)
Op topsym_scale_op(const double *proj, const double *sqrt_rhf, double *Pt, int numr, int numz, int mmax);
// TS is the FP64 table; with TS null, TS32 the FP32 one.  The blocked form
// (tb_blocked) scales proj by sqrt_rhf itself; the unblocked one reads Pt.
Op contract_topsym_op(const double *TS, const float *TS32, const double *proj, const double *sqrt_rhf,
                      const double *Pt, double *Ptb, int numr, int numz, int mmax);
Op topsym_reduce_op(const double *Ptb, const double *sqrt_rhf, double *Stb, int numr, int numz, int mmax);
Op contract_side_op(const double *S, const double *proj, double *Ss, int numr, int numz, int mmax,
                    const double *H = nullptr);
// faces in Fourier mode order FF[mode row][fi]: fi in [0,nri) top J=fi+2,
// [nri,2nri) bottom, [2nri,2nri+nzi) side, straight from the contraction sums;
// with X1, each value also placed in X1 (J-layout) and X2 (K-layout when
// x2_klayout, else J-layout; may be null), the outer corners set to zero.
// isym = 2, 3 have no bottom face (those values are not written) and
// isym = 3 mode slot m has frequency 2m.
// a cold mixed start (isym = 1): the face op also writes the first residual
// at the interior points next to a face (see RebuildArgs::r)
void face_modes_add_residual(Op *op, float *r, const float *rh, long rh_pitch, const double *alphar, double gamma,
                             double four_pi);
// reach_rows: launch over the rows boundary modes reach only (see
// RebuildArgs::kmap; taken when isym < 3 and mmax <= numphi / 2)
Op face_modes_op(const double *Stb, const double *Ss, const double *besc, const double *bess, double *FF, int numr,
                 int numz, int numphi, int mmax, double factor, int nsplit_tb, double *X1 = nullptr,
                 double *X2 = nullptr, int x2_klayout = 0, int isym = 1, bool reach_rows = false);

}  // namespace bh
