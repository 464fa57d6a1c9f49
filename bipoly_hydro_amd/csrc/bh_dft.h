// bh_dft.h -- real DFT along phi as panel-wise MFMA products (bh_dft.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "bh_ops.h"

namespace bh {

struct DftArgs {
    const double *x;  // input: N rows (phi, or modes for the inverse) x M columns, row pitch ldx
    long ldx;
    void *out;        // output: N rows x M columns, row pitch ldo (float for the FP32 forward)
    long ldo;
    const void *A;    // matrix operand in MFMA fragment order (dft_*_fragments)
    int N, M;
    int mask_numr, mask_numz;  // forward: zero the rhop layers (potential_solver.F90:70-78); 0 = no mask
    const float *xadd;  // inverse at N = 512: input is x + xadd (FP32), ...
    double *xstore;     // ... and x + xadd is also stored here (both null: plain input)
    // inverse at N = 512 with xadd, a cold start: on the interior points of
    // the mask_numr x mask_numz grid the input is xint + xadd (FP32 both,
    // each widened to FP64), and only the face points of x are read
    const float *xint;
    // inverse at N = 512: when nonzero, the grid's numr; the axis ghost
    // column J = 1 is then written from J = 2 of the plane rotated by pi
    // (helmadi.F90:396-405) in place of its own transform
    int axis_numr;
    int max_blocks;     // single launches at N = 512: cap on the persistent grid (0: one per panel)
    // forward FP32 at N = 512 with the mask: also rout[L][K-2][J-2] =
    // rscale * out for the interior points (null: not written)
    float *rout;
    double rscale;
    // inverse at N = 512 with xint and xadd and warm != 0: the input is
    // x + (xint + xadd) at every point (the FP32 corrections are zero on
    // the faces), and that sum is also stored to xstore
    int warm;
    // forward with the mask, isym = 2, 3: the layers K = 1, 2 next to the
    // equator keep their density (potential_solver.F90:76-78 zeroes them
    // for isym = 1 only)
    int mask_keep_bottom;
    // the cold inverse that keeps the state as corrections (x not stored):
    // *set_flag = 1 (see bh_plan::lazy_flag)
    float *set_flag;
};

enum { DFT_FWD_F32 = 0, DFT_FWD_F64 = 1, DFT_INV_F64 = 2 };
// N <= 512 (power of two >= 4)
constexpr int DFT_MAX_N = 512;
Op dft_op(const DftArgs &a, int kind);
size_t dft_lds_bytes(int kind, int N);
// Matrix operands in fragment order: forward [row tile][k step][lane] with
// cosine rows k = 0..N/2-1 then sine rows k = 1..N/2-1; inverse P rows then
// Q rows, n = 0..N/2-1.  At N = 512: the forward operand of size N/4 followed
// by the radix-4 twiddles (cos, sin)(2 pi k r / N), r = 0..3, k = 0..N/4;
// the inverse operand of size N/8 followed by the radix-8 twiddles, r =
// 0..7, k = 0..N/16.
void dft_forward_fragments(int N, std::vector<float> *f32, std::vector<double> *f64);
void dft_inverse_fragments(int N, std::vector<double> *f64);

}  // namespace bh
