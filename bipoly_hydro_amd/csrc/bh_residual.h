// bh_residual.h -- the FP64 residual tile of the mixed solve, shared by the
// residual kernels (bh_mixed.hip) and the fused refinement step
// (bh_spectral.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "bh_mixed.h"

namespace bh {

// ---------------------------------------------------------------------------
// r = -4 pi rho-hat - A x - w on the interior, with x = X + dlt when dlt is
// given (then also written to Xo).  A x = B_r x + B_z x as the ADI defines
// them (potsetup.F90:175-228): B_r x = brb x + betar x(J-1) + alphar x(J+1)
// (no J-1 term at J = 2), B_z x = 2 gamma x - gamma (x(K-1) + x(K+1)); face
// values enter through the neighbours.  All of X, dlt, w, rh are full
// (numr x numz) planes in J-layout; dlt and w have zero faces.  r is written
// compact: r[(K-2)*nri + (J-2)].
// ---------------------------------------------------------------------------
// Tile: 256 consecutive J of one plane and ROWS rows of K (2, 4 or 8: the
// op variant, an autotuning choice; the default RES_ROWS), the column's rows
// K0-1 .. K0+ROWS and the J coefficients loaded once.  Every point's
// arithmetic is the same for any ROWS.
// The variant fixes at compile time whether delta and w are present (UPD:
// the refinement pass), whether Xo is written (WX; a solve whose x is
// formed later by the inverse DFT from the corrections leaves it out),
// whether a second correction is added (ADD2) and whether X is a cold start
// (COLD), so that the row loop carries no pointer tests.  R0 (with COLD and
// UPD) takes the first residual r0 = g - A x0 in place of rh and of X: the
// corrections are then the whole of x, with zero faces, and
// r = r0 - w - A (dlt + dlt2), the same r as from X = x0 and rh up to the
// FP32 rounding of r0 (a warm solve's refinement, which then reads no x0).
// NOW (with UPD, COLD, WX and ADD2, no w): the first residual of a warm
// start whose state was left as its faces (in X) and the two corrections of
// the cold solve before it: x0 = X's faces + (dlt + dlt2) inside, written
// to Xo, and r = -4 pi rh - A x0.
constexpr int RES_ROWS = 4;

// tx is the thread's column in the tile; with lout the tile's r goes to
// lout[(K - 2) * lpitch + J - 2] (LDS, the fused spectral step) instead of r.
template <int ROWS, bool UPD, bool COLD, bool EDGE, bool WX, bool ADD2 = false, bool R0 = false, bool NOW = false>
__device__ __forceinline__ void residual_tile_v(const ResidualArgs &a, int bx, int by, int bz, int tx = threadIdx.x,
                                                float *lout = nullptr, int lpitch = 0)
{
    const int numr = a.numr, nri = numr - 2, nzi = a.numz - 2;
    const int L = bz;
    const long plane = (long)numr * a.numz;
    const int jr = bx * 256 + tx;
    if (jr >= nri) return;
    const int J = jr + 2;
    const int K0 = by * ROWS + 2, K1 = min(K0 + ROWS, nzi + 2);
    const double *__restrict__ X = a.X + L * plane;
    const float *__restrict__ dl = UPD ? a.dlt + L * plane : nullptr;
    const float *__restrict__ dl2 = ADD2 ? a.dlt2 + L * plane : nullptr;
    static_assert(!R0 || (COLD && UPD && !WX), "R0 forms the corrections' residual only");
    static_assert(!NOW || (COLD && UPD && WX && ADD2 && !R0), "NOW forms a warm start's first residual only");
    const float *__restrict__ rh = a.rh + L * (a.rh_pitch ? a.rh_pitch : plane);
    const float *__restrict__ r0 = R0 ? a.r0 + (long)L * nri * nzi : nullptr;
    const float *__restrict__ w = UPD && !NOW ? a.w + L * plane : nullptr;
    double *__restrict__ Xo = WX ? a.Xo + L * plane : nullptr;
    float *__restrict__ r32 = a.r + (long)L * nri * nzi;
    const double brbJ = a.brb[(long)L * (numr + 1) + J], aJ = a.alphar[J], bJ = a.betar[J];
    // cold start: the interior of X is zero and dlt (dlt2) is zero on the
    // faces, so interior values are the corrections alone and face values X
    // alone.  Two corrections are summed first (exact in FP64 for FP32
    // values this close), as the inverse DFT sums them.
    auto corr = [&](long i) { return ADD2 ? (double)dl[i] + (double)dl2[i] : (double)dl[i]; };
    auto val = [&](long i) { return UPD ? X[i] + corr(i) : X[i]; };
    auto inner = [&](long i) { return COLD ? (UPD ? corr(i) : 0.0) : val(i); };
    auto face = [&](long i) { return R0 ? 0.0 : COLD ? X[i] : val(i); };
    // the lane-dependent cases (the east face at J = numr - 1, no west term
    // at J = 2) are selects, not branches; every address read is inside the
    // plane.  All of the tile's loads are issued before the first use and
    // the stores follow the arithmetic (stores share the load counter, so
    // interleaving them would serialise the rows).
    const bool east_face = J + 1 == numr, has_west = J > 2;
    // interior tiles (not EDGE) have ROWS rows and no face row: their
    // loads are straight-line code (a branch around a load makes the
    // compiler wait for every load in flight where the paths join)
    const int nk = EDGE ? K1 - K0 : ROWS;
    const long c0 = (long)(K0 - 1) * numr + (J - 1);  // (J, K0)
    double xcv[ROWS + 2], xev[ROWS], xwv[ROWS];  // xcv[i]: row K0 - 1 + i
    float rv[ROWS], wv[ROWS];
#pragma unroll
    for (int i = 0; i < ROWS + 2; ++i) {
        const int K = K0 - 1 + i;
        const long ci = c0 + (long)(i - 1) * numr;
        if (!EDGE) {
            xcv[i] = inner(ci);
        } else {
            xcv[i] = 0.0;
            if (i <= nk + 1) xcv[i] = (K == 1 || K == a.numz) ? face(ci) : inner(ci);
        }
    }
#pragma unroll
    for (int i = 0; i < ROWS; ++i) {
        const long ci = c0 + (long)i * numr;
        xev[i] = 0.0;
        xwv[i] = 0.0;
        rv[i] = 0.f;
        wv[i] = 0.f;
        if (i < nk) {
            if (COLD) {
                const double xf = R0 ? 0.0 : X[ci - (J - 1) + numr - 1];  // the row's face value, one address per row
                const double xi = inner(ci + 1);
                xev[i] = east_face ? xf : xi;
            } else {
                xev[i] = val(ci + 1);
            }
            xwv[i] = inner(ci - 1);
            rv[i] = R0 ? r0[(long)(K0 - 2 + i) * nri + jr] : rh[ci];
            if (UPD && !NOW) wv[i] = w[ci];
        }
    }
#pragma unroll
    for (int i = 0; i < ROWS; ++i) {
        if (i < nk) {
            const long ci = c0 + (long)i * numr;
            const double xc = xcv[i + 1];
            double Ax = brbJ * xc + aJ * xev[i];
            Ax += has_west ? bJ * xwv[i] : 0.0;
            Ax += 2.0 * a.gamma * xc - a.gamma * (xcv[i] + xcv[i + 2]);
            double r = R0 ? (double)rv[i] - Ax : -a.four_pi * (double)rv[i] - Ax;
            if (UPD && !NOW) r -= (double)wv[i];
            if (WX) Xo[ci] = xc;
            if (lout) lout[(K0 - 2 + i) * lpitch + jr] = (float)r;
            else r32[(long)(K0 - 2 + i) * nri + jr] = (float)r;
        }
    }
}

}  // namespace bh
