// bh_mixed.h -- declarations of the mixed-precision device code (bh_mixed.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "bh_ops.h"

namespace bh {

// Sine mode k (0-based) held in spectral slot m of the mixed solve's
// transforms: the ceil(n/2) even modes first, then the odd ones, so that
// the folded sine transform writes and reads each half as one block.
inline int spec_mode(int m, int n)
{
    const int he = (n + 1) / 2;
    return m < he ? 2 * m : 2 * (m - he) + 1;
}

// Entry (a, k) of the orthonormal sine basis of tridiag(-1, 2, -1) of order
// n: sqrt(2/(n+1)) sin(pi (a+1)(k+1)/(n+1)), exactly zero where the sine is.
inline double sine_basis(int a, int k, int n)
{
    const long ab = ((long)(a + 1) * (k + 1)) % (2 * (n + 1));
    if (ab % (n + 1) == 0) return 0.0;
    return std::sqrt(2.0 / (n + 1)) * std::sin(3.141592653589793238 * (double)ab / (n + 1));
}

// dlt and w are given together (the refinement pass; Xo = X + dlt is then
// written unless Xo is null) or both null (the first residual).  dlt2 (with
// dlt, and Xo null) is a second FP32 correction: x = X + (dlt + dlt2).
// dlt, dlt2 and Xo with w null and cold set: the first residual of a warm
// start from a cold solve's kept corrections, x = X's faces + (dlt + dlt2)
// inside, also written to Xo.
struct ResidualArgs {
    const double *X; const float *dlt; const float *w; const float *rh;
    double *Xo; float *r;
    const float *dlt2;
    const double *brb, *alphar, *betar;
    double gamma, four_pi;
    int numr, numz;
    int cold;  // X is zero on the interior (a cold start): read it only on the faces
    long rh_pitch;  // floats from one plane of rh to the next (0: numr x numz)
    // the fused refinement step of a warm solve: the first residual
    // r0 = g - A x0 (FP32, compact like r), taken in place of X and rh, so
    // that r = r0 - w - A (dlt + dlt2) with zero faces (cold must be 1)
    const float *r0;
    // with dlt, dlt2 and Xo (a warm start of a plan whose cold solves may
    // keep the state as corrections): write x0 out only when *lazy_flag is
    // set, else the plain first residual from X
    const float *lazy_flag;
};
// Xo = X + (dlt + dlt2) on the interior (dlt2 may be null); cold: X is zero
// on the interior and is not read
struct UpdateArgs {
    const double *X;
    const float *dlt;
    double *Xo;
    int numr, numz;
    const float *dlt2;
    int cold;
};
struct TridiagArgs {
    const float *cin;
    float *yout;
    const float *ibet, *gam;
    const double *betar;
    int nri, nzi;
};
// rows: K rows per tile, 2, 4 or 8 (0: the default); the same results for any
Op residual_op(const ResidualArgs &a, int nplanes, int rows = 0);
Op add_update_op(const double *X, const float *dlt, double *Xo, int numr, int numz, int nplanes,
                 const float *dlt2 = nullptr, int cold = 0);
Op tridiag_shift_op(const float *cin, float *yout, const float *ibet, const float *gam, const double *betar, int nri,
                    int nzi, int nplanes);

struct MrArgs {
    const double *brb, *alphar, *betar, *dt;
    int numr, nri, nsteps, mode0, nmodes;
    double *scratch_v, *scratch_u, *scratch_g;  // nmodes*nri*nri each
    float *Mr;                                   // [numphi][nri][nri]
    int frag;  // store M_r(L) as bf16 hi/lo pairs in the bf16 MFMA A-fragment order
    double *Mr64;  // when given (in place of Mr): M_r(L)^T in FP64, [mode][j'][j] = M_r[j][j']
};
void launch_build_mr(const MrArgs &a, hipStream_t s);
void launch_fillf(float *p, size_t n, float v, hipStream_t s);

}  // namespace bh
