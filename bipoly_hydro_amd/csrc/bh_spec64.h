// bh_spec64.h -- the FP64 spectral solve (bh_spec64.hip): the mixed solve's
// exact spectral form of the truncated Peaceman-Rachford iteration, with
// every product and sweep in FP64 and no refinement.
#pragma once
#include <hip/hip_runtime.h>

#include "bh_ops.h"

namespace bh {

// r0 = -4 pi rho-hat - A X0 on the interior, compact [L][K-2][J-2]; X0 in
// J-layout (cold: read only on the faces); rho-hat FP64 in J-layout.
struct Res64Args {
    const double *X, *rh;
    double *r;
    const double *brb, *alphar, *betar;
    double gamma, four_pi;
    int numr, numz, cold;
};
Op res64_op(const Res64Args &a, int nplanes);

// Per line (L, k) of T (= S r0) and U (= T M_r(L)^T), both compact
// [L][k][J-2]: C = T - mu_k U, then (B_r(L) + lambda_k) D = C along J, D
// into U (T holds the elimination factors afterwards).
struct Thomas64Args {
    double *T, *U;
    const double *mu, *lam;   // [nzi] per sine mode k
    const double *brb, *alphar, *betar;  // brb per plane [L][numr + 1]
    int numr, nri, nzi;
};
Op thomas64_op(const Thomas64Args &a, int nplanes);

// X[L] interior = (cold ? 0 : X) + delta[L][K-2][J-2]
struct Update64Args {
    double *X;
    const double *dlt;
    int numr, numz, cold;
};
Op update64_op(const Update64Args &a, int nplanes);

}  // namespace bh
