// bh_mixed.h -- declarations of the mixed-precision device code (bh_mixed.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace bh {

struct ResidualArgs {
    const double *X; const float *dlt; const float *w; const float *rh;
    double *Xo; float *r;
    const double *brb, *alphar, *betar;
    double gamma, four_pi;
    int numr, numz;
};
void launch_residual64(const ResidualArgs &a, int nplanes, hipStream_t s);
void launch_add_update(const double *X, const float *dlt, double *Xo, int numr, int numz, int nplanes,
                       hipStream_t s);
void launch_tridiag_shift(const float *cin, float *yout, const float *ibet, const float *gam,
                          const double *betar, int nri, int nzi, int nplanes, hipStream_t s);

struct MrArgs {
    const double *brb, *alphar, *betar, *dt;
    int numr, nri, nsteps, mode0, nmodes;
    double *scratch_v, *scratch_u, *scratch_g;  // nmodes*nri*nri each
    float *Mr;                                   // [numphi][nri][nri]
};
void launch_build_mr(const MrArgs &a, hipStream_t s);
void launch_fillf(float *p, size_t n, float v, hipStream_t s);
void launch_face_scatter_j(const double *FF, double *X, int numr, int numz, int numphi, hipStream_t s);

}  // namespace bh
