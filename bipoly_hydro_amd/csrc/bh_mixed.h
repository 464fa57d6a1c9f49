// bh_mixed.h -- declarations of the mixed-precision device code (bh_mixed.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace bh {

enum { EPI_PLAIN = 0, EPI_MASK = 1, EPI_FILTER = 2 };

// C[z] = op(A[z]) * op(B[z]), FP32 MFMA, row-major; op = transpose when the
// launch asks for it (A stored K x R, B stored M x K).  batch = gridDim.z.
struct GemmF32Args {
    const float *A; long lda, sA;
    const void *B; long ldb, sB;  // float or double
    float *C; long ldc, sC;
    int R, M, Kd, batch;
    int mask_numr, mask_numz;     // EPI_MASK
    const float *mu;              // EPI_FILTER: per-column factor
    const float *Rin; float *C2;  // EPI_FILTER: C2 = Rin - mu * acc (same indexing as C)
};
void launch_gemm_f32(const GemmF32Args &g, bool ta, bool tb, bool b_double, int epi, hipStream_t s);

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
void launch_tridiag_shift(const float *cin, float *yout, const double *ibet, const double *gam,
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
