// bh_spectral.h -- the fused per-mode spectral step of the mixed solve
// (bh_spectral.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "bh_mixed.h"
#include "bh_ops.h"

namespace bh {

// One workgroup per azimuthal mode L does, with the mode's data in LDS:
//   RT = r^T S                     folded sine transform (FP32 MFMA)
//   W = mu . (M_r RT), C = RT - W  when mr is given (FP32 MFMA)
//   (B_r + lambda) D = C           Thomas sweeps per spectral column (FP32)
//   delta = S D^T, w = S W^T       folded inverse transforms into planes
//                                  (or dlt += S D^T, the refinement passes
//                                  that accumulate their corrections)
// r is compact [L][nzi][nri]; delta and w are full numr x numz planes whose
// interior is written (faces untouched).
struct SpecArgs {
    const float *r;
    const float *sfb, *sfa;       // folded sine table in B- and A-fragment order
    const float *mr;              // [mode] M_r as split-bf16 A fragments, or null
    const float *mu;              // [nzi] filter factors, spectral slot order
    const float *ibet;            // [mode][nri][nzi] pivots 1/bet of B_r + lambda
    const double *betar, *alphar;
    float *dlt, *w;
    int numr, numz, nphi;         // mode number of plane L: L, or L - nphi/2 past nphi/2
    int acc;                      // without mr: add delta to dlt's interior instead of storing it
    const float *sfb16, *sfa16;   // nzi >= 64: the same tables split into bf16 hi, lo fragments
    // a correction (no mr) with FP32 transforms instead of split bf16: the
    // solve's only refinement pass, whose error no later pass contracts
    int fp32t;
    // a refinement pass with split-bf16 transforms (nzi >= 64) may form its
    // own residual r = g - w - A x from ra (the residual kernel's arguments;
    // ra.r unused) into LDS instead of reading r: fuse_res != 0
    int fuse_res;
    ResidualArgs ra;
};

// Sizes the fused kernel takes: nri a multiple of 64 up to 256 and nzi one
// of 32, 64, 128; the environment variable BH_SPEC_KERNEL=0 turns it off.
bool spectral_fused_ok(int nri, int nzi);
Op spectral_op(const SpecArgs &a, int nplanes, int nzi);
// The folded sine table (ceil(n/2) x n, spectral slot columns) in the two
// fragment orders the kernel reads, in FP32 and (nzi >= 64; else left
// empty) split into bf16 hi and lo parts.
void spectral_fragments(int nzi, std::vector<float> *sfb, std::vector<float> *sfa, std::vector<float> *sfb16,
                        std::vector<float> *sfa16);

}  // namespace bh
