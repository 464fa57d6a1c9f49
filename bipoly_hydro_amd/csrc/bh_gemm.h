// bh_gemm.h -- the MFMA GEMM template of bh_gemm.hip.
#pragma once
#include <hip/hip_runtime.h>

namespace bh {

enum { B_PLAIN = 0, B_FOLD_PLUS = 1, B_FOLD_MINUS = 2, B_ADD_F32 = 3 };
enum { E_PLAIN = 0, E_MASK = 1, E_FILTER = 2, E_INV_COMBINE = 3 };

// C[z] = op(A[z]) * op(B'[z]) in row-major storage; z = 0..batch-1 with
// strides sA, sB, sC (elements).  op = transpose when requested (A stored
// Kd x R, B stored M x Kd).  B' is formed on load according to the B mode.
struct GemmArgs {
    const void *A; long lda, sA;
    const void *B; long ldb, sB;
    const float *Badd;            // B_ADD_F32: FP32 addend, same indexing as B
    void *C; long ldc, sC;
    int R, M, Kd, batch;
    int nfold;                    // N of the folded real DFT (B_FOLD_*, E_INV_COMBINE)
    int mask_numr, mask_numz;     // E_MASK
    const float *mu;              // E_FILTER: per-column factor
    const void *Rin; void *C2;    // E_FILTER: C2 = Rin - mu * acc, same indexing as C
    int tile;                     // 0 wide 64x256, 1 square 128x128, 2 small 64x64
};

void gemm64(const GemmArgs &g, int bmode, int epi, hipStream_t s);
void gemm32(const GemmArgs &g, int bmode, int epi, bool ta, bool tb, bool b_double, hipStream_t s);

}  // namespace bh
