// bh_gemm.h -- the MFMA GEMM template of bh_gemm.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "bh_ops.h"

namespace bh {

enum { B_PLAIN = 0, B_FOLD_PLUS = 1, B_FOLD_MINUS = 2, B_ADD_F32 = 3 };
enum { E_PLAIN = 0, E_MASK = 1, E_FILTER = 2, E_INV_COMBINE = 3 };
// Folded sine transforms of length n = nfold (the DST-I of the mixed solve):
// DST_FWD  C[r][m'] = sum_a (A[a][r] +- A[n-1-a][r]) B[a][m'], A stored
//          n x R (transposed), the sign + for the first ceil(n/2) output
//          columns (even sine modes) and - for the other floor(n/2) (odd);
//          K = ceil(n/2) or floor(n/2), B the folded table (ldb >= n).
// DST_INV  C[z][c] = E + O and C[n-1-z][c] = E - O for z < ceil(n/2), with
//          E = sum_{m'<ceil(n/2)} A[z][m'] B[c][m'] and O the same over the
//          floor(n/2) odd columns; A the folded table, B stored M x n.
enum { DST_NONE = 0, DST_FWD = 1, DST_INV = 2 };

// C[z] = op(A[z]) * op(B'[z]) in row-major storage; z = 0..batch-1 with
// strides sA, sB, sC (elements).  op = transpose when requested (A stored
// Kd x R, B stored M x Kd).  B' is formed on load according to the B mode.
struct GemmArgs {
    const void *A; long lda, sA;
    const void *B; long ldb, sB;
    const float *Badd;            // B_ADD_F32: FP32 addend, same indexing as B
    void *C; long ldc, sC;
    int R, M, Kd, batch;
    int nfold;                    // N of the folded real DFT (B_FOLD_*, E_INV_COMBINE) or sine transform
    int mask_numr, mask_numz;     // E_MASK
    const float *mu;              // E_FILTER: per-column factor
    const void *Rin; void *C2;    // E_FILTER: C2 = Rin - mu * acc, same indexing as C
    int tile;                     // 0: 64x256, 1: 128x128, 2: 64x64, 3: 64x128, 4: 128x64 (rows x cols)
    int col_fastest;              // tile numbering: 0 row tiles fastest, 1 column tiles fastest
    int mask_keep_bottom;         // E_MASK: keep the layers K = 1, 2 (isym = 2, 3)
};

// The instantiated (element types, B mode, epilogue, transposes) combinations.
enum GemmInst {
    G64_PLAIN = 0, G64_FOLDP_MASK, G64_FOLDM_MASK, G64_FOLDP, G64_FOLDM, G64_INV, G64_ADD, G64_ADD_INV,
    G32D_FOLDP_MASK, G32D_FOLDM_MASK, G32_TA, G32_TB, G32_FILTER, G32_PLAIN, G32_DST_FWD, G32_DST_INV, G_NINST
};
int gemm_inst64(int bmode, int epi);
int gemm_inst32(int bmode, int epi, bool ta, bool tb, bool b_double);
// The Op of one product (variant = inst * 8 + tile; grid from the tile shape).
Op gemm_op(const GemmArgs &g, int inst);
void gemm_tile_rows_cols(int tile, int *rows, int *cols);

void gemm64(const GemmArgs &g, int bmode, int epi, hipStream_t s);
void gemm32(const GemmArgs &g, int bmode, int epi, bool ta, bool tb, bool b_double, hipStream_t s);

}  // namespace bh
