// bh_ops.cpp -- dispatch of Ops to the kernels of bh_gemm.hip,
// bh_kernels.hip and bh_mixed.hip.
#include "bh_ops.h"

namespace bh {

void launch_op(const Op &op, hipStream_t s)
{
    if (op.kind == OP_MARK) return;
    if (gemm_launch(op, s)) return;
    if (kernels_launch(op, s)) return;
    if (mixed_launch(op, s)) return;
    (void)density_launch(op, s);
}

void launch_op_grouped(int kind, int variant, const void *d_args, const GroupSched &gs, int nblocks, hipStream_t s)
{
    if (kind == OP_MARK || nblocks <= 0 || gs.total <= 0) return;
    if (gemm_launch_grouped(kind, variant, d_args, gs, nblocks, s)) return;
    if (kernels_launch_grouped(kind, variant, d_args, gs, nblocks, s)) return;
    if (mixed_launch_grouped(kind, variant, d_args, gs, nblocks, s)) return;
    (void)density_launch_grouped(kind, variant, d_args, gs, nblocks, s);
}

int op_blocks_per_cu(int kind, int variant)
{
    int b = 0;
    if (gemm_occupancy(kind, variant, &b)) return b;
    if (kernels_occupancy(kind, variant, &b)) return b;
    if (mixed_occupancy(kind, variant, &b)) return b;
    if (density_occupancy(kind, variant, &b)) return b;
    return 0;
}

}  // namespace bh
