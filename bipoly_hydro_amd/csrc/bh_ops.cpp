// bh_ops.cpp -- dispatch of Ops to the kernels of bh_gemm.hip,
// bh_kernels.hip and bh_mixed.hip.
#include "bh_ops.h"

#include <algorithm>
#include <cstdlib>
#include <map>
#include <string>
#include <mutex>

namespace bh {

int persistent_blocks(long tiles, int blocks_per_cu)
{
    // One workgroup per tile, rounded up to a multiple of 8 for the XCD
    // chunking of BH_TILE_LOOP; the hardware keeps as many resident as fit.
    // (Capping the grid at the occupancy reported for these kernels ran the
    // DFT products up to 15 % slower.)
    (void)blocks_per_cu;
    static const bool linear = [] {
        const char *e = std::getenv("BH_TILE_ORDER");
        return e && std::string(e) == "linear";
    }();
    const long n = (tiles + 7) / 8 * 8;
    // BH_TILE_ORDER=linear: one more workgroup than a multiple of 8 turns
    // the XCD chunking off (tile t on workgroup t), for comparison
    if (linear) return (int)std::min(n + 1, 0x7ffffff9L);
    return (int)std::min(std::max(n, 8L), 0x7ffffff8L);
}

int kernel_occupancy(const void *fn, int threads, int lds)
{
    static std::mutex mu;
    static std::map<std::pair<const void *, int>, int> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto key = std::make_pair(fn, lds * 2048 + threads);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, fn, threads, lds) != hipSuccess) b = 0;
    cache[key] = b;
    return b;
}

void launch_op(const Op &op, hipStream_t s)
{
    if (op.kind == OP_MARK) return;
    if (gemm_launch(op, s)) return;
    if (kernels_launch(op, s)) return;
    if (mixed_launch(op, s)) return;
    if (density_launch(op, s)) return;
    if (spectral_launch(op, s)) return;
    if (spec64_launch(op, s)) return;
    (void)dft_launch(op, s);
}

void launch_op_grouped(int kind, int variant, const void *d_args, const GroupSched &gs, int nblocks, hipStream_t s)
{
    if (kind == OP_MARK || nblocks <= 0 || gs.total <= 0) return;
    if (gemm_launch_grouped(kind, variant, d_args, gs, nblocks, s)) return;
    if (kernels_launch_grouped(kind, variant, d_args, gs, nblocks, s)) return;
    if (mixed_launch_grouped(kind, variant, d_args, gs, nblocks, s)) return;
    if (density_launch_grouped(kind, variant, d_args, gs, nblocks, s)) return;
    if (spectral_launch_grouped(kind, variant, d_args, gs, nblocks, s)) return;
    if (spec64_launch_grouped(kind, variant, d_args, gs, nblocks, s)) return;
    (void)dft_launch_grouped(kind, variant, d_args, gs, nblocks, s);
}

int op_blocks_per_cu(int kind, int variant)
{
    int b = 0;
    if (gemm_occupancy(kind, variant, &b)) return b;
    if (kernels_occupancy(kind, variant, &b)) return b;
    if (mixed_occupancy(kind, variant, &b)) return b;
    if (density_occupancy(kind, variant, &b)) return b;
    if (dft_occupancy(kind, variant, &b)) return b;
    if (spectral_occupancy(kind, variant, &b)) return b;
    if (spec64_occupancy(kind, variant, &b)) return b;
    return 0;
}

}  // namespace bh
