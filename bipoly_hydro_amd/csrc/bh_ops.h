// bh_ops.h -- the solve as a list of kernel operations, and grouped launches.
//
// A solve is recorded as a short list of Op records (kernel kind, template
// variant, launch grid in tiles, argument block).  One grid's solve launches
// each Op on its own; a batch of grids launches, for every position in the
// list, one persistent grouped kernel per distinct (kind, variant) that runs
// the tiles of all the grids' Ops at that position.
#pragma once
#include <hip/hip_runtime.h>

#include <cstring>
#include <type_traits>

namespace bh {

// Group g of a grouped launch owns tiles [tile0, tile0 + nx*ny*nz); its local
// tile l decodes to (l % nx, (l / nx) % ny, l / (nx*ny)).
struct GroupTile {
    int tile0, nx, ny, nz;
};
struct GroupSched {
    const GroupTile *grp;
    int ngroups, total;
};

#if defined(__HIPCC__)
__device__ inline int bh_find_group(const GroupTile *grp, int ng, int t)
{
    int lo = 0, hi = ng - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (grp[mid].tile0 <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Persistent loop over the tiles of every group; BODY sees `ga` (this
// group's arguments, a register copy) and bx, by, bz.
#define BH_GROUP_LOOP(ARGS_T, args, gs, BODY)                                       \
    for (int t_ = blockIdx.x; t_ < (gs).total; t_ += gridDim.x) {                   \
        const int g_ = bh_find_group((gs).grp, (gs).ngroups, t_);                   \
        const GroupTile e_ = (gs).grp[g_];                                          \
        const int l_ = t_ - e_.tile0;                                               \
        const int bx = l_ % e_.nx, by = (l_ / e_.nx) % e_.ny, bz = l_ / (e_.nx * e_.ny); \
        const ARGS_T ga = (args)[g_];                                               \
        BODY;                                                                       \
        __syncthreads();                                                            \
    }
#endif

enum OpKind {
    OP_MARK = 0,      // stage boundary (timing event), no kernel
    OP_GEMM,          // bh_gemm.hip
    OP_FILL,          // bh_kernels.hip
    OP_TRANSPOSE,
    OP_SWEEP,
    OP_AXIS,
    OP_CONTRACT_TB,
    OP_CONTRACT_SIDE,
    OP_REBUILD,
    OP_SCATTER,
    OP_RESIDUAL,      // bh_mixed.hip
    OP_UPDATE,
    OP_TRIDIAG,
    OP_DENSITY,       // bh_density.hip
    OP_NKINDS
};

constexpr int OP_ARGS_MAX = 256;

struct Op {
    int kind = OP_MARK, variant = 0;
    int nx = 1, ny = 1, nz = 1;  // launch grid in tiles
    int argsz = 0;
    alignas(16) unsigned char args[OP_ARGS_MAX];

    template <class A>
    static Op make(int kind, int variant, const A &a, int nx, int ny, int nz)
    {
        static_assert(sizeof(A) <= OP_ARGS_MAX, "argument block too large");
        static_assert(std::is_trivially_copyable<A>::value, "argument block must be trivially copyable");
        Op o;
        o.kind = kind;
        o.variant = variant;
        o.nx = nx;
        o.ny = ny;
        o.nz = nz;
        o.argsz = (int)sizeof(A);
        std::memcpy(o.args, &a, sizeof(A));
        return o;
    }
    static Op mark(int index)
    {
        Op o;
        o.kind = OP_MARK;
        o.variant = index;
        return o;
    }
    template <class A>
    const A &as() const { return *reinterpret_cast<const A *>(args); }
    long tiles() const { return (long)nx * ny * nz; }
};

// One Op, alone, on stream s.
void launch_op(const Op &op, hipStream_t s);
// The Ops of one (kind, variant) for several problems: d_args holds their
// argument blocks (stride = the kind's argument size) in device memory.
void launch_op_grouped(int kind, int variant, const void *d_args, const GroupSched &gs, int nblocks,
                       hipStream_t s);
// Resident workgroups per CU of the grouped kernel (0 when unknown).
int op_blocks_per_cu(int kind, int variant);

// per-file dispatchers (return false when the kind is not theirs)
bool gemm_launch(const Op &op, hipStream_t s);
bool gemm_launch_grouped(int kind, int variant, const void *d_args, const GroupSched &gs, int nblocks, hipStream_t s);
bool gemm_occupancy(int kind, int variant, int *blocks);
bool kernels_launch(const Op &op, hipStream_t s);
bool kernels_launch_grouped(int kind, int variant, const void *d_args, const GroupSched &gs, int nblocks,
                            hipStream_t s);
bool kernels_occupancy(int kind, int variant, int *blocks);
bool mixed_launch(const Op &op, hipStream_t s);
bool mixed_launch_grouped(int kind, int variant, const void *d_args, const GroupSched &gs, int nblocks,
                          hipStream_t s);
bool mixed_occupancy(int kind, int variant, int *blocks);
bool density_launch(const Op &op, hipStream_t s);
bool density_launch_grouped(int kind, int variant, const void *d_args, const GroupSched &gs, int nblocks,
                            hipStream_t s);
bool density_occupancy(int kind, int variant, int *blocks);

}  // namespace bh
