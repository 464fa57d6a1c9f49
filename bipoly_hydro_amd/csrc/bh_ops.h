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

// Persistent loop over tiles 0..total-1.  Workgroups are dispatched to the
// 8 XCDs round-robin (workgroup b runs on XCD b % 8), and each XCD has its
// own L2.  When the grid is a multiple of 8, XCD x takes the contiguous
// chunk [x c, (x+1) c) of tiles, c = ceil(total / 8), so that neighbouring
// tiles, which share operands (a GEMM's column panel, a mode's factors, a
// sweep's halo rows), meet in one L2.
#define BH_TILE_LOOP(total, T_VAR, ...)                                                         \
    {                                                                                           \
        const int n_ = gridDim.x;                                                               \
        const int b_ = blockIdx.x;                                                              \
        int t0_ = b_;                                                                           \
        int t1_ = (total);                                                                      \
        int st_ = n_;                                                                           \
        if (n_ >= 8 && (n_ & 7) == 0) {                                                         \
            const int c_ = ((total) + 7) >> 3;                                                  \
            const int x_ = b_ & 7;                                                              \
            t0_ = x_ * c_ + (b_ >> 3);                                                          \
            t1_ = min((total), (x_ + 1) * c_);                                                  \
            st_ = n_ >> 3;                                                                      \
        }                                                                                       \
        for (int T_VAR = t0_; T_VAR < t1_; T_VAR += st_) {                                      \
            __VA_ARGS__;                                                                        \
            __syncthreads();                                                                    \
        }                                                                                       \
    }

// Single-problem form: tile t of an (nx, ny, nz) grid.
#define BH_SINGLE_LOOP(a, nx, ny, total, ...)                                                   \
    BH_TILE_LOOP(total, t_, {                                                                   \
        const int bx = t_ % (nx);                                                               \
        const int by = (t_ / (nx)) % (ny);                                                      \
        const int bz = t_ / ((nx) * (ny));                                                      \
        __VA_ARGS__;                                                                            \
    })

// Grouped form: the body sees `ga` (this group's arguments, a register
// copy) and bx, by, bz.
#define BH_GROUP_LOOP(ARGS_T, args, gs, ...)                                                    \
    BH_TILE_LOOP((gs).total, t_, {                                                              \
        const int g_ = bh_find_group((gs).grp, (gs).ngroups, t_);                               \
        const GroupTile e_ = (gs).grp[g_];                                                      \
        const int l_ = t_ - e_.tile0;                                                           \
        const int bx = l_ % e_.nx;                                                              \
        const int by = (l_ / e_.nx) % e_.ny;                                                    \
        const int bz = l_ / (e_.nx * e_.ny);                                                    \
        const ARGS_T ga = (args)[g_];                                                           \
        __VA_ARGS__;                                                                            \
    })
#endif

enum OpKind {
    OP_MARK = 0,      // stage boundary (timing event), no kernel
    OP_GEMM,          // bh_gemm.hip
    OP_FILL,          // bh_kernels.hip
    OP_TRANSPOSE,
    OP_SWEEP,
    OP_AXIS,
    OP_PROJ,
    OP_CONTRACT_SIDE,
    OP_TOPSYM_SCALE,
    OP_CONTRACT_TOPSYM,
    OP_TOPSYM_REDUCE,
    OP_FACE_MODES,
    OP_RESIDUAL,      // bh_mixed.hip
    OP_UPDATE,
    OP_TRIDIAG,
    OP_DENSITY,       // bh_density.hip
    OP_DFT,           // bh_dft.hip
    OP_SPECTRAL,      // bh_spectral.hip
    OP_RES64,         // bh_spec64.hip
    OP_THOMAS64,
    OP_UPDATE64,
    OP_NKINDS
};

// OP_MARK variants past the stage marks (0-7): ops between MARK_FORK and
// MARK_SIDE_END run on the plan's side stream, ordered after what precedes
// the fork; MARK_JOIN makes the main stream wait for them.  Grouped
// (batched) execution runs everything in order on one stream.
enum { MARK_FORK = 100, MARK_SIDE_END = 101, MARK_JOIN = 102 };

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

// Workgroups for a launch of `tiles` tiles through BH_TILE_LOOP.
int persistent_blocks(long tiles, int blocks_per_cu);
// Resident workgroups per CU of a kernel (cached per kernel and LDS size).
int kernel_occupancy(const void *fn, int threads, int lds);
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
bool dft_launch(const Op &op, hipStream_t s);
bool dft_launch_grouped(int kind, int variant, const void *d_args, const GroupSched &gs, int nblocks, hipStream_t s);
bool dft_occupancy(int kind, int variant, int *blocks);
bool spectral_launch(const Op &op, hipStream_t s);
bool spectral_launch_grouped(int kind, int variant, const void *d_args, const GroupSched &gs, int nblocks,
                             hipStream_t s);
bool spectral_occupancy(int kind, int variant, int *blocks);
bool spec64_launch(const Op &op, hipStream_t s);
bool spec64_launch_grouped(int kind, int variant, const void *d_args, const GroupSched &gs, int nblocks,
                           hipStream_t s);
bool spec64_occupancy(int kind, int variant, int *blocks);

}  // namespace bh
