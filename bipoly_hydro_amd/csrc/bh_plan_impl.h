// bh_plan_impl.h -- plan internals shared by the single-grid solve, the
// batch executor and the serving loop.
#pragma once
#include "../../include/bh_potential.h"
#include "bh_hostcopy.h"
#include "bh_internal.h"
#include "bh_ops.h"

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

namespace bh {

// Error reporting for the C ABI: the message of the last failure on this thread.
int fail(int code, const std::string &msg);
const char *last_error();

#define HIPCHK(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return ::bh::fail(BH_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Device allocations owned by one object, freed with it.
struct DevMem {
    std::vector<void *> bufs;
    size_t bytes = 0;
    DevMem() = default;
    DevMem(const DevMem &) = delete;
    DevMem &operator=(const DevMem &) = delete;
    ~DevMem();
    int alloc_bytes(void **out, size_t nbytes);
    template <class T>
    int alloc(T **out, size_t n)
    {
        void *p = nullptr;
        int rc = alloc_bytes(&p, n * sizeof(T));
        *out = (T *)p;
        return rc;
    }
};

// Host-side geometry and coefficients of one grid (setup_frac.F90:374-463,
// potsetup.F90:113-252, helmadi.F90:124-129).
struct HostGeo {
    bh_grid_desc d;
    Geo g;
    std::vector<double> r_g, rhf_g, rinv_g, rhfinv_g, zhf_g;  // 1-based
    std::vector<double> alphar, betar;                         // [numr + 1]
    std::vector<double> brb, elzb;                             // [numphi][numr + 1]
    std::vector<double> dt[2];                                 // cold, warm schedules
};
// Fills defaults (zeros take the reference's values) and validates.
int normalize_desc(const bh_grid_desc *in, int precision, bh_grid_desc *out);
// The mixed-precision solve (BH_PREC_MIXED or its FP32-table variant).
inline bool is_mixed(int precision) { return precision == BH_PREC_MIXED || precision == BH_PREC_MIXED_TABLE32; }
// The FP64 spectral solve (BH_PREC_SPECTRAL64).
inline bool is_spec64(int precision) { return precision == BH_PREC_SPECTRAL64; }
inline bool valid_precision(int precision)
{
    return precision == BH_PREC_FP64 || is_mixed(precision) || is_spec64(precision);
}
void host_geometry(const bh_grid_desc &d, HostGeo *h);

// Green's tables of one (numr, numz, dr, mmax) geometry on one device
// (tm.F90, sm.F90, bm.F90): shared by every plan of that geometry.
struct GeoTables {
    DevMem mem;
    double *T = nullptr;         // top (and mirrored bottom) table, blocked symmetric [m][pair][kz][16][16]
    float *T32 = nullptr;        // ... the same in FP32 (BH_PREC_MIXED_TABLE32 plans; T is then null)
    double *S = nullptr;         // side table, Toeplitz in z [m][d][jr]
    double *H = nullptr;         // isym = 2, 3: side image table, Hankel in z [m][kg + kl - 4][jr]
    double *sqrt_rhf = nullptr;  // sqrt(rhf) [numr + 1], 1-based J
};

// Transform matrices and ADI operator data of one (grid, schedule,
// precision) on one device: shared by every plan of that configuration.
struct ModeOps {
    DevMem mem;
    Geo g;
    std::vector<double> dt[2];
    double *Dc = nullptr, *Ds = nullptr, *Ec = nullptr, *Es = nullptr, *Pf = nullptr, *besc = nullptr, *bess = nullptr;
    double *alphar = nullptr, *betar = nullptr, *elzb = nullptr;
    // real DFT operands in MFMA fragment order (numphi <= DFT_MAX_N)
    float *dfwd32 = nullptr;
    double *dfwd64 = nullptr, *dinv64 = nullptr;
    // FP64 Thomas factors per pseudo-timestep
    double *rc_alpha[2] = {}, *rc_ibet[2] = {}, *rc_gback[2] = {};
    double *zc_alpha[2] = {}, *zc_ibet[2] = {}, *zc_gback[2] = {};
    // mixed precision
    float *Dc32 = nullptr, *Ds32 = nullptr, *S32 = nullptr, *mu32[2] = {}, *Mr[2] = {};
    float *tri_ibet = nullptr, *tri_gam = nullptr;
    double *brb_d = nullptr;
    // fused per-mode spectral step (then Mr holds split bf16 fragments)
    bool spec_fused = false;
    float *sfb = nullptr, *sfa = nullptr;
    float *sfb16 = nullptr, *sfa16 = nullptr;  // nzi >= 64: split-bf16 fragments (same sizes)
    // FP64 spectral solve: sine basis S [a][k] (symmetric), lambda_k and
    // mu_k per sine mode k, M_r(m)^T per azimuthal mode number m = 0..N/2
    // ([m][j'][j]) for both schedules
    double *S64 = nullptr, *lam64 = nullptr, *mu64[2] = {}, *MrT64[2] = {};
};

// Shared resources from the per-process cache (created on first use; kept
// while any plan holds them).  `created` reports whether this call built them.
int get_tables(const HostGeo &h, int precision, int device, std::shared_ptr<const GeoTables> *out, bool *created);
int get_mode_ops(const HostGeo &h, int precision, int device, std::shared_ptr<const ModeOps> *out, bool *created);
// Bytes of device memory a plan would add: its own buffers, and the shared
// resources not yet in the cache.
size_t plan_private_bytes(const bh_grid_desc &d, int precision);
size_t plan_shared_bytes(const bh_grid_desc &d, int precision);
size_t plan_new_shared_bytes(const bh_grid_desc &d, int precision, int device);

struct TimingSet {
    hipEvent_t ev[7];
};

}  // namespace bh

struct bh_plan {
    bh_grid_desc d;
    int precision = BH_PREC_FP64;
    size_t budget = 0;
    bh::Geo g;
    int device = 0;
    std::shared_ptr<const bh::GeoTables> tab;
    std::shared_ptr<const bh::ModeOps> ops;
    bh::DevMem mem;  // this plan's own buffers
    size_t bytes = 0;  // own + shared
    // per-solve state
    double *RJ = nullptr, *RK = nullptr, *XJ = nullptr, *XK = nullptr, *proj = nullptr;
    double *Stb = nullptr, *Ss = nullptr, *FF = nullptr;
    double *Pt = nullptr, *Ptb = nullptr;  // scaled projections, contraction partials (top/bottom)
    long rh_pitch = 0;  // floats per mode row of RH32: numr x numz rounded up to 32 (128-byte rows)
    float *RH32 = nullptr, *r32 = nullptr, *RT = nullptr, *WT = nullptr, *CT = nullptr, *DT = nullptr;
    float *D1 = nullptr, *Wf = nullptr;
    float *D2 = nullptr;  // sum of the refinement corrections (mixed, fused spectral step, passes >= 2)
    // FP64 spectral solve: r0 then delta, S r0, its product with M_r^T, each
    // compact [L][k or K][J-2]
    double *R64 = nullptr, *T64 = nullptr, *U64 = nullptr;
    // FP64 path: the interior between ADI half-steps in the tiled layouts of
    // LineArgs (TK after the r half-step, TJ after the z half-step)
    double *TJ = nullptr, *TK = nullptr;
    double *rho_in = nullptr, *pot_out = nullptr;  // host-path staging
    std::unique_ptr<bh::HostStager> stager;        // host-path copies (created on first use)
    // the source step (bh_hydro.hip): r, rhf, rinv, rhfinv [4][numr + 1]
    // (1-based J), and the centre-of-mass arrays cylin_a_com, cylin_v_com
    // [2][numphi][3]
    double *src_grid = nullptr, *src_com = nullptr;
    int passes = 2;                                // FP64 refinement passes (mixed; desc.refine_passes)
    int tile_dft = 2, tile_spec = 2, tile_dstf = 2, tile_dsti = 2;  // GEMM tile shapes, measured best
    int order_dft = 0, order_spec = 0;             // GEMM tile numbering (GemmArgs::col_fastest)
    bool use_dft = true;                           // panel DFT kernel (bh_dft.hip) when numphi allows
    // the forward DFT of rho runs on a side stream beside the projections
    // and the boundary (independent until the faces are scattered)
    bool overlap = true;
    int cus = 256;      // compute units of the device
    int side_cus = 0;   // workgroups of the side-stream DFT (0: two per CU)
    hipStream_t side = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    // the stream of the plan's last solve (every solve ends there, the side
    // stream joined): a solve on another stream first waits for an event
    // recorded on it then, and bh_set_potential synchronises with one, so
    // that calls on one plan never overlap on the device.  Solves on one
    // stream are ordered by the stream and record nothing.
    hipEvent_t done_ev = nullptr;
    hipStream_t last_stream = nullptr;
    bool has_last = false;
    bool has_solution = false;
    // mixed: the interior of the warm-start state XJ is D1 + (D2, or Wf at
    // one pass) -- a cold solve through the single-plan entry points leaves
    // it there instead of writing it; XJ's faces are current.  The next warm
    // solve's first residual writes it out (or materialize_state does).
    bool lazy_state = false;
    // the same on the device, for solves captured into graphs and replayed
    // (mixed plans that may keep the state so): [0] set by the cold inverse
    // that keeps it, [1] its value when the current solve began (the face
    // kernel moves [0] to [1] and clears [0]); a warm first residual writes
    // the state out when [1] is set
    float *lazy_flag = nullptr;
    // the faces of the mode rows no boundary mode reaches are zero in XJ
    // (as every solve leaves them; an import of the potential need not):
    // single-plan mixed solves then form the faces of the other rows only
    bool faces_clean = true;
    // owned by a batch: single-plan solves of it keep the state whole, so
    // that the batch's (possibly captured and replayed) launches never meet
    // a state left as corrections
    bool in_batch = false;
    // timing (HIP events at the stage marks of each timed solve)
    bool timing = false;
    std::vector<bh::TimingSet> tsets;
    int tcount = 0;
    double last_ms[6] = {0, 0, 0, 0, 0, 0};
    bool have_timing = false;
    ~bh_plan()
    {
        if (side) (void)hipStreamDestroy(side);
        if (fork_ev) (void)hipEventDestroy(fork_ev);
        if (join_ev) (void)hipEventDestroy(join_ev);
        if (done_ev) (void)hipEventDestroy(done_ev);
    }
};

namespace bh {
// Create a plan (validated descriptor).  Set-up work runs on its own stream.
int plan_create(const bh_grid_desc &d, int precision, size_t budget, bh_plan **out);
// The solve of one grid as Ops, with stage marks 0-4 (start, after the
// forward DFT, after the boundary, after the ADI, end) and 5-6 around the
// top/bottom boundary contraction.
// allow_lazy (the single-plan entry points): a cold mixed solve at N = 512
// may leave its warm-start state as its two corrections (lazy_cold).  A
// warm solve of a plan whose state is so left writes it out as it forms
// its first residual, whatever allow_lazy.
void build_ops(const bh_plan *p, const double *d_rho, double *d_pot, int tstep, std::vector<Op> *out,
               bool allow_lazy = false);
bool lazy_cold(const bh_plan *p, int tstep);
// Launch the Ops of one solve on s (records timing events when enabled).
int enqueue_ops(bh_plan *p, const std::vector<Op> &ops, hipStream_t s);
// Forward real DFT along phi (FP64) of the numphi rows of B (Mcols columns,
// row pitch ldb) into C (row pitch ldc), in helmadi's mode order; mask
// zeroes the rhop layers; max_blocks caps the persistent grid (0: none).
void dft_fwd64_ops(const bh_plan *p, const double *B, long ldb, int Mcols, double *C, long ldc, bool mask,
                   std::vector<Op> *out, int max_blocks = 0);
// Wait until the plan's last enqueued solve has finished.
int wait_plan_idle(bh_plan *p);
// Write out a warm-start state a cold solve left as its two corrections
// (enqueued on s, after the plan's last solve); nothing when it is whole.
int materialize_state(bh_plan *p, hipStream_t s);
}  // namespace bh
