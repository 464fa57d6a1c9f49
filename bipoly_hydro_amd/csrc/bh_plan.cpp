// bh_plan.cpp -- plan construction (the potsetup counterpart) and the solve
// sequence (the potential_solver counterpart) behind the C ABI of
// include/bh_potential.h.
#include "bh_plan_impl.h"
#include "bh_dft.h"
#include "bh_gemm.h"
#include "bh_mixed.h"
#include "bh_spec64.h"
#include "bh_spectral.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

using namespace bh;

// ---------------------------------------------------------------------------
// plan construction
// ---------------------------------------------------------------------------
namespace bh {

int plan_create(const bh_grid_desc &d, int precision, size_t budget, bh_plan **out)
{
    *out = nullptr;
    int device = 0;
    HIPCHK(hipGetDevice(&device));
    const size_t need = plan_private_bytes(d, precision) + plan_shared_bytes(d, precision);
    if (budget && need > budget)
        return fail(BH_E_NOMEM, "plan needs " + std::to_string(need) + " bytes of device memory, over the budget of " +
                                    std::to_string(budget));
    HostGeo h;
    host_geometry(d, &h);
    std::unique_ptr<bh_plan> p(new bh_plan());
    p->d = d;
    p->precision = precision;
    p->budget = budget;
    p->g = h.g;
    p->device = device;
    p->passes = d.refine_passes;
    // GEMM tile shapes, overridable for tuning sweeps
    if (const char *e = std::getenv("BH_GEMM_TILE_DFT")) p->tile_dft = std::atoi(e);
    if (const char *e = std::getenv("BH_GEMM_TILE_SPEC")) p->tile_spec = std::atoi(e);
    if (const char *e = std::getenv("BH_GEMM_TILE_DSTF")) p->tile_dstf = std::atoi(e);
    if (const char *e = std::getenv("BH_GEMM_TILE_DSTI")) p->tile_dsti = std::atoi(e);
    if (const char *e = std::getenv("BH_GEMM_ORDER_DFT")) p->order_dft = std::atoi(e);
    if (const char *e = std::getenv("BH_GEMM_ORDER_SPEC")) p->order_spec = std::atoi(e);
    if (const char *e = std::getenv("BH_DFT_KERNEL")) p->use_dft = std::atoi(e) != 0;
    if (const char *e = std::getenv("BH_OVERLAP")) p->overlap = std::atoi(e) != 0;
    if (const char *e = std::getenv("BH_SIDE_CUS")) p->side_cus = std::atoi(e);
    bool created;
    int rc;
    if ((rc = get_tables(h, precision, device, &p->tab, &created)) != BH_OK) return rc;
    if ((rc = get_mode_ops(h, precision, device, &p->ops, &created)) != BH_OK) return rc;

    const Geo &g = p->g;
    const int numr = g.numr, numz = g.numz, N = g.numphi, mmax = g.mmax, nri = g.nri, nzi = g.nzi;
    const size_t M = (size_t)numr * numz;
    const int nf = 2 * nri + nzi;
    DevMem &m = p->mem;
#define ALLOC(ptr, n) \
    if ((rc = m.alloc(&(ptr), (size_t)(n))) != BH_OK) return rc
    ALLOC(p->XJ, M * N);
    ALLOC(p->XK, M * N);
    ALLOC(p->proj, M * (2 * mmax - 1));
    ALLOC(p->Stb, (size_t)mmax * 4 * nri);
    {
        const size_t NB = (nri + 15) / 16;
        if (!tb_blocked(numr)) ALLOC(p->Pt, (size_t)mmax * nzi * 16 * NB * 4);  // the blocked form scales itself
        const size_t parts = tb_blocked(numr) ? tb_chunks(numz) : topsym_splits(numz) * ((NB + 1) / 2);
        ALLOC(p->Ptb, (size_t)mmax * parts * 16 * NB * 4);
    }
    ALLOC(p->Ss, (size_t)side_splits(numz) * mmax * 2 * nzi);
    ALLOC(p->FF, (size_t)N * nf);
    if (is_mixed(precision)) ALLOC(p->lazy_flag, 64);
    ALLOC(p->src_grid, 4 * (size_t)(numr + 1));
    ALLOC(p->src_com, 6 * (size_t)N);
    const bool with_d2 = is_mixed(precision) && p->ops->spec_fused && p->passes >= 2;
    if (is_mixed(precision)) {
        const size_t PI = (size_t)N * nri * nzi;
        // rho-hat's mode rows on 128-byte boundaries: the forward DFT writes
        // them in column panels, and aligned rows store ~40 % faster
        p->rh_pitch = (long)(M + 31) / 32 * 32;
        ALLOC(p->RH32, (size_t)p->rh_pitch * N);
        ALLOC(p->r32, PI);
        if (!p->ops->spec_fused) {  // spectral blocks between the separate kernels
            ALLOC(p->RT, PI);
            ALLOC(p->CT, PI);
            ALLOC(p->DT, 2 * PI);  // DT then WT, so that one product maps both back
            p->WT = p->DT + PI;
        }
        // D1 then Wf (then D2, the summed refinement corrections); faces zero
        ALLOC(p->D1, (with_d2 ? 3 : 2) * M * N);
        p->Wf = p->D1 + M * N;
        if (with_d2) p->D2 = p->Wf + M * N;
    } else if (is_spec64(precision)) {
        const size_t PI = (size_t)N * nri * nzi;
        ALLOC(p->RJ, M * N);
        ALLOC(p->R64, PI);
        ALLOC(p->T64, PI);
        ALLOC(p->U64, PI);
    } else {
        ALLOC(p->RJ, M * N);
        ALLOC(p->RK, (size_t)N * tiled_stride(nri, nzi));
        ALLOC(p->TJ, (size_t)N * tiled_stride(nzi, nri));
        ALLOC(p->TK, (size_t)N * tiled_stride(nri, nzi));
    }
#undef ALLOC
    p->bytes = p->mem.bytes + p->tab->mem.bytes + p->ops->mem.bytes;
    HIPCHK(hipDeviceGetAttribute(&p->cus, hipDeviceAttributeMultiprocessorCount, device));
    HIPCHK(hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking));
    // the fork and join events order the side stream against the solve's
    // stream on this device only: no system-scope fence (its cache
    // writeback and invalidation stalled the recording stream ~20 us)
    HIPCHK(hipEventCreateWithFlags(&p->fork_ev, hipEventDisableTiming | hipEventDisableSystemFence));
    HIPCHK(hipEventCreateWithFlags(&p->join_ev, hipEventDisableTiming | hipEventDisableSystemFence));
    HIPCHK(hipEventCreateWithFlags(&p->done_ev, hipEventDisableTiming));
    hipStream_t s;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    std::vector<double> sg(4 * (size_t)(numr + 1), 0.0);
    for (int J = 1; J <= numr; ++J) {
        sg[J] = h.r_g[J];
        sg[(numr + 1) + J] = h.rhf_g[J];
        sg[2 * (numr + 1) + J] = h.rinv_g[J];
        sg[3 * (numr + 1) + J] = h.rhfinv_g[J];
    }
    HIPCHK(hipMemcpyAsync(p->src_grid, sg.data(), sg.size() * sizeof(double), hipMemcpyHostToDevice, s));
    launch_fill(p->src_com, 6 * (size_t)N, 0.0, s);
    launch_fill(p->XJ, M * N, 0.0, s);
    launch_fill(p->XK, M * N, 0.0, s);
    launch_fill(p->FF, (size_t)N * nf, 0.0, s);  // rows the face kernel may skip stay zero
    if (p->lazy_flag) launch_fillf(p->lazy_flag, 64, 0.f, s);
    if (is_mixed(precision)) {
        launch_fillf(p->D1, (with_d2 ? 3 : 2) * M * N, 0.f, s);
    }
    hipError_t e = hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return fail(BH_E_DEVICE, std::string("plan setup: ") + hipGetErrorString(e));
    *out = p.release();
    return BH_OK;
}

// ---------------------------------------------------------------------------
// the solve as Ops
// ---------------------------------------------------------------------------
namespace {

// Workgroups of a DFT that runs on the side stream: two per CU (a multiple
// of 8 for the XCD-chunked panel order; the kernel fits four per CU, so the
// main stream's kernels keep room beside it); 0 (one per panel) without the
// side stream.
int side_blocks(const bh_plan *p)
{
    if (!p->overlap) return 0;
    return p->side_cus > 0 ? p->side_cus : std::max(8, 2 * p->cus / 8 * 8);
}

}  // namespace

// Forward real DFT along phi of the N rows of B (columns Mcols), written in
// helmadi's mode order (helmadi.F90:131-156) to C; cosine and sine halves.
void dft_fwd64_ops(const bh_plan *p, const double *B, long ldb, int Mcols, double *C, long ldc, bool mask,
                   std::vector<Op> *out, int max_blocks)
{
    const ModeOps &o = *p->ops;
    const int N = p->g.numphi, nh = N / 2 + 1, nsn = N / 2 - 1;
    if (o.dfwd64 && p->use_dft) {
        DftArgs d{B, ldb, C, ldc, o.dfwd64, N, Mcols, mask ? p->g.numr : 0, mask ? p->g.numz : 0};
        d.max_blocks = max_blocks;
        d.mask_keep_bottom = p->g.isym != 1;
        out->push_back(dft_op(d, DFT_FWD_F64));
        return;
    }
    GemmArgs a{};
    a.A = o.Dc; a.lda = nh; a.B = B; a.ldb = ldb; a.C = C; a.ldc = ldc;
    a.R = nh; a.M = Mcols; a.Kd = nh; a.batch = 1; a.nfold = N;
    a.mask_numr = p->g.numr; a.mask_numz = p->g.numz;
    a.mask_keep_bottom = p->g.isym != 1;
    a.tile = Mcols >= 4096 ? p->tile_dft : 2;
    a.col_fastest = p->order_dft;
    out->push_back(gemm_op(a, gemm_inst64(B_FOLD_PLUS, mask ? E_MASK : E_PLAIN)));
    a.A = o.Ds; a.lda = nsn; a.C = C + (long)nh * ldc; a.R = nsn; a.Kd = nsn;
    out->push_back(gemm_op(a, gemm_inst64(B_FOLD_MINUS, mask ? E_MASK : E_PLAIN)));
}

namespace {

// Inverse real DFT with the 2/N normalisation (helmadi.F90:357-392) of the
// N mode rows of X into N phi rows of out.
void dft_inv64_ops(const bh_plan *p, const double *X, long ldx, int Mcols, double *out, long ldo,
                   std::vector<Op> *ops)
{
    const ModeOps &o = *p->ops;
    const int N = p->g.numphi, nh = N / 2 + 1, nsn = N / 2 - 1;
    if (o.dinv64 && p->use_dft) {
        DftArgs d{X, ldx, out, ldo, o.dinv64, N, Mcols, 0, 0};
        ops->push_back(dft_op(d, DFT_INV_F64));
        return;
    }
    GemmArgs a{};
    a.A = o.Ec; a.lda = nh; a.B = X; a.ldb = ldx; a.C = out; a.ldc = ldo;
    a.R = nh; a.M = Mcols; a.Kd = nh; a.batch = 1; a.nfold = N;
    a.tile = Mcols >= 4096 ? p->tile_dft : 2;
    a.col_fastest = p->order_dft;
    ops->push_back(gemm_op(a, G64_PLAIN));
    a.A = o.Es; a.lda = nsn; a.B = X + (long)nh * ldx;
    a.R = nsn; a.Kd = nsn;
    ops->push_back(gemm_op(a, G64_INV));
}

// Boundary projections of rho onto the mmax azimuthal modes (bessel.F90:176-184, 226-235)
void proj64_ops(const bh_plan *p, const double *rho, std::vector<Op> *out)
{
    out->push_back(proj_op(rho, p->ops->Pf, p->proj, p->g.numr * p->g.numz, p->g.numphi, p->g.mmax, p->g.isym));
}

// Boundary potential (bessel.F90:186-488) -> Fourier-space faces FF, also
// placed in the faces of X1 (J-layout) and X2 (K-layout when x2_klayout,
// else J-layout).  The side contraction (short and latency-bound) runs on
// the solve's stream right after the projections, beside the side stream's
// forward DFT; the solve's stream then joins the side stream, so that the
// top/bottom contraction streams its table with HBM to itself.  (With the
// DFT beside the table stream instead the solve takes the same time: the
// boundary window is bound by its total HBM traffic.)
void boundary_ops(const bh_plan *p, double *X1, double *X2, int x2_klayout, std::vector<Op> *out,
                  const ResidualArgs *face_res = nullptr, bool reach_rows = false, float *flag = nullptr)
{
    const Geo &g = p->g;
    const int numr = g.numr, numz = g.numz, N = g.numphi, mmax = g.mmax;
    out->push_back(contract_side_op(p->tab->S, p->proj, p->Ss, numr, numz, mmax, p->tab->H));
    out->push_back(Op::mark(MARK_JOIN));
    out->push_back(Op::mark(5));
    if (!tb_blocked(numr)) out->push_back(topsym_scale_op(p->proj, p->tab->sqrt_rhf, p->Pt, numr, numz, mmax));
    out->push_back(contract_topsym_op(p->tab->T, p->tab->T32, p->proj, p->tab->sqrt_rhf, p->Pt, p->Ptb, numr, numz,
                                      mmax));
    // (no reduction launch: the face kernel sums the contraction's
    // partials itself, in the reduction's order)
    out->push_back(Op::mark(6));
    // factor of bessel.F90:170-174 (doubled for isym = 3)
    const double factor = g.isym == 3 ? -2.0 * g.dr * g.dz * g.dphi : -g.dr * g.dz * g.dphi;
    out->push_back(face_modes_op(p->Stb, p->Ss, p->ops->besc, p->ops->bess, p->FF, numr, numz, N, mmax, factor, 1,
                                 X1, X2, x2_klayout, g.isym, reach_rows));
    {
        RebuildArgs fa = out->back().as<RebuildArgs>();
        fa.flag = flag;
        fa.Ptb = p->Ptb;
        fa.sqrt_rhf = p->tab->sqrt_rhf;
        const int NB = (numr - 2 + 15) / 16;
        fa.nk = tb_blocked(numr) ? tb_chunks(numz) : topsym_splits(numz) * ((NB + 1) / 2);
        std::memcpy(out->back().args, &fa, sizeof fa);
    }
    if (face_res)
        face_modes_add_residual(&out->back(), face_res->r, face_res->rh, face_res->rh_pitch, face_res->alphar,
                                face_res->gamma, face_res->four_pi);
}

// The mixed-precision solve.  Evaluates the same truncated Peaceman-Rachford
// iteration as the FP64 path (helmadi.F90:195-355), as A x = g - M r0 (see
// bh_mixed.hip), with FP32 MFMA transforms and FP64 residuals.
// the second correction of a mixed solve: D2 with two or more passes, the w
// buffer at one pass (free once the one refinement residual has read w)
float *correction_sum(const bh_plan *p) { return p->passes >= 2 ? p->D2 : p->Wf; }

// a mixed plan whose cold solves at N = 512 may keep the state as their
// corrections (single-plan solves; not a plan a batch owns)
bool keeps_corrections(const bh_plan *p)
{
    return is_mixed(p->precision) && p->ops->spec_fused && p->ops->dinv64 && p->use_dft &&
           p->g.numphi == DFT_MAX_N && !p->in_batch && p->lazy_flag;
}

void mixed_ops(const bh_plan *p, const double *d_rho, double *d_pot, int tstep, std::vector<Op> *out,
               bool allow_lazy)
{
    const Geo &g = p->g;
    const ModeOps &o = *p->ops;
    const int numr = g.numr, numz = g.numz, N = g.numphi, nri = g.nri, nzi = g.nzi;
    const long M = (long)numr * numz;
    const long PIm = (long)nri * nzi;  // interior values per mode
    const int w = (tstep > 1) ? 1 : 0;
    out->push_back(Op::mark(0));
    // 1. rho-hat in FP32 (masked as rhop) on the side stream, on half the
    //    CUs, while the FP64 projections and the boundary (which do not need
    //    it) run on the main stream
    out->push_back(Op::mark(MARK_FORK));
    // a cold start at N = 512: the DFT also writes the residual's source
    // -4 pi rho-hat on the interior, so that after the faces only the points
    // next to a face need the stencil (done by the face kernel)
    const bool cold_split = tstep <= 1 && o.dfwd32 && p->use_dft && N == DFT_MAX_N;
    if (o.dfwd32 && p->use_dft) {
        DftArgs d{d_rho, M, p->RH32, p->rh_pitch, o.dfwd32, N, (int)M, numr, numz};
        d.max_blocks = side_blocks(p);
        if (cold_split) {
            d.rout = p->r32;
            d.rscale = -g.four_pi;
        }
        out->push_back(dft_op(d, DFT_FWD_F32));
        out->push_back(Op::mark(MARK_SIDE_END));
        proj64_ops(p, d_rho, out);
    } else {
        const int nh = N / 2 + 1, nsn = N / 2 - 1;
        GemmArgs a{};
        a.A = o.Dc32; a.lda = nh; a.B = d_rho; a.ldb = M; a.C = p->RH32; a.ldc = p->rh_pitch;
        a.R = nh; a.M = (int)M; a.Kd = nh; a.batch = 1; a.nfold = N;
        a.mask_numr = numr; a.mask_numz = numz;
        a.tile = M >= 4096 ? p->tile_dft : 2;
        a.col_fastest = p->order_dft;
        out->push_back(gemm_op(a, G32D_FOLDP_MASK));
        a.A = o.Ds32; a.lda = nsn; a.C = p->RH32 + (long)nh * p->rh_pitch; a.R = nsn; a.Kd = nsn;
        out->push_back(gemm_op(a, G32D_FOLDM_MASK));
        out->push_back(Op::mark(MARK_SIDE_END));
        proj64_ops(p, d_rho, out);
    }
    out->push_back(Op::mark(1));
    // 2. boundary faces (FP64), into both potential buffers; the starting
    //    potential is zero (cold) or the one kept from the last solve (warm).
    //    A cold start needs no zeroing: the residuals below read only the
    //    faces of Xa, and Xa's interior is rewritten by the final update.
    double *Xa = p->XJ, *Xb = p->XK;
    // 3. r0 = g - A x0: a cold start at N = 512 has -4 pi rho-hat in r from
    //    the DFT already, and the face kernel adds the face terms at the
    //    points next to a face as it forms the faces; otherwise a residual
    //    kernel after the faces
    ResidualArgs ra{};
    ra.brb = o.brb_d; ra.alphar = o.alphar; ra.betar = o.betar;
    ra.gamma = g.gamma; ra.four_pi = g.four_pi; ra.numr = numr; ra.numz = numz;
    ra.rh = p->RH32;
    ra.rh_pitch = p->rh_pitch;
    ra.X = Xa; ra.dlt = nullptr; ra.w = nullptr; ra.Xo = nullptr; ra.r = p->r32;
    ra.cold = tstep <= 1;
    // (the fused spectral path never reads Xb: its faces are only needed by
    // the separate kernels' second iterate)
    // (single-plan solves of a plan whose other rows' faces are zero: the
    // face kernel visits the rows boundary modes reach)
    boundary_ops(p, Xa, o.spec_fused ? nullptr : Xb, 0, out, cold_split ? &ra : nullptr,
                 allow_lazy && o.spec_fused && p->faces_clean,
                 keeps_corrections(p) ? p->lazy_flag : nullptr);  // joins the side stream
    out->push_back(Op::mark(2));
    if (!cold_split) {
        ResidualArgs r1 = ra;
        if (tstep > 1 && keeps_corrections(p)) {
            // when the last solve was a cold one that kept the state as its
            // delta1 + correction sum (lazy_flag, read on the device, so that
            // a replayed graph decides afresh): x0 inside is formed from
            // them and written to Xa, the same x0 as if stored whole
            r1.dlt = p->D1;
            r1.dlt2 = correction_sum(p);
            r1.Xo = Xa;
            r1.w = nullptr;
            r1.cold = 1;
            r1.lazy_flag = p->lazy_flag + 1;
        }
        out->push_back(residual_op(r1, N));
    }
    // folded sine transforms (spectral slots ordered by spec_mode):
    // forward RT[L] (nri x nzi) = r[L]^T S, inverse dst[L] interior
    // (nzi x nri) = S in[L]^T for nb consecutive modes of in and dst
    auto dst_fwd = [&](const float *r, float *dst) {
        GemmArgs a{};
        a.A = r; a.lda = nri; a.sA = PIm;
        a.B = o.S32; a.ldb = nzi; a.sB = 0;
        a.C = dst; a.ldc = nzi; a.sC = PIm;
        a.R = nri; a.M = nzi; a.Kd = nzi; a.batch = N; a.nfold = nzi;
        a.tile = p->tile_dstf;
        a.col_fastest = p->order_spec;
        out->push_back(gemm_op(a, G32_DST_FWD));
    };
    auto dst_inv = [&](const float *in, float *dst, int nb) {
        GemmArgs a{};
        a.A = o.S32; a.lda = nzi; a.sA = 0;
        a.B = in; a.ldb = nzi; a.sB = PIm;
        a.C = dst + numr + 1; a.ldc = numr; a.sC = M;
        a.R = (nzi + 1) / 2; a.M = nri; a.Kd = nzi; a.batch = nb; a.nfold = nzi;
        a.tile = p->tile_dsti;
        a.col_fastest = p->order_spec;
        out->push_back(gemm_op(a, G32_DST_INV));
    };
    const bool inv512 = o.dinv64 && p->use_dft && N == DFT_MAX_N;
    const bool cold = tstep <= 1;
    if (o.spec_fused) {
        // 4-5 in one kernel per mode, the block kept in LDS: delta1 into D1,
        // w into Wf
        SpecArgs sa{p->r32, o.sfb, o.sfa, o.Mr[w], o.mu32[w], o.tri_ibet, o.betar, o.alphar, p->D1, p->Wf,
                    numr, numz, N, 0, o.sfb16, o.sfa16};
        out->push_back(spectral_op(sa, N, nzi));
        // 6. refinement passes.  The iterate is never written: pass k forms
        //    x_k = X0 + (delta1 + Dacc) in its residual (X0's faces alone for
        //    a cold start, whose X0 is zero inside), r_k = g - w - A x_k in
        //    FP64, and adds the next correction to Dacc in FP32.  Every
        //    correction after the first is about 1e-7 of x, so their FP32 sum
        //    holds them to ~1e-14 of x.  With one pass Dacc is the w buffer,
        //    free once the one refinement residual has read w.
        float *Dacc = correction_sum(p);
        // with two or more passes (split-bf16 transforms, nzi >= 64) each
        // pass forms its residual inside the spectral kernel, in LDS: no r
        // written and read back, no residual launch.  r32 then keeps the
        // first residual r0 = g - A x0, from which a warm pass forms its
        // residual, r0 - w - A (delta1 + Dacc), without reading x0 (as a
        // cold one, whose x0 is zero inside, reads only x0's faces)
        const bool fuse = p->passes >= 2 && nzi >= 64;
        for (int it = 0; it < p->passes; ++it) {
            ra.X = Xa; ra.dlt = p->D1; ra.dlt2 = it > 0 ? Dacc : nullptr; ra.w = p->Wf; ra.Xo = nullptr;
            ra.r = p->r32; ra.cold = cold;
            if (fuse && !cold) {
                ra.r0 = p->r32;
                ra.cold = 1;
            }
            if (!fuse) out->push_back(residual_op(ra, N));
            SpecArgs sr{p->r32, o.sfb, o.sfa, nullptr, nullptr, o.tri_ibet, o.betar, o.alphar, Dacc, nullptr,
                        numr, numz, N, it > 0 ? 1 : 0, o.sfb16, o.sfa16, p->passes == 1 ? 1 : 0};
            sr.fuse_res = fuse ? 1 : 0;
            sr.ra = ra;
            out->push_back(spectral_op(sr, N, nzi));
        }
        // 7. x = X0 + (delta1 + Dacc) into Xa, the warm-start state, then the
        //    inverse DFT (FP64) and the axis ghost; at N = 512 the panel DFT
        //    does all three in one pass (in place is safe: every element of
        //    Xa is read and written by the same thread), and a lazy cold
        //    start (lazy_cold) leaves x inside as delta1 and Dacc instead
        if (inv512) {
            out->push_back(Op::mark(3));
            DftArgs d{Xa, M, d_pot, M, o.dinv64, N, (int)M, numr, numz, Dacc, Xa};
            d.xint = p->D1;
            d.warm = cold ? 0 : 1;
            // a lazy cold start keeps x inside as delta1 and Dacc, which
            // stay in place (the faces are in Xa already): x is not stored
            if (allow_lazy && lazy_cold(p, tstep) && keeps_corrections(p)) {
                d.xstore = nullptr;
                d.set_flag = p->lazy_flag;
            }
            d.axis_numr = numr;
            out->push_back(dft_op(d, DFT_INV_F64));
        } else {
            out->push_back(add_update_op(Xa, p->D1, Xa, numr, numz, N, Dacc, cold ? 1 : 0));
            out->push_back(Op::mark(3));
            dft_inv64_ops(p, Xa, M, (int)M, d_pot, M, out);
            out->push_back(axis_op(d_pot, numr, numz, N));
        }
    } else {
        // the separate transform, filter and Thomas kernels (grid sizes the
        // fused step does not take): each pass writes its iterate
        double *Xcur = Xa, *Xnext = Xb;
        dst_fwd(p->r32, p->RT);
        // 4. W = mu_z . (M_r RT), C = RT - W
        {
            GemmArgs a{};
            a.A = o.Mr[w]; a.lda = nri; a.sA = (long)nri * nri;
            a.B = p->RT; a.ldb = nzi; a.sB = PIm;
            a.C = p->WT; a.ldc = nzi; a.sC = PIm;
            a.R = nri; a.M = nzi; a.Kd = nri; a.batch = N;
            a.mu = o.mu32[w]; a.Rin = p->RT; a.C2 = p->CT;
            a.tile = p->tile_spec;
            a.col_fastest = p->order_spec;
            out->push_back(gemm_op(a, G32_FILTER));
        }
        // 5. (B_r + lambda_z) dhat = C, back to physical space; delta = S dhat
        //    and w = S W in one launch: WT follows DT and Wf follows D1
        out->push_back(tridiag_shift_op(p->CT, p->DT, o.tri_ibet, o.tri_gam, o.betar, nri, nzi, N));
        dst_inv(p->DT, p->D1, 2 * N);
        // 6. refinement: x += delta, r = g - w - A x (FP64), solve again
        for (int it = 0; it < p->passes; ++it) {
            ra.X = Xcur; ra.dlt = p->D1; ra.dlt2 = nullptr; ra.w = p->Wf; ra.Xo = Xnext; ra.r = p->r32;
            ra.cold = (cold && it == 0);  // Xcur = Xa until the first pass writes Xnext
            out->push_back(residual_op(ra, N));
            dst_fwd(p->r32, p->RT);
            out->push_back(tridiag_shift_op(p->RT, p->DT, o.tri_ibet, o.tri_gam, o.betar, nri, nzi, N));
            dst_inv(p->DT, p->D1, N);
            std::swap(Xcur, Xnext);
        }
        // 7. x = Xcur + D1 into Xa, the warm-start state, then the inverse DFT
        //    (FP64) and the axis ghost; the panel DFT does both in one pass
        if (inv512) {
            out->push_back(Op::mark(3));
            DftArgs d{Xcur, M, d_pot, M, o.dinv64, N, (int)M, 0, 0, p->D1, Xa};
            d.axis_numr = numr;
            out->push_back(dft_op(d, DFT_INV_F64));
        } else {
            out->push_back(add_update_op(Xcur, p->D1, Xa, numr, numz, N));  // pointwise: in place is safe
            out->push_back(Op::mark(3));
            dft_inv64_ops(p, Xa, M, (int)M, d_pot, M, out);
            out->push_back(axis_op(d_pot, numr, numz, N));
        }
    }
    out->push_back(Op::mark(4));
}

// The plain FP64 solve: the reference's sequence (potential_solver.F90:66-106).
void fp64_ops(const bh_plan *p, const double *d_rho, double *d_pot, int tstep, std::vector<Op> *out)
{
    const Geo &g = p->g;
    const ModeOps &o = *p->ops;
    const int numr = g.numr, numz = g.numz, N = g.numphi, nri = g.nri, nzi = g.nzi;
    const long M = (long)numr * numz;
    const bool warm = tstep > 1;
    const int w = warm ? 1 : 0;
    out->push_back(Op::mark(0));
    // 1. rho-hat (with rhop's zeroed layers, potential_solver.F90:70-78) on
    //    the side stream, on half the CUs, beside the boundary projections
    //    of rho (bessel.F90:179-185, 228-237) and the boundary
    out->push_back(Op::mark(MARK_FORK));
    dft_fwd64_ops(p, d_rho, M, (int)M, p->RJ, M, true, out, side_blocks(p));
    out->push_back(Op::mark(MARK_SIDE_END));
    proj64_ops(p, d_rho, out);
    out->push_back(Op::mark(1));
    // 2. starting potential: zero (cold, potential_solver.F90:88-93: the
    //    first half-step reads no interior) or the previous solution kept in
    //    Fourier space (warm); then the boundary potential -> Fourier-space
    //    faces, placed in both layouts
    boundary_ops(p, p->XJ, p->XK, 1, out);
    // 3. rho-hat's compact K-layout interior for the z half-steps, after the
    //    boundary rather than beside it on the side stream (there it slowed
    //    the FP64-table contraction by more than it takes alone)
    out->push_back(transpose_op(p->RJ, p->RK, numr, numz, N));
    out->push_back(Op::mark(2));
    // 4. Peaceman-Rachford iterations (helmadi.F90:195-355).  Between the
    //    half-steps the interior lives in the tiled buffers (the r half-step
    //    writes TK, the z half-step TJ); the first half-step reads XJ, the
    //    last writes XJ, the warm-start state; the faces stay in XJ and XK.
    const std::vector<double> &dt = o.dt[w];
    const int nst = (int)dt.size();
    for (int st = 0; st < nst; ++st) {
        LineArgs r{};
        r.xin = p->XJ; r.xout = p->XK; r.rh = p->RJ;
        r.SA = M; r.PA = numr; r.SB = M; r.PB = numz;
        r.xt = p->TJ; r.STi = tiled_stride(nzi, nri); r.in_tiled = st > 0; r.cold = st == 0 && !warm;
        r.yt = p->TK; r.STo = tiled_stride(nri, nzi); r.out_tiled = 1;
        r.nlines = nzi; r.len = nri;
        r.alpha = o.rc_alpha[w] + (size_t)st * N * nri;
        r.ibet = o.rc_ibet[w] + (size_t)st * N * nri;
        r.gback = o.rc_gback[w] + (size_t)st * N * nri;
        r.coef_stride = nri;
        r.dir = 0; r.dt = dt[st]; r.gamma = g.gamma; r.four_pi = g.four_pi;
        r.alphar = o.alphar; r.betar = o.betar; r.elzb = o.elzb; r.numr = numr;
        r.eqsym = g.isym != 1;
        out->push_back(sweep_op(r, N, 0));
        LineArgs z = r;
        z.xin = p->XK; z.xout = p->XJ; z.rh = p->RK;
        z.PA = numz; z.PB = numr;
        z.xt = p->TK; z.STi = tiled_stride(nri, nzi); z.in_tiled = 1; z.cold = 0;
        z.rh_compact = 1; z.STr = tiled_stride(nri, nzi);
        z.yt = p->TJ; z.STo = tiled_stride(nzi, nri); z.out_tiled = st + 1 < nst;
        z.nlines = nri; z.len = nzi;
        z.alpha = o.zc_alpha[w] + (size_t)st * nzi;
        z.ibet = o.zc_ibet[w] + (size_t)st * nzi;
        z.gback = o.zc_gback[w] + (size_t)st * nzi;
        z.coef_stride = 0;
        z.dir = 1;
        out->push_back(sweep_op(z, N, 0));
    }
    out->push_back(Op::mark(3));
    // 5. inverse DFT with the 2/N normalisation, then the axis ghost (a
    //    mirror for isym = 3) and, for isym = 2, 3, the equator
    //    (helmadi.F90:396-411)
    dft_inv64_ops(p, p->XJ, M, (int)M, d_pot, M, out);
    out->push_back(axis_op(d_pot, numr, numz, N, g.isym == 3 ? AXIS_MIRROR : AXIS_ROTATE));
    if (g.isym != 1) out->push_back(axis_op(d_pot, numr, numz, N, EQUATOR));
    out->push_back(Op::mark(4));
}

// The FP64 spectral solve: the truncated Peaceman-Rachford iteration of
// helmadi.F90:195-355 as x_N = X0 + delta, A delta = (I - M) r0 (the mixed
// solve's form), evaluated once in FP64 (bh_spec64.hip).
void spec64_ops(const bh_plan *p, const double *d_rho, double *d_pot, int tstep, std::vector<Op> *out)
{
    const Geo &g = p->g;
    const ModeOps &o = *p->ops;
    const int numr = g.numr, numz = g.numz, N = g.numphi, nri = g.nri, nzi = g.nzi, nh = N / 2 + 1;
    const long M = (long)numr * numz, PIm = (long)nri * nzi;
    const int w = tstep > 1 ? 1 : 0, cold = tstep <= 1 ? 1 : 0;
    out->push_back(Op::mark(0));
    // 1-2. rho-hat (masked as rhop) on the side stream beside the
    //      projections; the boundary faces into XJ (XK takes a copy)
    out->push_back(Op::mark(MARK_FORK));
    dft_fwd64_ops(p, d_rho, M, (int)M, p->RJ, M, true, out, side_blocks(p));
    out->push_back(Op::mark(MARK_SIDE_END));
    proj64_ops(p, d_rho, out);
    out->push_back(Op::mark(1));
    boundary_ops(p, p->XJ, p->XK, 0, out);
    out->push_back(Op::mark(2));
    // 3. r0 = g - A X0 (X0 zero inside for a cold start)
    Res64Args ra{p->XJ, p->RJ, p->R64, o.brb_d, o.alphar, o.betar, g.gamma, g.four_pi, numr, numz, cold};
    out->push_back(res64_op(ra, N));
    // 4. T = S r0 per plane
    GemmArgs a{};
    a.A = o.S64; a.lda = nzi; a.sA = 0;
    a.B = p->R64; a.ldb = nri; a.sB = PIm;
    a.C = p->T64; a.ldc = nri; a.sC = PIm;
    a.R = nzi; a.M = nri; a.Kd = nzi; a.batch = N;
    a.tile = 2;
    out->push_back(gemm_op(a, G64_PLAIN));
    // 5. U = T M_r^T: planes 0..N/2 take mode numbers 0..N/2, planes
    //    N/2+1..N-1 mode numbers 1..N/2-1
    GemmArgs b{};
    b.A = p->T64; b.lda = nri; b.sA = PIm;
    b.B = o.MrT64[w]; b.ldb = nri; b.sB = (long)nri * nri;
    b.C = p->U64; b.ldc = nri; b.sC = PIm;
    b.R = nzi; b.M = nri; b.Kd = nri; b.batch = nh;
    b.tile = 2;
    out->push_back(gemm_op(b, G64_PLAIN));
    if (N - nh > 0) {
        b.A = p->T64 + nh * PIm; b.B = o.MrT64[w] + (long)nri * nri; b.C = p->U64 + nh * PIm;
        b.batch = N - nh;
        out->push_back(gemm_op(b, G64_PLAIN));
    }
    // 6. (B_r + lambda_k) D = T - mu_k U along J, D into U
    Thomas64Args ta{p->T64, p->U64, o.mu64[w], o.lam64, o.brb_d, o.alphar, o.betar, numr, nri, nzi};
    out->push_back(thomas64_op(ta, N));
    // 7. delta = S D, then X = X0 + delta
    a.B = p->U64; a.C = p->R64;
    out->push_back(gemm_op(a, G64_PLAIN));
    Update64Args ua{p->XJ, p->R64, numr, numz, cold};
    out->push_back(update64_op(ua, N));
    out->push_back(Op::mark(3));
    // 8. inverse DFT and the axis ghost (helmadi.F90:357-411)
    dft_inv64_ops(p, p->XJ, M, (int)M, d_pot, M, out);
    out->push_back(axis_op(d_pot, numr, numz, N));
    out->push_back(Op::mark(4));
}

int collect_timing(bh_plan *p)
{
    if (!p->timing || p->tcount == 0) return BH_OK;
    HIPCHK(hipEventSynchronize(p->tsets[p->tcount - 1].ev[4]));
    double sum[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < p->tcount; ++i) {
        const TimingSet &t = p->tsets[i];
        float d[4], tot, ct;
        for (int k = 0; k < 4; ++k) HIPCHK(hipEventElapsedTime(&d[k], t.ev[k], t.ev[k + 1]));
        HIPCHK(hipEventElapsedTime(&tot, t.ev[0], t.ev[4]));
        HIPCHK(hipEventElapsedTime(&ct, t.ev[5], t.ev[6]));
        // order: boundary, forward DFT, ADI, inverse DFT, total, top/bottom contraction
        sum[0] += d[1]; sum[1] += d[0]; sum[2] += d[2]; sum[3] += d[3]; sum[4] += tot; sum[5] += ct;
    }
    for (int k = 0; k < 6; ++k) p->last_ms[k] = sum[k] / p->tcount;
    p->tcount = 0;
    p->have_timing = true;
    return BH_OK;
}

}  // namespace

bool lazy_cold(const bh_plan *p, int tstep)
{
    return tstep <= 1 && is_mixed(p->precision) && p->ops->spec_fused && p->ops->dinv64 && p->use_dft &&
           p->g.numphi == DFT_MAX_N;
}

void build_ops(const bh_plan *p, const double *d_rho, double *d_pot, int tstep, std::vector<Op> *out,
               bool allow_lazy)
{
    out->clear();
    if (is_mixed(p->precision)) mixed_ops(p, d_rho, d_pot, tstep, out, allow_lazy);
    else if (is_spec64(p->precision)) spec64_ops(p, d_rho, d_pot, tstep, out);
    else fp64_ops(p, d_rho, d_pot, tstep, out);
}

int enqueue_ops(bh_plan *p, const std::vector<Op> &ops, hipStream_t s)
{
    TimingSet *ts = nullptr;
    if (p->timing && p->tcount < 1024) {
        if (p->tcount == (int)p->tsets.size()) {
            TimingSet t;
            for (auto &e : t.ev) HIPCHK(hipEventCreate(&e));
            p->tsets.push_back(t);
        }
        ts = &p->tsets[p->tcount++];
    }
    // after the plan's previous solve, whatever stream it was enqueued on
    if (p->has_last && p->last_stream != s) {
        HIPCHK(hipEventRecord(p->done_ev, p->last_stream));
        HIPCHK(hipStreamWaitEvent(s, p->done_ev, 0));
    }
    hipStream_t cur = s;
    for (const Op &op : ops) {
        if (op.kind == OP_MARK) {
            if (op.variant == MARK_FORK) {
                if (!p->overlap) continue;
                HIPCHK(hipEventRecord(p->fork_ev, s));
                HIPCHK(hipStreamWaitEvent(p->side, p->fork_ev, 0));
                cur = p->side;
            } else if (op.variant == MARK_SIDE_END) {
                cur = s;
            } else if (op.variant == MARK_JOIN) {
                if (!p->overlap) continue;
                HIPCHK(hipEventRecord(p->join_ev, p->side));
                HIPCHK(hipStreamWaitEvent(s, p->join_ev, 0));
            } else if (ts) {
                HIPCHK(hipEventRecord(ts->ev[op.variant], s));
            }
        } else if (!(op.kind == OP_FILL && op.as<FillArgs>().n == 0)) {
            launch_op(op, cur);
        }
    }
    HIPCHK(hipGetLastError());
    p->last_stream = s;
    p->has_last = true;
    p->has_solution = true;
    return BH_OK;
}

int materialize_state(bh_plan *p, hipStream_t s)
{
    if (!p->lazy_state) return BH_OK;
    const Geo &g = p->g;
    const std::vector<Op> ops{add_update_op(p->XJ, p->D1, p->XJ, g.numr, g.numz, g.numphi, correction_sum(p), 1)};
    const bool timing = p->timing;
    p->timing = false;  // not a solve: no stage events
    const int rc = enqueue_ops(p, ops, s);
    p->timing = timing;
    if (rc != BH_OK) return rc;
    launch_fillf(p->lazy_flag, 1, 0.f, s);  // written out: the next warm start reads it whole
    p->lazy_state = false;
    return BH_OK;
}

int wait_plan_idle(bh_plan *p)
{
    if (p->has_last) {
        HIPCHK(hipEventRecord(p->done_ev, p->last_stream));
        HIPCHK(hipEventSynchronize(p->done_ev));
    }
    return BH_OK;
}

}  // namespace bh

namespace {

int enqueue_solve(bh_plan *p, const double *d_rho, double *d_pot, int tstep, hipStream_t s)
{
    std::vector<Op> ops;
    // a solve captured into a graph may be replayed after other calls (an
    // import among them): it keeps the state whole and forms every row's
    // faces.  Whether its warm start writes out a state left as corrections
    // is decided on the device (lazy_flag) at each replay.
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (s) HIPCHK(hipStreamIsCapturing(s, &cs));
    const bool capturing = cs != hipStreamCaptureStatusNone;
    const bool lazy_ok = !p->in_batch && !capturing;
    build_ops(p, d_rho, d_pot, tstep, &ops, lazy_ok);
    const int rc = enqueue_ops(p, ops, s);
    p->lazy_state = rc == BH_OK && lazy_ok && lazy_cold(p, tstep) && keeps_corrections(p);
    p->faces_clean = rc == BH_OK;  // every row's faces formed (or kept at zero)
    return rc;
}

// the host-path copy threads and their pinned ring (no exception leaves
// the library: thread or memory exhaustion is an error code)
int ensure_stager(bh_plan *p)
{
    if (p->stager) return BH_OK;
    try {
        p->stager.reset(new HostStager(p->device, host_copy_workers()));
    } catch (const std::exception &e) {
        return fail(BH_E_DEVICE, std::string("could not start the host copy threads: ") + e.what());
    }
    return BH_OK;
}

int ensure_staging(bh_plan *p, size_t n)
{
    if (p->rho_in) return BH_OK;
    int rc;
    if ((rc = p->mem.alloc(&p->rho_in, n)) != BH_OK) return rc;
    if ((rc = p->mem.alloc(&p->pot_out, n)) != BH_OK) return rc;
    p->bytes += 2 * n * sizeof(double);
    return BH_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" const char *bh_last_error(void) { return bh::last_error(); }

extern "C" int bh_plan_create(const bh_grid_desc *desc, int precision, size_t hbm_budget, bh_plan **out)
{
    if (!desc || !out) return fail(BH_E_INVALID, "null argument");
    *out = nullptr;
    bh_grid_desc d;
    int rc = normalize_desc(desc, precision, &d);
    if (rc != BH_OK) return rc;
    return plan_create(d, precision, hbm_budget, out);
}

extern "C" int bh_plan_destroy(bh_plan *p)
{
    if (!p) return BH_OK;
    (void)hipDeviceSynchronize();
    for (auto &t : p->tsets)
        for (auto &e : t.ev) (void)hipEventDestroy(e);
    delete p;
    return BH_OK;
}

extern "C" size_t bh_plan_device_bytes(const bh_plan *p) { return p ? p->bytes : 0; }

extern "C" int bh_plan_set_timing(bh_plan *p, int on)
{
    if (!p) return fail(BH_E_INVALID, "null plan");
    p->timing = on != 0;
    p->tcount = 0;
    return BH_OK;
}

extern "C" int bh_last_timings(const bh_plan *p, double *ms, int n)
{
    if (!p || !ms) return fail(BH_E_INVALID, "null argument");
    if (!p->have_timing) return fail(BH_E_STATE, "no timed solve yet (bh_plan_set_timing, then solve)");
    for (int i = 0; i < n && i < 6; ++i) ms[i] = p->last_ms[i];
    return BH_OK;
}

extern "C" int bh_sync_timing(bh_plan *p) { return p ? collect_timing(p) : fail(BH_E_INVALID, "null plan"); }

extern "C" int bh_solve_device(bh_plan *p, const double *d_rho, double *d_pot, int tstep, void *stream)
{
    if (!p || !d_rho || !d_pot) return fail(BH_E_INVALID, "null argument");
    return enqueue_solve(p, d_rho, d_pot, tstep, (hipStream_t)stream);
}

extern "C" int bh_solve(bh_plan *p, const double *rho, double *pot, int tstep)
{
    if (!p || !rho || !pot) return fail(BH_E_INVALID, "null argument");
    const size_t n = (size_t)p->g.numr * p->g.numz * p->g.numphi;
    int rc = ensure_staging(p, n);
    if (rc != BH_OK) return rc;
    if ((rc = ensure_stager(p)) != BH_OK) return rc;
    // rho in through the pinned ring, checked for finiteness as it is
    // copied; the solve is enqueued behind the transfers and phi comes back
    // the same way
    long bad = -1;
    rc = p->stager->to_device(p->rho_in, rho, n, 0, true, &bad);
    if (rc != BH_OK) return fail(rc, "copy of rho to the device failed");
    if (bad >= 0) return fail(BH_E_NONFINITE, "rho holds a non-finite value at index " + std::to_string(bad));
    rc = enqueue_solve(p, p->rho_in, p->pot_out, tstep, 0);
    if (rc != BH_OK) return rc;
    rc = p->stager->to_host(pot, p->pot_out, n, 0);
    if (rc != BH_OK) return fail(rc, "copy of the potential to the host failed");
    return collect_timing(p);
}

extern "C" int bh_set_potential(bh_plan *p, const double *pot, int on_device)
{
    if (!p || !pot) return fail(BH_E_INVALID, "null argument");
    const Geo &g = p->g;
    const long M = (long)g.numr * g.numz;
    const size_t n = (size_t)M * g.numphi;
    const double *src = pot;
    int rc = wait_plan_idle(p);  // a solve still running reads and writes XJ
    if (rc != BH_OK) return rc;
    p->lazy_state = false;  // XJ is rewritten whole
    p->faces_clean = false;
    if (p->lazy_flag) launch_fillf(p->lazy_flag, 1, 0.f, 0);
    if (!on_device) {
        rc = ensure_staging(p, n);
        if (rc != BH_OK) return rc;
        if ((rc = ensure_stager(p)) != BH_OK) return rc;
        rc = p->stager->to_device(p->pot_out, pot, n, 0, false, nullptr);  // through the pinned ring
        if (rc != BH_OK) return fail(rc, "copy of the potential to the device failed");
        src = p->pot_out;
    }
    std::vector<Op> ops;
    dft_fwd64_ops(p, src, M, (int)M, p->XJ, M, false, &ops);
    for (const Op &op : ops) launch_op(op, 0);
    HIPCHK(hipDeviceSynchronize());
    p->has_solution = true;
    return BH_OK;
}

// ---------------------------------------------------------------------------
// Test hooks (not part of the public ABI): run the GEMM on device buffers and
// copy a plan's internal buffers to the host, for stage-by-stage checks.
// ---------------------------------------------------------------------------
extern "C" int bhx_gemm_f64(const double *A, long lda, const double *B, long ldb, int R, int M, int Kd, double *C,
                            long ldc)
{
    GemmArgs a{};
    a.A = A; a.lda = lda; a.B = B; a.ldb = ldb; a.C = C; a.ldc = ldc; a.R = R; a.M = M; a.Kd = Kd; a.batch = 1;
    a.tile = (M >= 256 && R >= 64) ? 0 : 2;
    gemm64(a, B_PLAIN, E_PLAIN, 0);
    HIPCHK(hipDeviceSynchronize());
    return BH_OK;
}

// which: 0 RJ, 1 RK (compact K-layout interior), 2 XJ, 3 XK, 4 proj, 5 T (blocked symmetric), 6 S, 7 Stb, 8 Ss, 9 (unused), 10 FF
extern "C" int bhx_snapshot(bh_plan *p, int which, double *host, size_t n)
{
    if (!p || !host) return fail(BH_E_INVALID, "null argument");
    if (which == 2) {  // the warm-start state, whole
        const int rc = materialize_state(p, 0);
        if (rc != BH_OK) return rc;
    }
    const Geo &g = p->g;
    const size_t M = (size_t)g.numr * g.numz, N = g.numphi, mm = g.mmax, nri = g.nri, nzi = g.nzi;
    const size_t nf = 2 * nri + nzi;
    const double *src[] = {p->RJ, p->RK, p->XJ, p->XK, p->proj, p->tab->T, p->tab->S, p->Stb, p->Ss, nullptr, p->FF};
    const size_t cnt[] = {M * N, p->RK ? N * tiled_stride(g.nri, g.nzi) : 0, M * N, M * N, M * (2 * mm - 1), topsym_table_elems(g.numr, g.numz, g.mmax),
                          mm * (2 * nzi - 1) * nri, mm * 4 * nri, side_splits(g.numz) * mm * 2 * nzi, N * nf, N * nf};
    if (which < 0 || which > 10) return fail(BH_E_INVALID, "which");
    if (!src[which]) return fail(BH_E_STATE, "buffer not used by this plan's precision");
    if (n < cnt[which]) return fail(BH_E_INVALID, "buffer too small: need " + std::to_string(cnt[which]));
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(host, src[which], cnt[which] * sizeof(double), hipMemcpyDeviceToHost));
    return BH_OK;
}

// forward (inverse = 0) or inverse (inverse = 1) real DFT of nplanes x ncols
// device data with the plan's matrices; test hook for the folded DFT GEMMs.
extern "C" int bhx_dft64(const bh_plan *p, const double *in, double *out, int ncols, int inverse)
{
    if (!p || !in || !out) return fail(BH_E_INVALID, "null argument");
    std::vector<Op> ops;
    if (inverse) dft_inv64_ops(p, in, ncols, ncols, out, ncols, &ops);
    else dft_fwd64_ops(p, in, ncols, ncols, out, ncols, false, &ops);
    for (const Op &op : ops) launch_op(op, 0);
    HIPCHK(hipDeviceSynchronize());
    return BH_OK;
}

// forward real DFT in FP32 (the mixed plan's transform of rho, unmasked) of
// nplanes x ncols device data; test hook.
extern "C" int bhx_dft32(const bh_plan *p, const double *in, float *out, int ncols)
{
    if (!p || !in || !out) return fail(BH_E_INVALID, "null argument");
    if (!p->ops->dfwd32) return fail(BH_E_STATE, "plan has no FP32 transform");
    DftArgs d{in, ncols, out, ncols, p->ops->dfwd32, p->g.numphi, ncols, 0, 0};
    launch_op(dft_op(d, DFT_FWD_F32), 0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    return BH_OK;
}

// azimuthal mode projections of numphi x numr*numz device data into
// (2 mmax - 1) x numr*numz; test hook.
extern "C" int bhx_proj(const bh_plan *p, const double *rho, double *proj)
{
    if (!p || !rho || !proj) return fail(BH_E_INVALID, "null argument");
    launch_op(proj_op(rho, p->ops->Pf, proj, p->g.numr * p->g.numz, p->g.numphi, p->g.mmax), 0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    return BH_OK;
}

// FP64 / FP32 GEMM with an explicit tile shape (0..4) and transposes (FP32
// only); test hook.
extern "C" int bhx_gemm_tile(int fp32, const void *A, long lda, const void *B, long ldb, int R, int M, int Kd, void *C,
                             long ldc, int tile, int ta, int tb)
{
    GemmArgs a{};
    a.A = A; a.lda = lda; a.B = B; a.ldb = ldb; a.C = C; a.ldc = ldc; a.R = R; a.M = M; a.Kd = Kd; a.batch = 1;
    a.tile = tile;
    if (fp32) gemm32(a, B_PLAIN, E_PLAIN, ta != 0, tb != 0, false, 0);
    else gemm64(a, B_PLAIN, E_PLAIN, 0);
    HIPCHK(hipDeviceSynchronize());
    return BH_OK;
}

// Folded sine transform of length n over batch blocks with the folded table
// Sf (ceil(n/2) x n, columns in spectral slot order); test hook.  Forward:
// out[b] (nrows x n) = in[b]^T S with in[b] n x nrows; inverse: out[b]
// (n x nrows) = S in[b]^T with in[b] nrows x n.
extern "C" int bhx_dst32(const float *Sf, int n, const float *in, float *out, int nrows, int batch, int inverse,
                         int tile)
{
    GemmArgs a{};
    const long blk = (long)n * nrows;
    a.nfold = n; a.batch = batch; a.tile = tile; a.Kd = n;
    if (inverse) {
        a.A = Sf; a.lda = n; a.B = in; a.ldb = n; a.sB = blk; a.C = out; a.ldc = nrows; a.sC = blk;
        a.R = (n + 1) / 2; a.M = nrows;
        launch_op(gemm_op(a, G32_DST_INV), 0);
    } else {
        a.A = in; a.lda = nrows; a.sA = blk; a.B = Sf; a.ldb = n; a.C = out; a.ldc = n; a.sC = blk;
        a.R = nrows; a.M = n;
        launch_op(gemm_op(a, G32_DST_FWD), 0);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    return BH_OK;
}
