// bh_hydro.hip -- the hydro step that consumes the potential, on the device:
// the momentum source step of bipoly_hydro (source.F90 with
// make_source_temp.F90, source_bc.F90, make_source_temp_bc.F90), for one
// global block (isym = 1).  With the potential left in HBM by the solve, the
// hydro code's source step needs no 137 MB download per step.
//
// Arrays are the reference's (numr, numz, numphi) Fortran-order fields,
// ghost layers included, indexed [L][K][J] with J fastest.  The interior
// updates are pointwise given the fields they read, one thread per (J, K,
// L), J along the lanes (coalesced); the boundary conditions follow in the
// reference's order (axis, bottom, edge, top), one launch each.  The axis
// averages sum over L in the reference's sequential order, so every value
// is formed with the reference's operations in the reference's order.
#include "bh_plan_impl.h"

#include <hip/hip_runtime.h>

#include <string>

namespace bh {
namespace {

struct SrcArgs {
    double *s, *t, *a, *temps, *tempa;
    const double *rho, *p, *pot;
    const double *gr;            // [4][numr + 1], 1-based J: r, rhf, rinv, rhfinv
    const double *cac, *cvc;     // [numphi][3]: cylin_a_com, cylin_v_com
    double dt, omega, R_com, dr, drinv, dzinv, dphiinv;
    int numr, numz, numphi, com; // com: tstep > 1000 (the centre-of-mass terms)
    int bc[3];                   // bottom, edge, top: 1 wall, 2 free, else Dirichlet
};

__device__ __forceinline__ long ix(const SrcArgs &g, int J, int K, int L)
{
    return ((long)(L - 1) * g.numz + (K - 1)) * g.numr + (J - 1);
}

// phieff = pot - 0.5 Omega^2 rhf^2 (source.F90:149-160)
__device__ __forceinline__ double phieff(const SrcArgs &g, int J, int K, int L)
{
    const double hf = g.gr[(g.numr + 1) + J];
    return g.pot[ix(g, J, K, L)] - 0.5 * g.omega * g.omega * hf * hf;
}

// sa - dt * (radial pressure + gravity gradient - centrifugal and Coriolis
// terms of the azimuthal momentum src at the radial faces), zero where the
// face density vanishes (source.F90:225-284; make_source_temp.F90:139-208)
__device__ __forceinline__ double radial_term(const SrcArgs &g, const double *src, double sa, int J, int K, int L)
{
    const int np1 = g.numr + 1;
    const double rj = g.gr[J], ri = g.gr[2 * np1 + J], q = 0.25 * g.dr;
    const int Lp = L < g.numphi ? L + 1 : 1;
    const double rh = 0.5 * ri * (g.rho[ix(g, J, K, L)] * (rj + q) + g.rho[ix(g, J - 1, K, L)] * (rj - q));
    if (!(rh > 0.0)) return 0.0;
    const double tp = 0.25 * ri * ((src[ix(g, J, K, L)] + src[ix(g, J, K, Lp)]) * (rj + q) +
                                   (src[ix(g, J - 1, K, L)] + src[ix(g, J - 1, K, Lp)]) * (rj - q));
    double f = g.drinv * ((g.p[ix(g, J, K, L)] - g.p[ix(g, J - 1, K, L)]) +
                          rh * (phieff(g, J, K, L) - phieff(g, J - 1, K, L))) -
               tp * tp * ri * ri * ri / rh - 2.0 * g.omega * tp * ri;
    if (g.com) f = f + rh * (g.cac[3 * (L - 1)] - g.omega * g.omega * g.R_com - 2 * g.omega * g.cvc[3 * (L - 1) + 1]);
    return sa - g.dt * f;
}

// aa - dt * (azimuthal pressure + gravity gradient + Coriolis term of the
// radial momentum src at the azimuthal faces) (source.F90:286-375;
// make_source_temp.F90:210-297)
__device__ __forceinline__ double azimuthal_term(const SrcArgs &g, const double *src, double aa, int J, int K, int L)
{
    const int np1 = g.numr + 1;
    const int Lm = L > 1 ? L - 1 : g.numphi;
    const double hf = g.gr[np1 + J], hfi = g.gr[3 * np1 + J], q = 0.25 * g.dr;
    const double rh = 0.5 * (g.rho[ix(g, J, K, L)] + g.rho[ix(g, J, K, Lm)]);
    const double tp = 0.25 * hfi * ((src[ix(g, J + 1, K, L)] + src[ix(g, J + 1, K, Lm)]) * (hf + q) +
                                    (src[ix(g, J, K, L)] + src[ix(g, J, K, Lm)]) * (hf - q));
    double f = g.dphiinv * ((g.p[ix(g, J, K, L)] - g.p[ix(g, J, K, Lm)]) +
                            rh * (phieff(g, J, K, L) - phieff(g, J, K, Lm))) +
               2.0 * g.omega * hf * tp;
    if (g.com) f = f + rh * hf * (g.cac[3 * (L - 1) + 1] + 2 * g.omega * g.cvc[3 * (L - 1)]);
    return aa - g.dt * f;
}

// make_source_temp's interior: temps from s and a, tempa from a and s
__global__ __launch_bounds__(256) void source_temp_kernel(SrcArgs g)
{
    const int J = blockIdx.x * 256 + threadIdx.x + 2, K = blockIdx.y + 2, L = blockIdx.z + 1;
    if (J > g.numr - 1) return;
    const long c = ix(g, J, K, L);
    g.temps[c] = radial_term(g, g.a, g.s[c], J, K, L);
    g.tempa[c] = azimuthal_term(g, g.s, g.a[c], J, K, L);
}

// source's interior: t from the vertical gradient (source.F90:172-203), s
// from tempa at the radial faces, a from temps at the azimuthal faces; each
// thread reads only its own s, t, a, so in place is safe
__global__ __launch_bounds__(256) void source_kernel(SrcArgs g)
{
    const int J = blockIdx.x * 256 + threadIdx.x + 2, K = blockIdx.y + 2, L = blockIdx.z + 1;
    if (J > g.numr - 1) return;
    const long c = ix(g, J, K, L), cm = ix(g, J, K - 1, L);
    const double rh = 0.5 * (g.rho[c] + g.rho[cm]);
    const double grad = (g.p[c] - g.p[cm]) + rh * (phieff(g, J, K, L) - phieff(g, J, K - 1, L));
    const double tv = g.t[c];
    g.t[c] = g.com ? tv - g.dt * (g.dzinv * grad + rh * g.cac[3 * (L - 1) + 2]) : tv - g.dt * g.dzinv * grad;
    const double sv = radial_term(g, g.tempa, g.s[c], J, K, L);
    const double av = azimuthal_term(g, g.temps, g.a[c], J, K, L);
    g.s[c] = sv;
    g.a[c] = av;
}

// Boundary conditions of a radial / azimuthal pair (sr, sa) and, for the
// momenta, t (source_bc.F90; make_source_temp_bc.F90 with t null).
// Axis, one workgroup per K = 2..numz-1 (source_bc.F90:45-93): the radial
// component vanishes at J = 2, the azimuthal one (and t) take their mean
// over L, summed in order by one thread from LDS; the J = 1 ghosts are the
// pi rotation of J = 3 (radial, negated) and J = 2.
constexpr int AXIS_THREADS = 256;
__global__ __launch_bounds__(AXIS_THREADS) void bc_axis_kernel(SrcArgs g, double *sr, double *tt, double *sa)
{
    extern __shared__ double axis_lds[];  // [2][numphi]
    const int K = blockIdx.x + 2, N = g.numphi, h = N / 2;
    double *va = axis_lds, *vt = axis_lds + N;
    for (int L = 1 + threadIdx.x; L <= N; L += AXIS_THREADS) {
        va[L - 1] = sa[ix(g, 2, K, L)];
        if (tt) vt[L - 1] = tt[ix(g, 2, K, L)];
    }
    __syncthreads();
    __shared__ double mean[2];
    if (threadIdx.x == 0) {
        const double ninv = 1.0 / (double)N;
        double sum = 0.0;
        for (int L = 0; L < N; ++L) sum = sum + va[L];
        mean[0] = sum * ninv;
        if (tt) {
            sum = 0.0;
            for (int L = 0; L < N; ++L) sum = sum + vt[L];
            mean[1] = sum * ninv;
        }
    }
    __syncthreads();
    for (int L = 1 + threadIdx.x; L <= N; L += AXIS_THREADS) {
        const int Lr = L <= h ? L + h : L - h;
        sr[ix(g, 2, K, L)] = 0.0;
        sa[ix(g, 2, K, L)] = mean[0];
        sr[ix(g, 1, K, L)] = -sr[ix(g, 3, K, Lr)];
        sa[ix(g, 1, K, L)] = mean[0];
        if (tt) {
            tt[ix(g, 2, K, L)] = mean[1];
            tt[ix(g, 1, K, L)] = mean[1];
        }
    }
}

// bottom (K = 1 ghost), one thread per (J = 2..numr-1, L)
// (source_bc.F90:95-148, make_source_temp_bc.F90:76-110, isym = 1)
__global__ __launch_bounds__(256) void bc_bottom_kernel(SrcArgs g, double *sr, double *tt, double *sa)
{
    const int J = blockIdx.x * 256 + threadIdx.x + 2, L = blockIdx.y + 1;
    if (J > g.numr - 1) return;
    const long c1 = ix(g, J, 1, L), c2 = ix(g, J, 2, L);
    if (g.bc[0] == 1) {
        sr[c1] = sr[c2];
        if (tt) {
            tt[c1] = -tt[ix(g, J, 3, L)];
            tt[c2] = 0.0;
        }
        sa[c1] = sa[c2];
    } else if (g.bc[0] == 2) {
        sr[c1] = sr[c2];
        sa[c1] = sa[c2];
        if (tt) tt[c1] = tt[c2];
    } else {
        sr[c1] = 0.0;
        sa[c1] = 0.0;
        if (tt) {
            tt[c1] = 0.0;
            tt[c2] = 0.5 * (fabs(tt[c2]) - tt[c2]);
        }
    }
}

// edge (J = numr ghost), one thread per (K = 2..numz-1, L)
// (source_bc.F90:165-196, make_source_temp_bc.F90:112-143)
__global__ __launch_bounds__(256) void bc_edge_kernel(SrcArgs g, double *sr, double *tt, double *sa)
{
    const int K = blockIdx.x * 256 + threadIdx.x + 2, L = blockIdx.y + 1;
    if (K > g.numz - 1) return;
    const int nr = g.numr;
    const long ce = ix(g, nr, K, L), ci = ix(g, nr - 1, K, L);
    if (g.bc[1] == 1) {
        sr[ce] = -sr[ix(g, nr - 2, K, L)];
        sr[ci] = 0.0;
        sa[ce] = sa[ci];
        if (tt) tt[ce] = tt[ci];
    } else if (g.bc[1] == 2) {
        sr[ce] = sr[ci];
        if (tt) tt[ce] = tt[ci];
        sa[ce] = sa[ci];
    } else {
        sr[ce] = 0.0;
        sr[ci] = 0.5 * (fabs(sr[ci]) + sr[ci]);
        sa[ce] = 0.0;
        if (tt) tt[ce] = 0.0;
    }
}

// top (K = numz ghost), one thread per (J = 2..numr-1, L)
// (source_bc.F90:198-230, make_source_temp_bc.F90:145-176)
__global__ __launch_bounds__(256) void bc_top_kernel(SrcArgs g, double *sr, double *tt, double *sa)
{
    const int J = blockIdx.x * 256 + threadIdx.x + 2, L = blockIdx.y + 1;
    if (J > g.numr - 1) return;
    const int nz = g.numz;
    const long ct = ix(g, J, nz, L), ci = ix(g, J, nz - 1, L);
    if (g.bc[2] == 1) {
        sr[ct] = sr[ci];
        sa[ct] = sa[ci];
        if (tt) {
            tt[ct] = -tt[ix(g, J, nz - 2, L)];
            tt[ci] = 0.0;
        }
    } else if (g.bc[2] == 2) {
        sr[ct] = sr[ci];
        if (tt) tt[ct] = tt[ci];
        sa[ct] = sa[ci];
    } else {
        sr[ct] = 0.0;
        sa[ct] = 0.0;
        if (tt) {
            tt[ct] = 0.0;
            tt[ci] = 0.5 * (fabs(tt[ci]) + tt[ci]);
        }
    }
}

void launch_bc(const SrcArgs &g, double *sr, double *tt, double *sa, hipStream_t s)
{
    const int nri = g.numr - 2, nzi = g.numz - 2;
    hipLaunchKernelGGL(bc_axis_kernel, dim3(nzi), dim3(AXIS_THREADS), 2 * g.numphi * sizeof(double), s, g, sr, tt, sa);
    hipLaunchKernelGGL(bc_bottom_kernel, dim3((nri + 255) / 256, g.numphi), dim3(256), 0, s, g, sr, tt, sa);
    hipLaunchKernelGGL(bc_edge_kernel, dim3((nzi + 255) / 256, g.numphi), dim3(256), 0, s, g, sr, tt, sa);
    hipLaunchKernelGGL(bc_top_kernel, dim3((nri + 255) / 256, g.numphi), dim3(256), 0, s, g, sr, tt, sa);
}

int make_args(bh_plan *p, const bh_hydro_fields *f, const bh_source_params *prm, hipStream_t s, SrcArgs *g,
              bool need_momenta)
{
    if (!p || !f || !prm) return fail(BH_E_INVALID, "null argument");
    if (!f->temps || !f->tempa || !f->rho || !f->p || !f->pot || !f->s || !f->a || (need_momenta && !f->t))
        return fail(BH_E_INVALID, "a field pointer is null");
    const Geo &G = p->g;
    if (G.isym != 1)
        return fail(BH_E_INVALID, "the source step is restated for isym = 1 only (its isym = 2, 3 walls are not)");
    *g = SrcArgs{};
    g->s = f->s; g->t = f->t; g->a = f->a; g->temps = f->temps; g->tempa = f->tempa;
    g->rho = f->rho; g->p = f->p; g->pot = f->pot;
    g->gr = p->src_grid; g->cac = p->src_com; g->cvc = p->src_com + 3 * G.numphi;
    g->dt = prm->dt; g->omega = prm->omega_frame; g->R_com = prm->R_com;
    g->dr = G.dr;
    // dz = dr, dphi = 2 pi / numphi (setup_frac.F90:374-379)
    g->drinv = 1.0 / G.dr; g->dzinv = 1.0 / G.dz; g->dphiinv = 1.0 / G.dphi;
    g->numr = G.numr; g->numz = G.numz; g->numphi = G.numphi;
    g->com = prm->tstep > 1000 ? 1 : 0;
    for (int i = 0; i < 3; ++i) g->bc[i] = prm->boundary_condition[i];
    if (g->com) {
        if (!prm->cylin_a_com || !prm->cylin_v_com)
            return fail(BH_E_INVALID, "tstep > 1000 needs cylin_a_com and cylin_v_com");
        // the plan's copy is rewritten once the stream's earlier steps (which
        // may read it) are done; the caller's host arrays are then free
        HIPCHK(hipStreamSynchronize(s));
        HIPCHK(hipMemcpy(p->src_com, prm->cylin_a_com, 3 * G.numphi * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(p->src_com + 3 * G.numphi, prm->cylin_v_com, 3 * G.numphi * sizeof(double),
                         hipMemcpyHostToDevice));
    }
    return BH_OK;
}

void launch_source_temp(const SrcArgs &g, hipStream_t s)
{
    hipLaunchKernelGGL(source_temp_kernel, dim3((g.numr - 2 + 255) / 256, g.numz - 2, g.numphi), dim3(256), 0, s, g);
    launch_bc(g, g.temps, nullptr, g.tempa, s);
}

}  // namespace
}  // namespace bh

using namespace bh;

extern "C" int bh_make_source_temp_device(bh_plan *plan, const bh_hydro_fields *f, const bh_source_params *prm,
                                          void *stream)
{
    const hipStream_t s = (hipStream_t)stream;
    SrcArgs g;
    int rc = make_args(plan, f, prm, s, &g, false);
    if (rc != BH_OK) return rc;
    launch_source_temp(g, s);
    HIPCHK(hipGetLastError());
    return BH_OK;
}

extern "C" int bh_source_device(bh_plan *plan, int source_time, const bh_hydro_fields *f,
                                const bh_source_params *prm, void *stream)
{
    if (source_time != 1 && source_time != 2) return fail(BH_E_INVALID, "source_time must be 1 or 2");
    const hipStream_t s = (hipStream_t)stream;
    SrcArgs g;
    int rc = make_args(plan, f, prm, s, &g, true);
    if (rc != BH_OK) return rc;
    // make_source_temp first (source.F90:206): it reads s and a before this
    // step changes them, and t's update reads neither temps nor tempa
    if (source_time == 2) launch_source_temp(g, s);
    hipLaunchKernelGGL(source_kernel, dim3((g.numr - 2 + 255) / 256, g.numz - 2, g.numphi), dim3(256), 0, s, g);
    launch_bc(g, g.s, g.t, g.a, s);
    HIPCHK(hipGetLastError());
    return BH_OK;
}
