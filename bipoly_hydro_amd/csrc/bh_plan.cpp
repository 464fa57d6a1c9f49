// bh_plan.cpp -- plan construction (the potsetup counterpart) and the solve
// sequence (the potential_solver counterpart) behind the C ABI of
// include/bh_potential.h.
#include "../../include/bh_potential.h"
#include "bh_internal.h"
#include "bh_gemm.h"
#include "bh_mixed.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

using namespace bh;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(BH_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));     \
    } while (0)

// ln Gamma in the 6-term Lanczos form (gamma = 5) published in Numerical
// Recipes, 2nd ed., sec. 6.1.  The Green's-function series coefficients are
// built from it as in tm.F90:125-152; its own error (|eps| < 2e-10) reaches
// the tables, so the same approximation is used here (gammln.F90:14-26).
double lanczos_lngamma(double xx)
{
    static const double c[6] = {76.18009172947146, -86.50532032941677, 24.01409824083091,
                                -1.231739572450155, 0.1208650973866179e-2, -0.5395239384953e-5};
    double x = xx, y = xx;
    double tmp = x + 5.5;
    tmp = (x + 0.5) * std::log(tmp) - tmp;
    double ser = 1.000000000190015;
    for (int j = 0; j < 6; ++j) {
        y = y + 1.0;
        ser = ser + c[j] / y;
    }
    return tmp + std::log(2.5066282746310005 * ser / x);
}

struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
};

}  // namespace

struct bh_plan {
    bh_grid_desc d;
    int precision;
    size_t budget;
    Geo g;
    int nsteps_cold, nsteps_warm;
    // device memory
    std::vector<DevBuf> bufs;
    size_t bytes = 0;
    double *RJ = nullptr, *RK = nullptr, *XJ = nullptr, *XK = nullptr, *proj = nullptr;
    double *T = nullptr, *S = nullptr, *Stb = nullptr, *Ss = nullptr, *F = nullptr, *FF = nullptr;
    // real DFT in halves: Dc/Ds forward (cosine/sine), Ec/Es inverse; Pm projections
    double *Dc = nullptr, *Ds = nullptr, *Ec = nullptr, *Es = nullptr, *Pm = nullptr, *besc = nullptr, *bess = nullptr;
    double *alphar = nullptr, *betar = nullptr, *elzb = nullptr;
    double *rc_alpha[2] = {nullptr, nullptr}, *rc_ibet[2] = {nullptr, nullptr}, *rc_gback[2] = {nullptr, nullptr};
    double *zc_alpha[2] = {nullptr, nullptr}, *zc_ibet[2] = {nullptr, nullptr}, *zc_gback[2] = {nullptr, nullptr};
    double *rho_in = nullptr, *pot_out = nullptr;  // host-path staging
    // mixed precision (BH_PREC_MIXED)
    int passes = 1;                                // FP64 refinement passes
    int tile_dft = 2, tile_spec = 3;               // GEMM tile shapes (bh_gemm.hip), measured best
    float *Dc32 = nullptr, *Ds32 = nullptr, *S32 = nullptr, *mu32[2] = {nullptr, nullptr}, *Mr[2] = {nullptr, nullptr};
    float *tri_ibet = nullptr, *tri_gam = nullptr;
    float *RH32 = nullptr, *r32 = nullptr, *RT = nullptr, *WT = nullptr, *CT = nullptr, *DT = nullptr;
    float *D1 = nullptr, *Wf = nullptr;
    double *brb_d = nullptr;
    std::vector<double> dt[2];                     // 0 cold, 1 warm
    bool has_solution = false;
    bool timing = false;
    hipEvent_t ev[5] = {};
    double last_ms[5] = {0, 0, 0, 0, 0};
    bool have_timing = false;
    int device = 0;
};

namespace {

int dalloc(bh_plan *p, double **out, size_t n)
{
    size_t bytes = n * sizeof(double);
    if (p->budget && p->bytes + bytes > p->budget)
        return fail(BH_E_NOMEM, "plan exceeds the HBM budget (" + std::to_string(p->budget) + " bytes)");
    void *ptr = nullptr;
    hipError_t e = hipMalloc(&ptr, bytes ? bytes : 8);
    if (e != hipSuccess) return fail(BH_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    p->bufs.push_back({ptr, bytes});
    p->bytes += bytes;
    *out = (double *)ptr;
    return BH_OK;
}

int upload(double *dst, const std::vector<double> &src)
{
    HIPCHK(hipMemcpy(dst, src.data(), src.size() * sizeof(double), hipMemcpyHostToDevice));
    return BH_OK;
}

int upload_f(float *dst, const std::vector<float> &src)
{
    HIPCHK(hipMemcpy(dst, src.data(), src.size() * sizeof(float), hipMemcpyHostToDevice));
    return BH_OK;
}

void free_all(bh_plan *p)
{
    for (auto &b : p->bufs) (void)hipFree(b.p);
    p->bufs.clear();
    for (auto &e : p->ev)
        if (e) (void)hipEventDestroy(e);
}

}  // namespace

extern "C" const char *bh_last_error(void) { return g_err.c_str(); }

extern "C" int bh_plan_create(const bh_grid_desc *desc, int precision, size_t hbm_budget, bh_plan **out)
{
    if (!desc || !out) return fail(BH_E_INVALID, "null argument");
    *out = nullptr;
    bh_grid_desc d = *desc;
    if (d.isym == 0) d.isym = 1;
    if (d.mmax == 0) d.mmax = 30;
    if (d.dr == 0.0) d.dr = 1.0 / (130.0 - 3.0);
    if (d.nsteps_cold == 0) d.nsteps_cold = 20;
    if (d.nsteps_warm == 0) d.nsteps_warm = 5;
    if (d.isym != 1) return fail(BH_E_INVALID, "isym: only 1 (no symmetry) is supported");
    if (d.numr < 5 || d.numz < 5) return fail(BH_E_INVALID, "numr and numz must be >= 5");
    if (d.numphi < 4 || (d.numphi & (d.numphi - 1))) return fail(BH_E_INVALID, "numphi must be a power of two >= 4");
    if (d.numz % 2) return fail(BH_E_INVALID, "numz must be even for isym = 1 (symmetric z grid)");
    if (d.mmax < 1 || d.mmax > 250) return fail(BH_E_INVALID, "mmax must be in [1, 250]");
    if (d.numr - 2 > 512 || d.numz - 2 > 512) return fail(BH_E_INVALID, "numr-2 and numz-2 must be <= 512");
    if (d.nsteps_cold < 2 || d.nsteps_warm < 2) return fail(BH_E_INVALID, "nsteps must be >= 2");
    if (!(d.dr > 0.0) || !std::isfinite(d.dr)) return fail(BH_E_INVALID, "dr must be positive");
    if (precision != BH_PREC_FP64 && precision != BH_PREC_MIXED)
        return fail(BH_E_INVALID, "precision must be BH_PREC_FP64 or BH_PREC_MIXED");

    bh_plan *p = new bh_plan();
    p->d = d;
    // GEMM tile shapes, overridable for tuning sweeps
    if (const char *e = std::getenv("BH_GEMM_TILE_DFT")) p->tile_dft = std::atoi(e);
    if (const char *e = std::getenv("BH_GEMM_TILE_SPEC")) p->tile_spec = std::atoi(e);
    p->precision = precision;
    p->budget = hbm_budget;
    (void)hipGetDevice(&p->device);
    const int numr = d.numr, numz = d.numz, N = d.numphi, mmax = d.mmax;
    const int nri = numr - 2, nzi = numz - 2;
    const long M = (long)numr * numz;
    Geo &g = p->g;
    g.numr = numr; g.numz = numz; g.numphi = N; g.mmax = mmax; g.nri = nri; g.nzi = nzi;
    // constants and grid: setup_frac.F90:345, 374-379, 428-463; potsetup.F90:113-114, 151-153
    const double pi = std::acos(-1.0);
    g.dr = d.dr; g.dz = d.dr;
    g.dphi = 2.0 * pi * (1.0 / (double)N);
    g.piinv = 1.0 / pi;
    g.four_pi = 4.0 * pi;
    const double drinv = 1.0 / g.dr, dzinv = 1.0 / g.dz, dphiinv = 1.0 / g.dphi;
    g.gamma = dzinv * dzinv;
    const double drinv2 = drinv * drinv, dphiinv2 = dphiinv * dphiinv;
    std::vector<double> r_g(numr + 1), rhf_g(numr + 1), rinv_g(numr + 1), rhfinv_g(numr + 1), zhf_g(numz + 1);
    for (int J = 1; J <= numr; ++J) {
        double x = (double)J;
        r_g[J] = (x - 2.0) * g.dr;
        rhf_g[J] = r_g[J] + 0.5 * g.dr;
        rinv_g[J] = (r_g[J] != 0.0) ? 1.0 / r_g[J] : 0.0;
        rhfinv_g[J] = 1.0 / rhf_g[J];
    }
    {
        double offset = -((double)(numz / 2) * g.dz);
        for (int K = 1; K <= numz; ++K) zhf_g[K] = offset + ((double)K - 0.5) * g.dz;
    }
    // ADI coefficients: potsetup.F90:155-252 (one block: global radial index)
    std::vector<double> alphar(numr + 1, 0.0), betar(numr + 1, 0.0);
    for (int J = 2; J <= numr - 1; ++J) {
        alphar[J] = -r_g[J + 1] * rhfinv_g[J] * drinv2;
        betar[J] = -r_g[J] * rhfinv_g[J] * drinv2;
    }
    const std::vector<double> &ar = betar, &cr = alphar;
    std::vector<double> brb((size_t)N * (numr + 1), 0.0), elzb((size_t)N * (numr + 1), 0.0);
    for (int L = 1; L <= N; ++L) {
        int lstop = N / 2 + 1;
        int mode = (L <= lstop) ? L - 1 : L - lstop;
        double m1mode = (mode % 2 == 0) ? 1.0 : -1.0;
        double elm = std::cos((double)mode * g.dphi);
        double *b = &brb[(size_t)(L - 1) * (numr + 1)], *e = &elzb[(size_t)(L - 1) * (numr + 1)];
        for (int J = 3; J <= numr - 1; ++J)
            b[J] = 2.0 * drinv2 - 2.0 * (elm - 1.0) * dphiinv2 * rhfinv_g[J] * rhfinv_g[J];
        b[2] = -alphar[2] + (m1mode - 1.0) * betar[2] - 2.0 * (elm - 1.0) * dphiinv2 * rhfinv_g[2] * rhfinv_g[2];
        for (int J = 2; J <= numr - 1; ++J)
            e[J] = -2.0 * drinv2 + 2.0 * (elm - 1.0) * dphiinv2 * rhfinv_g[J] * rhfinv_g[J];
        e[2] = alphar[2] - (m1mode - 1.0) * betar[2] + 2.0 * (elm - 1.0) * dphiinv2 * rhfinv_g[2] * rhfinv_g[2];
    }
    // pseudo-timesteps: helmadi.F90:124-129
    p->nsteps_cold = d.nsteps_cold;
    p->nsteps_warm = d.nsteps_warm;
    for (int w = 0; w < 2; ++w) {
        int ns = w ? d.nsteps_warm : d.nsteps_cold;
        std::vector<double> &dt = p->dt[w];
        dt.assign(ns, 0.0);
        dt[ns - 1] = 4.0 * (rinv_g[numr] * rinv_g[numr]);
        double alph = std::pow(r_g[numr] * rinv_g[3], 2.0 / ((double)ns - 1.0));
        for (int i = ns - 2; i >= 0; --i) dt[i] = alph * dt[i + 1];
    }

    int rc;
#define ALLOC(ptr, n)                                     \
    do {                                                  \
        if ((rc = dalloc(p, &(ptr), (size_t)(n))) != BH_OK) { \
            free_all(p);                                  \
            delete p;                                     \
            return rc;                                    \
        }                                                 \
    } while (0)
    const bool mixed = (precision == BH_PREC_MIXED);
    if (!mixed) {
        ALLOC(p->RJ, M * N);
        ALLOC(p->RK, M * N);
    }
    ALLOC(p->XJ, M * N);
    ALLOC(p->XK, M * N);
    ALLOC(p->proj, M * (2 * mmax - 1));
    ALLOC(p->T, (size_t)mmax * nri * nri * nzi);
    ALLOC(p->S, (size_t)mmax * (2 * nzi - 1) * nri);
    ALLOC(p->Stb, (size_t)CT_SPLIT * mmax * 4 * nri);
    ALLOC(p->Ss, (size_t)CT_SPLIT * mmax * 2 * nzi);
    const int nf = 2 * nri + nzi;
    ALLOC(p->F, (size_t)N * nf);
    ALLOC(p->FF, (size_t)N * nf);
    const int nh = N / 2 + 1, ns_ = N / 2 - 1;  // sizes of the cosine and sine halves
    ALLOC(p->Dc, (size_t)nh * nh);
    ALLOC(p->Ds, (size_t)ns_ * ns_);
    ALLOC(p->Ec, (size_t)nh * nh);
    ALLOC(p->Es, (size_t)ns_ * ns_);
    ALLOC(p->Pm, (size_t)(2 * mmax - 1) * N);
    ALLOC(p->besc, (size_t)mmax * N);
    ALLOC(p->bess, (size_t)mmax * N);
    ALLOC(p->alphar, numr + 1);
    ALLOC(p->betar, numr + 1);
    ALLOC(p->elzb, (size_t)N * (numr + 1));
    for (int w = 0; w < 2 && !mixed; ++w) {
        int ns = (int)p->dt[w].size();
        ALLOC(p->rc_alpha[w], (size_t)ns * N * nri);
        ALLOC(p->rc_ibet[w], (size_t)ns * N * nri);
        ALLOC(p->rc_gback[w], (size_t)ns * N * nri);
        ALLOC(p->zc_alpha[w], (size_t)ns * nzi);
        ALLOC(p->zc_ibet[w], (size_t)ns * nzi);
        ALLOC(p->zc_gback[w], (size_t)ns * nzi);
    }
    double *d_rhf = nullptr, *d_rhfinv = nullptr, *d_zhf = nullptr, *d_coefh = nullptr, *d_cser = nullptr, *d_nu = nullptr;
    const int nser = 111;  // hypr_upr_bnd + 1 (pot.h:7)
    ALLOC(d_rhf, numr + 1);
    ALLOC(d_rhfinv, numr + 1);
    ALLOC(d_zhf, numz + 1);
    ALLOC(d_coefh, mmax);
    ALLOC(d_cser, (size_t)mmax * nser);
    ALLOC(d_nu, mmax + 1);
#undef ALLOC

#define UPLOAD(dst, vec)                              \
    do {                                              \
        if ((rc = upload((dst), (vec))) != BH_OK) {   \
            free_all(p);                              \
            delete p;                                 \
            return rc;                                \
        }                                             \
    } while (0)

    // DFT matrices.  Forward rows in helmadi's mode order (helmadi.F90:131-156):
    // k = 0..N/2 -> sum x cos(2 pi k n/N); N/2+k -> -sum x sin(2 pi k n/N).
    // Then the projection rows of bessel (potsetup.F90:131-139): m = 0 (sum),
    // cos(m (L-1/2) dphi) for m = 1..mmax-1, then sin for m = 1..mmax-1.
    {
        // a_k = sum_n x_n cos(2 pi k n/N) = x_0 + (-1)^k x_{N/2} + sum_{n<N/2} (x_n + x_{N-n}) cos(...)
        // -b_k = -sum_n x_n sin(2 pi k n/N) = -sum_{n<N/2} (x_n - x_{N-n}) sin(...)
        std::vector<double> Dc((size_t)nh * nh), Ds((size_t)ns_ * ns_), Ec((size_t)nh * nh), Es((size_t)ns_ * ns_);
        auto cosr = [&](long a) { return std::cos(2.0 * pi * (double)(a % N) / (double)N); };
        auto sinr = [&](long a) { return std::sin(2.0 * pi * (double)(a % N) / (double)N); };
        for (int k = 0; k < nh; ++k)
            for (int n = 0; n < nh; ++n) Dc[(size_t)k * nh + n] = cosr((long)k * n);
        for (int k = 0; k < ns_; ++k)
            for (int n = 0; n < ns_; ++n) Ds[(size_t)k * ns_ + n] = -sinr((long)(k + 1) * (n + 1));
        // inverse (realft isign = -1 with the 2/N of helmadi.F90:386-392):
        // x_n = P_n - Q_n, x_{N-n} = P_n + Q_n
        for (int n = 0; n < nh; ++n)
            for (int k = 0; k < nh; ++k) {
                double v;
                if (k == 0) v = 1.0 / N;
                else if (k == N / 2) v = ((n % 2) ? -1.0 : 1.0) / N;
                else v = 2.0 * cosr((long)k * n) / N;
                Ec[(size_t)n * nh + k] = v;
            }
        for (int n = 0; n < ns_; ++n)
            for (int k = 0; k < ns_; ++k) Es[(size_t)n * ns_ + k] = 2.0 * sinr((long)(k + 1) * (n + 1)) / N;
        // projection rows of bessel (potsetup.F90:131-139): m = 0 (sum),
        // cos(m (L-1/2) dphi) for m = 1..mmax-1, then sin for m = 1..mmax-1
        std::vector<double> bc((size_t)mmax * N), bs((size_t)mmax * N), Pm((size_t)(2 * mmax - 1) * N);
        for (int M1 = 1; M1 <= mmax; ++M1)
            for (int L = 1; L <= N; ++L) {
                double a = 0.5 * g.dphi * ((double)M1 - 1.0) * (2.0 * (double)L - 1.0);
                bc[(size_t)(M1 - 1) * N + (L - 1)] = std::cos(a);
                bs[(size_t)(M1 - 1) * N + (L - 1)] = std::sin(a);
            }
        for (int n = 0; n < N; ++n) {
            Pm[n] = 1.0;
            for (int m = 1; m < mmax; ++m) {
                Pm[(size_t)m * N + n] = bc[(size_t)m * N + n];
                Pm[(size_t)(mmax - 1 + m) * N + n] = bs[(size_t)m * N + n];
            }
        }
        UPLOAD(p->Dc, Dc);
        UPLOAD(p->Ds, Ds);
        UPLOAD(p->Ec, Ec);
        UPLOAD(p->Es, Es);
        UPLOAD(p->Pm, Pm);
        UPLOAD(p->besc, bc);
        UPLOAD(p->bess, bs);
    }
    UPLOAD(p->alphar, alphar);
    UPLOAD(p->betar, betar);
    UPLOAD(p->elzb, elzb);
    // Thomas factors per pseudo-timestep (tridagr.F90:41-60, tridagz.F90:115-132)
    for (int w = 0; w < 2 && !mixed; ++w) {
        const std::vector<double> &dt = p->dt[w];
        int ns = (int)dt.size();
        std::vector<double> ra((size_t)ns * N * nri), ri((size_t)ns * N * nri), rg((size_t)ns * N * nri);
        std::vector<double> za((size_t)ns * nzi), zi((size_t)ns * nzi), zg((size_t)ns * nzi);
        std::vector<double> bet(numr + 1), gam(numr + 2);
        for (int s = 0; s < ns; ++s) {
            for (int L = 1; L <= N; ++L) {
                const double *b = &brb[(size_t)(L - 1) * (numr + 1)];
                bet[2] = b[2] + dt[s];
                for (int J = 3; J <= numr - 1; ++J) {
                    gam[J] = cr[J - 1] / bet[J - 1];
                    bet[J] = (b[J] + dt[s]) - ar[J] * gam[J];
                }
                size_t off = ((size_t)s * N + (L - 1)) * nri;
                for (int J = 2; J <= numr - 1; ++J) {
                    double ib = 1.0 / bet[J];
                    ra[off + J - 2] = (J == 2) ? 0.0 : -ar[J] * ib;
                    ri[off + J - 2] = ib;
                    rg[off + J - 2] = (J <= numr - 2) ? -gam[J + 1] : 0.0;
                }
            }
            std::vector<double> bz(numz + 1), betz(numz + 1), gamz(numz + 2);
            const double az = -g.gamma, cz = -g.gamma;
            for (int K = 2; K <= numz - 1; ++K) bz[K] = 2.0 * g.gamma + dt[s];
            betz[2] = bz[2];
            for (int K = 3; K <= numz - 1; ++K) {
                gamz[K] = cz / betz[K - 1];
                betz[K] = bz[K] - az * gamz[K];
            }
            for (int K = 2; K <= numz - 1; ++K) {
                double ib = 1.0 / betz[K];
                za[(size_t)s * nzi + K - 2] = (K == 2) ? 0.0 : -az * ib;
                zi[(size_t)s * nzi + K - 2] = ib;
                zg[(size_t)s * nzi + K - 2] = (K <= numz - 2) ? -gamz[K + 1] : 0.0;
            }
        }
        UPLOAD(p->rc_alpha[w], ra);
        UPLOAD(p->rc_ibet[w], ri);
        UPLOAD(p->rc_gback[w], rg);
        UPLOAD(p->zc_alpha[w], za);
        UPLOAD(p->zc_ibet[w], zi);
        UPLOAD(p->zc_gback[w], zg);
    }
    // Green's tables on the device (tm.F90:103-200, sm.F90:118-215)
    {
        std::vector<double> nu(mmax + 1, 0.0), coefh(mmax), cser((size_t)mmax * nser);
        for (int m = 3; m <= mmax; ++m) nu[m] = (double)m - 2.5;
        double glh = lanczos_lngamma(0.5);
        for (int m = 1; m <= mmax; ++m) {
            double mm = (double)m - 1.0;
            double aa = 0.5 * (mm + 1.5), bb = 0.5 * (mm + 0.5), cc = mm + 1.0;
            coefh[m - 1] = glh + lanczos_lngamma(mm + 0.5) - lanczos_lngamma(aa) - lanczos_lngamma(bb);
            for (int ir = 0; ir < nser; ++ir) {
                double ri = (double)ir;
                double dc = lanczos_lngamma(aa + ri) + lanczos_lngamma(bb + ri) - lanczos_lngamma(cc + ri) -
                            lanczos_lngamma(ri + 1.0);
                cser[(size_t)(m - 1) * nser + ir] = std::exp(dc);
            }
        }
        UPLOAD(d_rhf, rhf_g);
        UPLOAD(d_rhfinv, rhfinv_g);
        UPLOAD(d_zhf, zhf_g);
        UPLOAD(d_coefh, coefh);
        UPLOAD(d_cser, cser);
        UPLOAD(d_nu, nu);
        TableArgs ta;
        ta.numr = numr; ta.numz = numz; ta.mmax = mmax; ta.nser = nser;
        ta.hypr_cutoff = 1.025;  // pot.h:5
        ta.piinv = g.piinv; ta.dz = g.dz;
        ta.rhf_g = d_rhf; ta.rhfinv_g = d_rhfinv; ta.zhf_g = d_zhf;
        ta.coefh = d_coefh; ta.cser = d_cser; ta.nu = d_nu;
        launch_gen_top_table(ta, p->T, 0);
        launch_gen_side_table(ta, p->S, 0);
    }

    if (mixed) {
#define MALLOCF(ptr, n)                                              \
    do {                                                             \
        double *tmp_ = nullptr;                                      \
        if ((rc = dalloc(p, &tmp_, ((size_t)(n) + 1) / 2)) != BH_OK) { \
            free_all(p);                                             \
            delete p;                                                \
            return rc;                                               \
        }                                                            \
        (ptr) = (float *)tmp_;                                       \
    } while (0)
#define MALLOCD(ptr, n)                                      \
    do {                                                     \
        if ((rc = dalloc(p, &(ptr), (size_t)(n))) != BH_OK) { \
            free_all(p);                                     \
            delete p;                                        \
            return rc;                                       \
        }                                                    \
    } while (0)
        const size_t PI = (size_t)N * nri * nzi;  // interior values, all modes
        MALLOCF(p->Dc32, (size_t)nh * nh);
        MALLOCF(p->Ds32, (size_t)ns_ * ns_);
        MALLOCF(p->S32, (size_t)nzi * nzi);
        MALLOCF(p->mu32[0], nzi);
        MALLOCF(p->mu32[1], nzi);
        MALLOCF(p->Mr[0], (size_t)N * nri * nri);
        MALLOCF(p->Mr[1], (size_t)N * nri * nri);
        MALLOCF(p->tri_ibet, PI);
        MALLOCF(p->tri_gam, PI);
        MALLOCF(p->RH32, (size_t)M * N);
        MALLOCF(p->r32, PI);
        MALLOCF(p->RT, PI);
        MALLOCF(p->WT, PI);
        MALLOCF(p->CT, PI);
        MALLOCF(p->DT, PI);
        MALLOCF(p->D1, (size_t)M * N);
        MALLOCF(p->Wf, (size_t)M * N);
        MALLOCD(p->brb_d, (size_t)N * (numr + 1));
        UPLOAD(p->brb_d, brb);
        // FP32 forward DFT rows (helmadi's mode order), sine transform,
        // lambda_z, mu_z for both schedules
        {
            std::vector<float> Dc32((size_t)nh * nh), Ds32((size_t)ns_ * ns_), S((size_t)nzi * nzi);
            for (int k = 0; k < nh; ++k)
                for (int n = 0; n < nh; ++n)
                    Dc32[(size_t)k * nh + n] = (float)std::cos(2.0 * pi * (double)(((long)k * n) % N) / N);
            for (int k = 0; k < ns_; ++k)
                for (int n = 0; n < ns_; ++n)
                    Ds32[(size_t)k * ns_ + n] = (float)-std::sin(2.0 * pi * (double)(((long)(k + 1) * (n + 1)) % N) / N);
            const double sn = std::sqrt(2.0 / (nzi + 1));
            for (int a = 0; a < nzi; ++a)
                for (int b = 0; b < nzi; ++b) {
                    long ab = ((long)(a + 1) * (b + 1)) % (2 * (nzi + 1));
                    S[(size_t)a * nzi + b] = (float)(sn * std::sin(pi * (double)ab / (nzi + 1)));
                }
            if ((rc = upload_f(p->Dc32, Dc32)) != BH_OK || (rc = upload_f(p->Ds32, Ds32)) != BH_OK ||
                (rc = upload_f(p->S32, S)) != BH_OK) {
                free_all(p);
                delete p;
                return rc;
            }
            std::vector<double> lamz(nzi);
            for (int k = 0; k < nzi; ++k) lamz[k] = 2.0 * g.gamma * (1.0 - std::cos(pi * (k + 1) / (nzi + 1)));
            for (int w = 0; w < 2; ++w) {
                std::vector<float> mu(nzi);
                for (int k = 0; k < nzi; ++k) {
                    double m = 1.0;
                    for (double d : p->dt[w]) m *= (d - lamz[k]) / (d + lamz[k]);
                    mu[k] = (float)m;
                }
                if ((rc = upload_f(p->mu32[w], mu)) != BH_OK) {
                    free_all(p);
                    delete p;
                    return rc;
                }
            }
            // LU factors of B_r(L) + lambda_z(k), layout [L][J'][k]
            std::vector<float> ib(PI), gm(PI);
            for (int L = 0; L < N; ++L) {
                const double *b = &brb[(size_t)L * (numr + 1)];
                for (int k = 0; k < nzi; ++k) {
                    size_t base = (size_t)L * nri * nzi + k;
                    double bet = b[2] + lamz[k];
                    ib[base] = (float)(1.0 / bet);
                    gm[base] = 0.f;
                    for (int j = 1; j < nri; ++j) {
                        int J = j + 2;
                        double gg = alphar[J - 1] / bet;
                        bet = (b[J] + lamz[k]) - betar[J] * gg;
                        ib[base + (size_t)j * nzi] = (float)(1.0 / bet);
                        gm[base + (size_t)j * nzi] = (float)gg;
                    }
                }
            }
            if ((rc = upload_f(p->tri_ibet, ib)) != BH_OK || (rc = upload_f(p->tri_gam, gm)) != BH_OK) {
                free_all(p);
                delete p;
                return rc;
            }
        }
        // M_r(L) for both schedules, on the device in chunks of modes
        {
            const int chunk = 64;
            double *sv = nullptr, *su = nullptr, *sg = nullptr, *ddt = nullptr;
            size_t nsc = (size_t)chunk * nri * nri;
            const size_t ndt = std::max(p->dt[0].size(), p->dt[1].size());
            hipError_t e1 = hipMalloc(&sv, nsc * 8), e2 = hipMalloc(&su, nsc * 8), e3 = hipMalloc(&sg, nsc * 8),
                       e4 = hipMalloc(&ddt, ndt * 8);
            if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess || e4 != hipSuccess) {
                (void)hipFree(sv); (void)hipFree(su); (void)hipFree(sg); (void)hipFree(ddt);
                free_all(p);
                delete p;
                return fail(BH_E_NOMEM, "plan setup: scratch for M_r");
            }
            for (int w = 0; w < 2; ++w) {
                (void)hipMemcpy(ddt, p->dt[w].data(), p->dt[w].size() * 8, hipMemcpyHostToDevice);
                for (int m0 = 0; m0 < N; m0 += chunk) {
                    MrArgs ma;
                    ma.brb = p->brb_d; ma.alphar = p->alphar; ma.betar = p->betar; ma.dt = ddt;
                    ma.numr = numr; ma.nri = nri; ma.nsteps = (int)p->dt[w].size();
                    ma.mode0 = m0; ma.nmodes = std::min(chunk, N - m0);
                    ma.scratch_v = sv; ma.scratch_u = su; ma.scratch_g = sg; ma.Mr = p->Mr[w];
                    launch_build_mr(ma, 0);
                }
                (void)hipDeviceSynchronize();
            }
            (void)hipFree(sv); (void)hipFree(su); (void)hipFree(sg); (void)hipFree(ddt);
        }
        launch_fillf(p->D1, (size_t)M * N, 0.f, 0);
        launch_fillf(p->Wf, (size_t)M * N, 0.f, 0);
#undef MALLOCF
#undef MALLOCD
    }
#undef UPLOAD
    launch_fill(p->XJ, (size_t)M * N, 0.0, 0);
    launch_fill(p->XK, (size_t)M * N, 0.0, 0);
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) {
        free_all(p);
        delete p;
        return fail(BH_E_DEVICE, std::string("plan setup: ") + hipGetErrorString(e));
    }
    *out = p;
    return BH_OK;
}

extern "C" int bh_plan_destroy(bh_plan *p)
{
    if (!p) return BH_OK;
    (void)hipDeviceSynchronize();
    free_all(p);
    delete p;
    return BH_OK;
}

extern "C" size_t bh_plan_device_bytes(const bh_plan *p) { return p ? p->bytes : 0; }

extern "C" int bh_plan_set_timing(bh_plan *p, int on)
{
    if (!p) return fail(BH_E_INVALID, "null plan");
    if (on && !p->ev[0])
        for (auto &e : p->ev) HIPCHK(hipEventCreate(&e));
    p->timing = on != 0;
    return BH_OK;
}

extern "C" int bh_last_timings(const bh_plan *p, double *ms, int n)
{
    if (!p || !ms) return fail(BH_E_INVALID, "null argument");
    if (!p->have_timing) return fail(BH_E_STATE, "no timed solve yet");
    for (int i = 0; i < n && i < 5; ++i) ms[i] = p->last_ms[i];
    return BH_OK;
}

namespace {


// Forward real DFT along phi of the N rows of B (columns Mcols), written in
// helmadi's mode order (helmadi.F90:131-156) to C; cosine and sine halves.
void dft_fwd64(const bh_plan *p, const double *B, long ldb, int Mcols, double *C, long ldc, bool mask, hipStream_t s)
{
    const int N = p->g.numphi, nh = N / 2 + 1, nsn = N / 2 - 1;
    GemmArgs a{};
    a.A = p->Dc; a.lda = nh; a.B = B; a.ldb = ldb; a.C = C; a.ldc = ldc;
    a.R = nh; a.M = Mcols; a.Kd = nh; a.batch = 1; a.nfold = N;
    a.mask_numr = p->g.numr; a.mask_numz = p->g.numz;
    a.tile = Mcols >= 4096 ? p->tile_dft : 2;
    gemm64(a, B_FOLD_PLUS, mask ? E_MASK : E_PLAIN, s);
    a.A = p->Ds; a.lda = nsn; a.C = C + (long)nh * ldc; a.R = nsn; a.Kd = nsn;
    gemm64(a, B_FOLD_MINUS, mask ? E_MASK : E_PLAIN, s);
}

// Inverse real DFT with the 2/N normalisation (helmadi.F90:357-392) of the
// N mode rows of X (+ Xadd in FP32 when given) into N phi rows of out.
void dft_inv64(const bh_plan *p, const double *X, const float *Xadd, long ldx, int Mcols, double *out, long ldo,
               hipStream_t s)
{
    const int N = p->g.numphi, nh = N / 2 + 1, nsn = N / 2 - 1;
    GemmArgs a{};
    a.A = p->Ec; a.lda = nh; a.B = X; a.ldb = ldx; a.Badd = Xadd; a.C = out; a.ldc = ldo;
    a.R = nh; a.M = Mcols; a.Kd = nh; a.batch = 1; a.nfold = N;
    a.tile = Mcols >= 4096 ? p->tile_dft : 2;
    gemm64(a, Xadd ? B_ADD_F32 : B_PLAIN, E_PLAIN, s);
    a.A = p->Es; a.lda = nsn; a.B = X + (long)nh * ldx; a.Badd = Xadd ? Xadd + (long)nh * ldx : nullptr;
    a.R = nsn; a.Kd = nsn;
    gemm64(a, Xadd ? B_ADD_F32 : B_PLAIN, E_INV_COMBINE, s);
}

// Boundary projections of rho onto the mmax azimuthal modes (bessel.F90:179-185, 228-237)
void proj64(const bh_plan *p, const double *rho, hipStream_t s)
{
    const int N = p->g.numphi;
    const long M = (long)p->g.numr * p->g.numz;
    GemmArgs a{};
    a.A = p->Pm; a.lda = N; a.B = rho; a.ldb = M; a.C = p->proj; a.ldc = M;
    a.R = 2 * p->g.mmax - 1; a.M = (int)M; a.Kd = N; a.batch = 1;
    a.tile = M >= 4096 ? p->tile_dft : 2;
    gemm64(a, B_PLAIN, E_PLAIN, s);
}

// The mixed-precision solve.  Evaluates the same truncated Peaceman-Rachford
// iteration as the FP64 path (helmadi.F90:195-355), as A x = g - M r0 (see
// bh_mixed.hip), with FP32 MFMA transforms and FP64 residuals.
int enqueue_solve_mixed(bh_plan *p, const double *d_rho, double *d_pot, int tstep, hipStream_t s)
{
    const Geo &g = p->g;
    const int numr = g.numr, numz = g.numz, N = g.numphi, mmax = g.mmax, nri = g.nri, nzi = g.nzi;
    const long M = (long)numr * numz;
    const long PIm = (long)nri * nzi;  // interior values per mode
    const int w = (tstep > 1) ? 1 : 0;
    if (p->timing) HIPCHK(hipEventRecord(p->ev[0], s));
    // 1. rho-hat in FP32 (masked as rhop), projections in FP64
    {
        const int nh = N / 2 + 1, nsn = N / 2 - 1;
        GemmArgs a{};
        a.A = p->Dc32; a.lda = nh; a.B = d_rho; a.ldb = M; a.C = p->RH32; a.ldc = M;
        a.R = nh; a.M = (int)M; a.Kd = nh; a.batch = 1; a.nfold = N;
        a.mask_numr = numr; a.mask_numz = numz;
        a.tile = M >= 4096 ? p->tile_dft : 2;
        gemm32(a, B_FOLD_PLUS, E_MASK, false, false, true, s);
        a.A = p->Ds32; a.lda = nsn; a.C = p->RH32 + (long)nh * M; a.R = nsn; a.Kd = nsn;
        gemm32(a, B_FOLD_MINUS, E_MASK, false, false, true, s);
        proj64(p, d_rho, s);
    }
    if (p->timing) HIPCHK(hipEventRecord(p->ev[1], s));
    // 2. boundary faces (FP64), into both potential buffers
    launch_contract_topbot(p->T, p->proj, p->Stb, numr, numz, mmax, s);
    launch_contract_side(p->S, p->proj, p->Ss, numr, numz, mmax, s);
    launch_face_rebuild(p->Stb, p->Ss, p->besc, p->bess, p->F, numr, numz, N, mmax, -g.dr * g.dz * g.dphi, s);
    const int nf = 2 * nri + nzi;
    dft_fwd64(p, p->F, nf, nf, p->FF, nf, false, s);
    double *Xa = p->XJ, *Xb = p->XK;
    if (tstep <= 1) launch_fill(Xa, (size_t)M * N, 0.0, s);
    launch_face_scatter_j(p->FF, Xa, numr, numz, N, s);
    launch_face_scatter_j(p->FF, Xb, numr, numz, N, s);
    if (p->timing) HIPCHK(hipEventRecord(p->ev[2], s));
    // 3. r0 = g - A x0
    ResidualArgs ra{};
    ra.brb = p->brb_d; ra.alphar = p->alphar; ra.betar = p->betar;
    ra.gamma = g.gamma; ra.four_pi = g.four_pi; ra.numr = numr; ra.numz = numz;
    ra.rh = p->RH32;
    ra.X = Xa; ra.dlt = nullptr; ra.w = nullptr; ra.Xo = nullptr; ra.r = p->r32;
    launch_residual64(ra, N, s);
    // sine transform: RT[L] (nri x nzi) = r[L]^T S
    auto dst_fwd = [&](const float *r, float *out) {
        GemmArgs a{};
        a.A = r; a.lda = nri; a.sA = PIm;
        a.B = p->S32; a.ldb = nzi; a.sB = 0;
        a.C = out; a.ldc = nzi; a.sC = PIm;
        a.R = nri; a.M = nzi; a.Kd = nzi; a.batch = N;
        a.tile = p->tile_spec;
        gemm32(a, B_PLAIN, E_PLAIN, true, false, false, s);
    };
    // inverse: full-plane out[L] interior (nzi x nri) = S * in[L]^T
    auto dst_inv = [&](const float *in, float *out) {
        GemmArgs a{};
        a.A = p->S32; a.lda = nzi; a.sA = 0;
        a.B = in; a.ldb = nzi; a.sB = PIm;
        a.C = out + numr + 1; a.ldc = numr; a.sC = M;
        a.R = nzi; a.M = nri; a.Kd = nzi; a.batch = N;
        a.tile = p->tile_spec;
        gemm32(a, B_PLAIN, E_PLAIN, false, true, false, s);
    };
    dst_fwd(p->r32, p->RT);
    // 4. W = mu_z . (M_r RT), C = RT - W
    {
        GemmArgs a{};
        a.A = p->Mr[w]; a.lda = nri; a.sA = (long)nri * nri;
        a.B = p->RT; a.ldb = nzi; a.sB = PIm;
        a.C = p->WT; a.ldc = nzi; a.sC = PIm;
        a.R = nri; a.M = nzi; a.Kd = nri; a.batch = N;
        a.mu = p->mu32[w]; a.Rin = p->RT; a.C2 = p->CT;
        a.tile = p->tile_spec;
        gemm32(a, B_PLAIN, E_FILTER, false, false, false, s);
    }
    // 5. (B_r + lambda_z) dhat = C, back to physical space
    launch_tridiag_shift(p->CT, p->DT, p->tri_ibet, p->tri_gam, p->betar, nri, nzi, N, s);
    dst_inv(p->DT, p->D1);
    dst_inv(p->WT, p->Wf);
    // 6. refinement: x += delta, r = g - w - A x (FP64), solve again
    double *Xcur = Xa, *Xnext = Xb;
    for (int it = 0; it < p->passes; ++it) {
        ra.X = Xcur; ra.dlt = p->D1; ra.w = p->Wf; ra.Xo = Xnext; ra.r = p->r32;
        launch_residual64(ra, N, s);
        dst_fwd(p->r32, p->RT);
        launch_tridiag_shift(p->RT, p->DT, p->tri_ibet, p->tri_gam, p->betar, nri, nzi, N, s);
        dst_inv(p->DT, p->D1);
        std::swap(Xcur, Xnext);
    }
    // final update into Xa, the warm-start state (pointwise, so in place is safe)
    launch_add_update(Xcur, p->D1, Xa, numr, numz, N, s);
    if (p->timing) HIPCHK(hipEventRecord(p->ev[3], s));
    // 7. inverse DFT (FP64) and axis ghost
    dft_inv64(p, Xa, nullptr, M, (int)M, d_pot, M, s);
    launch_axis(d_pot, numr, numz, N, s);
    if (p->timing) HIPCHK(hipEventRecord(p->ev[4], s));
    HIPCHK(hipGetLastError());
    p->has_solution = true;
    return BH_OK;
}

// The solve sequence (potential_solver.F90:58-106 with bessel and helmadi).
int enqueue_solve(bh_plan *p, const double *d_rho, double *d_pot, int tstep, hipStream_t s)
{
    if (p->precision == BH_PREC_MIXED) return enqueue_solve_mixed(p, d_rho, d_pot, tstep, s);
    const Geo &g = p->g;
    const int numr = g.numr, numz = g.numz, N = g.numphi, mmax = g.mmax, nri = g.nri, nzi = g.nzi;
    const long M = (long)numr * numz;
    const bool warm = (tstep > 1) && true;
    const int w = warm ? 1 : 0;
    if (p->timing) HIPCHK(hipEventRecord(p->ev[0], s));
    // 1. one product gives rho-hat (with rhop's zeroed layers) and the
    //    boundary projections of rho (bessel.F90:179-185, 228-237)
    dft_fwd64(p, d_rho, M, (int)M, p->RJ, M, true, s);
    proj64(p, d_rho, s);
    launch_transpose_jk(p->RJ, p->RK, numr, numz, N, s);
    if (p->timing) HIPCHK(hipEventRecord(p->ev[1], s));
    // 2. boundary potential (bessel.F90:186-488) -> Fourier-space faces
    launch_contract_topbot(p->T, p->proj, p->Stb, numr, numz, mmax, s);
    launch_contract_side(p->S, p->proj, p->Ss, numr, numz, mmax, s);
    const double factor = -g.dr * g.dz * g.dphi;
    launch_face_rebuild(p->Stb, p->Ss, p->besc, p->bess, p->F, numr, numz, N, mmax, factor, s);
    const int nf = 2 * nri + nzi;
    dft_fwd64(p, p->F, nf, nf, p->FF, nf, false, s);
    // 3. starting potential: zero (cold, potential_solver.F90:88-93) or the
    //    previous solution kept in Fourier space (warm)
    if (!warm) launch_fill(p->XJ, (size_t)M * N, 0.0, s);
    launch_face_scatter(p->FF, p->XJ, p->XK, numr, numz, N, s);
    if (p->timing) HIPCHK(hipEventRecord(p->ev[2], s));
    // 4. Peaceman-Rachford iterations (helmadi.F90:195-355)
    const std::vector<double> &dt = p->dt[w];
    for (int st = 0; st < (int)dt.size(); ++st) {
        LineArgs r{};
        r.xin = p->XJ; r.xout = p->XK; r.rh = p->RJ;
        r.SA = M; r.PA = numr; r.SB = M; r.PB = numz;
        r.nlines = nzi; r.len = nri;
        r.alpha = p->rc_alpha[w] + (size_t)st * N * nri;
        r.ibet = p->rc_ibet[w] + (size_t)st * N * nri;
        r.gback = p->rc_gback[w] + (size_t)st * N * nri;
        r.coef_stride = nri;
        r.dir = 0; r.dt = dt[st]; r.gamma = g.gamma; r.four_pi = g.four_pi;
        r.alphar = p->alphar; r.betar = p->betar; r.elzb = p->elzb; r.numr = numr;
        launch_line_sweep(r, N, s);
        LineArgs z = r;
        z.xin = p->XK; z.xout = p->XJ; z.rh = p->RK;
        z.PA = numz; z.PB = numr;
        z.nlines = nri; z.len = nzi;
        z.alpha = p->zc_alpha[w] + (size_t)st * nzi;
        z.ibet = p->zc_ibet[w] + (size_t)st * nzi;
        z.gback = p->zc_gback[w] + (size_t)st * nzi;
        z.coef_stride = 0;
        z.dir = 1;
        launch_line_sweep(z, N, s);
    }
    if (p->timing) HIPCHK(hipEventRecord(p->ev[3], s));
    // 5. inverse DFT with the 2/N normalisation, then the axis ghost
    dft_inv64(p, p->XJ, nullptr, M, (int)M, d_pot, M, s);
    launch_axis(d_pot, numr, numz, N, s);
    if (p->timing) HIPCHK(hipEventRecord(p->ev[4], s));
    HIPCHK(hipGetLastError());
    p->has_solution = true;
    return BH_OK;
}

int collect_timing(bh_plan *p)
{
    if (!p->timing) return BH_OK;
    HIPCHK(hipEventSynchronize(p->ev[4]));
    float t[4];
    for (int i = 0; i < 4; ++i) HIPCHK(hipEventElapsedTime(&t[i], p->ev[i], p->ev[i + 1]));
    // order: boundary, forward DFT, ADI, inverse DFT, total
    p->last_ms[0] = t[1];
    p->last_ms[1] = t[0];
    p->last_ms[2] = t[2];
    p->last_ms[3] = t[3];
    float tot;
    HIPCHK(hipEventElapsedTime(&tot, p->ev[0], p->ev[4]));
    p->last_ms[4] = tot;
    p->have_timing = true;
    return BH_OK;
}

}  // namespace

extern "C" int bh_solve_device(bh_plan *p, const double *d_rho, double *d_pot, int tstep, void *stream)
{
    if (!p || !d_rho || !d_pot) return fail(BH_E_INVALID, "null argument");
    int rc = enqueue_solve(p, d_rho, d_pot, tstep, (hipStream_t)stream);
    if (rc != BH_OK) return rc;
    return BH_OK;
}

extern "C" int bh_solve(bh_plan *p, const double *rho, double *pot, int tstep)
{
    if (!p || !rho || !pot) return fail(BH_E_INVALID, "null argument");
    const size_t n = (size_t)p->g.numr * p->g.numz * p->g.numphi;
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(rho[i])) return fail(BH_E_NONFINITE, "rho holds a non-finite value at index " + std::to_string(i));
    if (!p->rho_in) {
        int rc;
        if ((rc = dalloc(p, &p->rho_in, n)) != BH_OK) return rc;
        if ((rc = dalloc(p, &p->pot_out, n)) != BH_OK) return rc;
    }
    HIPCHK(hipMemcpy(p->rho_in, rho, n * sizeof(double), hipMemcpyHostToDevice));
    int rc = enqueue_solve(p, p->rho_in, p->pot_out, tstep, 0);
    if (rc != BH_OK) return rc;
    HIPCHK(hipMemcpy(pot, p->pot_out, n * sizeof(double), hipMemcpyDeviceToHost));
    return collect_timing(p);
}

extern "C" int bh_set_potential(bh_plan *p, const double *pot, int on_device)
{
    if (!p || !pot) return fail(BH_E_INVALID, "null argument");
    const Geo &g = p->g;
    const long M = (long)g.numr * g.numz;
    const size_t n = (size_t)M * g.numphi;
    const double *src = pot;
    if (!on_device) {
        if (!p->rho_in) {
            int rc;
            if ((rc = dalloc(p, &p->rho_in, n)) != BH_OK) return rc;
            if ((rc = dalloc(p, &p->pot_out, n)) != BH_OK) return rc;
        }
        HIPCHK(hipMemcpy(p->pot_out, pot, n * sizeof(double), hipMemcpyHostToDevice));
        src = p->pot_out;
    }
    dft_fwd64(p, src, M, (int)M, p->XJ, M, false, 0);
    HIPCHK(hipDeviceSynchronize());
    p->has_solution = true;
    return BH_OK;
}

extern "C" int bh_sync_timing(bh_plan *p) { return p ? collect_timing(p) : fail(BH_E_INVALID, "null plan"); }

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

// which: 0 RJ, 1 RK, 2 XJ, 3 XK, 4 proj, 5 T, 6 S, 7 Stb, 8 Ss, 9 F, 10 FF
extern "C" int bhx_snapshot(const bh_plan *p, int which, double *host, size_t n)
{
    if (!p || !host) return fail(BH_E_INVALID, "null argument");
    const Geo &g = p->g;
    const size_t M = (size_t)g.numr * g.numz, N = g.numphi, mm = g.mmax, nri = g.nri, nzi = g.nzi;
    const size_t nf = 2 * nri + nzi;
    const double *src[] = {p->RJ, p->RK, p->XJ, p->XK, p->proj, p->T, p->S, p->Stb, p->Ss, p->F, p->FF};
    const size_t cnt[] = {M * N, M * N, M * N, M * N, M * (2 * mm - 1), mm * nri * nri * nzi,
                          mm * (2 * nzi - 1) * nri, CT_SPLIT * mm * 4 * nri, CT_SPLIT * mm * 2 * nzi, N * nf, N * nf};
    if (which < 0 || which > 10) return fail(BH_E_INVALID, "which");
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
    if (inverse) dft_inv64(p, in, nullptr, ncols, ncols, out, ncols, 0);
    else dft_fwd64(p, in, ncols, ncols, out, ncols, false, 0);
    HIPCHK(hipDeviceSynchronize());
    return BH_OK;
}

// FP64 / FP32 GEMM with an explicit tile shape (0 wide, 1 square, 2 small)
// and transposes (FP32 only); test hook.
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
