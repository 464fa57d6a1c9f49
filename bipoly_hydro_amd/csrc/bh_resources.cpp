// bh_resources.cpp -- grid geometry, and the per-geometry device resources
// (the potsetup counterpart: Green's tables, transform matrices, ADI
// operator data) kept in a per-process cache so that plans of the same grid
// share them.
#include "bh_plan_impl.h"
#include "bh_dft.h"
#include "bh_mixed.h"
#include "bh_spectral.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <mutex>

namespace bh {

namespace {
thread_local std::string g_err;
}

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

const char *last_error() { return g_err.c_str(); }

DevMem::~DevMem()
{
    for (void *p : bufs) (void)hipFree(p);
}

int DevMem::alloc_bytes(void **out, size_t nbytes)
{
    void *ptr = nullptr;
    hipError_t e = hipMalloc(&ptr, nbytes ? nbytes : 8);
    if (e != hipSuccess) {
        *out = nullptr;
        return fail(BH_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    bufs.push_back(ptr);
    bytes += nbytes;
    *out = ptr;
    return BH_OK;
}

int normalize_desc(const bh_grid_desc *desc, int precision, bh_grid_desc *out)
{
    if (!desc) return fail(BH_E_INVALID, "null argument");
    bh_grid_desc d = *desc;
    if (d.isym == 0) d.isym = 1;
    if (d.mmax == 0) d.mmax = 30;                     // pot.h:3
    if (d.dr == 0.0) d.dr = 1.0 / (130.0 - 3.0);      // oldnumr = 130 (runhydro.h:4)
    if (d.nsteps_cold == 0) d.nsteps_cold = 20;       // potential_solver.F90:88-93
    if (d.nsteps_warm == 0) d.nsteps_warm = 5;
    if (d.refine_passes == 0) d.refine_passes = 2;
    if (d.isym < 1 || d.isym > 3) return fail(BH_E_INVALID, "isym must be 1, 2 or 3");
    if (d.isym != 1 && (is_mixed(precision) || is_spec64(precision)))
        return fail(BH_E_INVALID, "isym = 2, 3 need precision BH_PREC_FP64 (the spectral solves' sine basis in z "
                                  "is the one of two Dirichlet faces)");
    if (d.numr < 5 || d.numz < 5) return fail(BH_E_INVALID, "numr and numz must be >= 5");
    if (d.numphi < 4 || (d.numphi & (d.numphi - 1))) return fail(BH_E_INVALID, "numphi must be a power of two >= 4");
    if (d.isym == 1 && d.numz % 2) return fail(BH_E_INVALID, "numz must be even for isym = 1 (symmetric z grid)");
    if (d.mmax < 1 || d.mmax > 512) return fail(BH_E_INVALID, "mmax must be in [1, 512]");
    if (d.numr - 2 > 512 || d.numz - 2 > 512) return fail(BH_E_INVALID, "numr-2 and numz-2 must be <= 512");
    if (d.nsteps_cold < 2 || d.nsteps_warm < 2) return fail(BH_E_INVALID, "nsteps must be >= 2");
    if (!(d.dr > 0.0) || !std::isfinite(d.dr)) return fail(BH_E_INVALID, "dr must be positive");
    if (d.refine_passes < 1 || d.refine_passes > 4) return fail(BH_E_INVALID, "refine_passes must be in [1, 4]");
    if (!valid_precision(precision))
        return fail(BH_E_INVALID, "precision must be BH_PREC_FP64, BH_PREC_MIXED, BH_PREC_MIXED_TABLE32 or "
                                  "BH_PREC_SPECTRAL64");
    *out = d;
    return BH_OK;
}

void host_geometry(const bh_grid_desc &d, HostGeo *h)
{
    h->d = d;
    const int numr = d.numr, numz = d.numz, N = d.numphi, mmax = d.mmax;
    Geo &g = h->g;
    g.numr = numr; g.numz = numz; g.numphi = N; g.mmax = mmax; g.nri = numr - 2; g.nzi = numz - 2;
    g.isym = d.isym;
    // constants and grid: setup_frac.F90:345, 374-379, 428-463; potsetup.F90:113-114, 151-153
    const double pi = std::acos(-1.0);
    g.dr = d.dr; g.dz = d.dr;
    g.dphi = 2.0 * pi * (1.0 / (double)N);
    g.piinv = 1.0 / pi;
    g.four_pi = 4.0 * pi;
    const double drinv = 1.0 / g.dr, dzinv = 1.0 / g.dz, dphiinv = 1.0 / g.dphi;
    g.gamma = dzinv * dzinv;
    const double drinv2 = drinv * drinv, dphiinv2 = dphiinv * dphiinv;
    h->r_g.assign(numr + 1, 0.0); h->rhf_g.assign(numr + 1, 0.0); h->rinv_g.assign(numr + 1, 0.0);
    h->rhfinv_g.assign(numr + 1, 0.0); h->zhf_g.assign(numz + 1, 0.0);
    for (int J = 1; J <= numr; ++J) {
        double x = (double)J;
        h->r_g[J] = (x - 2.0) * g.dr;
        h->rhf_g[J] = h->r_g[J] + 0.5 * g.dr;
        h->rinv_g[J] = (h->r_g[J] != 0.0) ? 1.0 / h->r_g[J] : 0.0;
        h->rhfinv_g[J] = 1.0 / h->rhf_g[J];
    }
    if (d.isym != 1) {  // the grid starts at the equator; K = 1 mirrors K = 2 (setup_frac.F90:452-456)
        for (int K = 1; K <= numz; ++K) h->zhf_g[K] = ((double)K - 1.5) * g.dz;
    } else {
        double offset = -((double)(numz / 2) * g.dz);
        for (int K = 1; K <= numz; ++K) h->zhf_g[K] = offset + ((double)K - 0.5) * g.dz;
    }
    // ADI coefficients: potsetup.F90:155-252 (one block: global radial index)
    std::vector<double> &alphar = h->alphar, &betar = h->betar;
    alphar.assign(numr + 1, 0.0);
    betar.assign(numr + 1, 0.0);
    for (int J = 2; J <= numr - 1; ++J) {
        alphar[J] = -h->r_g[J + 1] * h->rhfinv_g[J] * drinv2;
        betar[J] = -h->r_g[J] * h->rhfinv_g[J] * drinv2;
    }
    h->brb.assign((size_t)N * (numr + 1), 0.0);
    h->elzb.assign((size_t)N * (numr + 1), 0.0);
    for (int L = 1; L <= N; ++L) {
        int lstop = N / 2 + 1;
        int mode = (L <= lstop) ? L - 1 : L - lstop;
        if (d.isym == 3) mode = mode * mode;  // potsetup.F90:158-163 squares the index for isym = 3
        double m1mode = (mode % 2 == 0) ? 1.0 : -1.0;
        double elm = std::cos((double)mode * g.dphi);
        double *b = &h->brb[(size_t)(L - 1) * (numr + 1)], *e = &h->elzb[(size_t)(L - 1) * (numr + 1)];
        const std::vector<double> &rhfinv = h->rhfinv_g;
        for (int J = 3; J <= numr - 1; ++J)
            b[J] = 2.0 * drinv2 - 2.0 * (elm - 1.0) * dphiinv2 * rhfinv[J] * rhfinv[J];
        // the axis row (potsetup.F90:211-223, 239-251): the ghost J = 1 is
        // the cell rotated by pi, or for isym = 3 the cell itself
        if (d.isym == 3) b[2] = -alphar[2] - 2.0 * betar[2] - 2.0 * (elm - 1.0) * dphiinv2 * rhfinv[2] * rhfinv[2];
        else b[2] = -alphar[2] + (m1mode - 1.0) * betar[2] - 2.0 * (elm - 1.0) * dphiinv2 * rhfinv[2] * rhfinv[2];
        for (int J = 2; J <= numr - 1; ++J)
            e[J] = -2.0 * drinv2 + 2.0 * (elm - 1.0) * dphiinv2 * rhfinv[J] * rhfinv[J];
        if (d.isym == 3) e[2] = alphar[2] + 2.0 * betar[2] + 2.0 * (elm - 1.0) * dphiinv2 * rhfinv[2] * rhfinv[2];
        else e[2] = alphar[2] - (m1mode - 1.0) * betar[2] + 2.0 * (elm - 1.0) * dphiinv2 * rhfinv[2] * rhfinv[2];
    }
    // pseudo-timesteps: helmadi.F90:124-129
    for (int w = 0; w < 2; ++w) {
        int ns = w ? d.nsteps_warm : d.nsteps_cold;
        std::vector<double> &dt = h->dt[w];
        dt.assign(ns, 0.0);
        dt[ns - 1] = 4.0 * (h->rinv_g[numr] * h->rinv_g[numr]);
        double alph = std::pow(h->r_g[numr] * h->rinv_g[3], 2.0 / ((double)ns - 1.0));
        for (int i = ns - 2; i >= 0; --i) dt[i] = alph * dt[i + 1];
    }
}

namespace {

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

// A private non-blocking stream for set-up work, so that building a plan
// does not wait on (or for) solves running on other streams.
struct SetupStream {
    hipStream_t s = nullptr;
    SetupStream() { (void)hipStreamCreateWithFlags(&s, hipStreamNonBlocking); }
    ~SetupStream()
    {
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    }
};

template <class T>
int upload(T *dst, const std::vector<T> &src, hipStream_t s)
{
    HIPCHK(hipMemcpyAsync(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return BH_OK;
}

std::mutex g_cache_mu;
std::map<std::string, std::weak_ptr<const GeoTables>> g_tables;
std::map<std::string, std::weak_ptr<const ModeOps>> g_modeops;

// the opt-in variant contracts an FP32 copy of the top table: its rounding
// (2^-24 relative per entry, independent) moves the potential by about
// 1e-10 relative and halves the bytes of the HBM-bound contraction; every
// other plan keeps the table in FP64
bool table32(int precision) { return precision == BH_PREC_MIXED_TABLE32; }
// the transform and operator data do not depend on the table precision
int ops_precision(int precision) { return is_mixed(precision) ? BH_PREC_MIXED : precision; }

std::string tables_key(const bh_grid_desc &d, int precision, int device)
{
    char b[160];
    std::snprintf(b, sizeof b, "%d|%d|%d|%.17g|%d|%d|%d", device, d.numr, d.numz, d.dr, d.mmax,
                  table32(precision) ? 32 : 64, d.isym);
    return b;
}

std::string modeops_key(const bh_grid_desc &d, int precision, int device)
{
    char b[200];
    std::snprintf(b, sizeof b, "%d|%d|%d|%d|%.17g|%d|%d|%d|%d|%d|%d", device, d.numr, d.numz, d.numphi, d.dr, d.mmax,
                  d.nsteps_cold, d.nsteps_warm, ops_precision(precision),
                  is_mixed(precision) && spectral_fused_ok(d.numr - 2, d.numz - 2) ? 1 : 0, d.isym);
    return b;
}

size_t tables_bytes(const bh_grid_desc &d, int precision)
{
    const size_t nri = d.numr - 2, nzi = d.numz - 2, mm = d.mmax;
    return (table32(precision) ? 4 : 8) * topsym_table_elems(d.numr, d.numz, d.mmax) +
           8 * ((d.isym == 1 ? 1 : 2) * mm * (2 * nzi - 1) * nri + d.numr + 1);
}

size_t modeops_bytes(const bh_grid_desc &d, int precision)
{
    const size_t numr = d.numr, N = d.numphi, mm = d.mmax, nri = numr - 2, nzi = d.numz - 2;
    const size_t nh = N / 2 + 1, ns = N / 2 - 1;
    size_t b = 8 * (2 * nh * nh + 2 * ns * ns + (size_t)proj_coef_elems((int)N, (int)mm) + 2 * mm * N + 2 * (numr + 1) + N * (numr + 1));
    if (N <= 512) {
        const size_t kg = (N / 2 + 14) / 16, rt = (N / 2 + 15) / 16 + (N / 2 + 14) / 16;
        b += 8 * 2 * rt * kg * 256 + (is_mixed(precision) ? 4 * rt * kg * 256 : 0);
    }
    if (is_mixed(precision)) {
        b += 4 * (nh * nh + ns * ns + (nzi + 1) / 2 * nzi + 2 * nzi + 2 * N * nri * nri + 2 * N * nri * nzi);
        if (spectral_fused_ok((int)nri, (int)nzi))  // tables; no g; M_r and pivots per mode number
            b += 4 * (nzi >= 64 ? 4 : 2) * (nzi / 2) * nzi - 4 * N * nri * nzi -
                 4 * (N / 2 - 1) * (2 * nri * nri + nri * nzi);
        b += 8 * N * (numr + 1);
    } else if (is_spec64(precision)) {
        b += 8 * (nzi * nzi + 3 * nzi + 2 * nh * nri * nri + N * (numr + 1));
    } else {
        b += 8 * 3 * ((size_t)d.nsteps_cold + d.nsteps_warm) * (N * nri + nzi);
    }
    return b;
}

int build_tables(const HostGeo &h, bool f32, GeoTables *t)
{
    const int numr = h.g.numr, numz = h.g.numz, mmax = h.g.mmax, nri = h.g.nri, nzi = h.g.nzi;
    int rc;
    if ((rc = f32 ? t->mem.alloc(&t->T32, topsym_table_elems(numr, numz, mmax))
                  : t->mem.alloc(&t->T, topsym_table_elems(numr, numz, mmax))) != BH_OK)
        return rc;
    if ((rc = t->mem.alloc(&t->S, (size_t)mmax * (2 * nzi - 1) * nri)) != BH_OK) return rc;
    if (h.g.isym != 1 && (rc = t->mem.alloc(&t->H, (size_t)mmax * (2 * nzi - 1) * nri)) != BH_OK) return rc;
    if ((rc = t->mem.alloc(&t->sqrt_rhf, numr + 1)) != BH_OK) return rc;
    // Green's tables on the device (tm.F90:103-363, sm.F90:118-382); for
    // isym = 3 slot m holds the even mode 2m (tm.F90:105-133)
    const int nser = 111;  // hypr_upr_bnd + 1 (pot.h:7)
    const bool even = h.g.isym == 3;
    // the series and the upward recurrence serve the reference's 30 slots;
    // slots beyond take the downward ratios (d_qhigh_store)
    const int mref = std::min(mmax, 30);
    std::vector<double> nu(mmax + 1, 0.0), coefh(mmax), cser((size_t)mmax * nser);
    for (int m = 3; m <= mref; ++m) nu[m] = even ? 2.0 * (double)m - 4.5 : (double)m - 2.5;
    double glh = lanczos_lngamma(0.5);
    for (int m = 1; m <= mref; ++m) {
        double mm = even ? 2.0 * ((double)m - 1.0) : (double)m - 1.0;
        double aa = 0.5 * (mm + 1.5), bb = 0.5 * (mm + 0.5), cc = mm + 1.0;
        coefh[m - 1] = glh + lanczos_lngamma(mm + 0.5) - lanczos_lngamma(aa) - lanczos_lngamma(bb);
        for (int ir = 0; ir < nser; ++ir) {
            double ri = (double)ir;
            double dc = lanczos_lngamma(aa + ri) + lanczos_lngamma(bb + ri) - lanczos_lngamma(cc + ri) -
                        lanczos_lngamma(ri + 1.0);
            cser[(size_t)(m - 1) * nser + ir] = std::exp(dc);
        }
    }
    DevMem tmp;
    double *d_rhf, *d_rhfinv, *d_zhf, *d_coefh, *d_cser, *d_nu;
    if ((rc = tmp.alloc(&d_rhf, numr + 1)) != BH_OK || (rc = tmp.alloc(&d_rhfinv, numr + 1)) != BH_OK ||
        (rc = tmp.alloc(&d_zhf, numz + 1)) != BH_OK || (rc = tmp.alloc(&d_coefh, mmax)) != BH_OK ||
        (rc = tmp.alloc(&d_cser, (size_t)mmax * nser)) != BH_OK || (rc = tmp.alloc(&d_nu, mmax + 1)) != BH_OK)
        return rc;
    SetupStream ss;
    if (!ss.s) return fail(BH_E_DEVICE, "plan setup: stream creation failed");
    if ((rc = upload(d_rhf, h.rhf_g, ss.s)) != BH_OK || (rc = upload(d_rhfinv, h.rhfinv_g, ss.s)) != BH_OK ||
        (rc = upload(d_zhf, h.zhf_g, ss.s)) != BH_OK || (rc = upload(d_coefh, coefh, ss.s)) != BH_OK ||
        (rc = upload(d_cser, cser, ss.s)) != BH_OK || (rc = upload(d_nu, nu, ss.s)) != BH_OK)
        return rc;
    TableArgs ta;
    ta.numr = numr; ta.numz = numz; ta.mmax = mmax; ta.nser = nser; ta.isym = h.g.isym;
    ta.hypr_cutoff = 1.025;  // pot.h:5
    ta.piinv = h.g.piinv; ta.dz = h.g.dz;
    ta.rhf_g = d_rhf; ta.rhfinv_g = d_rhfinv; ta.zhf_g = d_zhf;
    ta.coefh = d_coefh; ta.cser = d_cser; ta.nu = d_nu;
    launch_gen_topsym_table(ta, t->T, t->T32, ss.s);
    launch_gen_side_table(ta, t->S, t->H, ss.s);
    std::vector<double> sq(numr + 1, 0.0);
    for (int J = 1; J <= numr; ++J) sq[J] = std::sqrt(h.rhf_g[J]);
    if ((rc = upload(t->sqrt_rhf, sq, ss.s)) != BH_OK) return rc;
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ss.s));
    return BH_OK;
}

int build_modeops(const HostGeo &h, int precision, ModeOps *o)
{
    const Geo &g = h.g;
    o->g = g;
    o->dt[0] = h.dt[0];
    o->dt[1] = h.dt[1];
    const int numr = g.numr, numz = g.numz, N = g.numphi, mmax = g.mmax, nri = g.nri, nzi = g.nzi;
    (void)numz;
    const bool mixed = is_mixed(precision);
    const int nh = N / 2 + 1, ns_ = N / 2 - 1;  // sizes of the cosine and sine halves
    const double pi = std::acos(-1.0);
    DevMem &m = o->mem;
    int rc;
#define ALLOC(ptr, n) \
    if ((rc = m.alloc(&(ptr), (size_t)(n))) != BH_OK) return rc
    ALLOC(o->Dc, (size_t)nh * nh);
    ALLOC(o->Ds, (size_t)ns_ * ns_);
    ALLOC(o->Ec, (size_t)nh * nh);
    ALLOC(o->Es, (size_t)ns_ * ns_);
    ALLOC(o->Pf, (size_t)proj_coef_elems(N, mmax));
    ALLOC(o->besc, (size_t)mmax * N);
    ALLOC(o->bess, (size_t)mmax * N);
    ALLOC(o->alphar, numr + 1);
    ALLOC(o->betar, numr + 1);
    ALLOC(o->elzb, (size_t)N * (numr + 1));
    SetupStream ss;
    if (!ss.s) return fail(BH_E_DEVICE, "plan setup: stream creation failed");
    hipStream_t s = ss.s;
    // host vectors must outlive the asynchronous uploads: keep them here
    std::vector<std::vector<double>> keep;
    std::vector<std::vector<float>> keepf;
    auto up = [&](double *dst, std::vector<double> &&v) {
        keep.push_back(std::move(v));
        return upload(dst, keep.back(), s);
    };
    auto upf = [&](float *dst, std::vector<float> &&v) {
        keepf.push_back(std::move(v));
        return upload(dst, keepf.back(), s);
    };
    keep.reserve(32);
    keepf.reserve(16);

    // DFT matrices.  Forward rows in helmadi's mode order (helmadi.F90:131-156):
    // k = 0..N/2 -> sum x cos(2 pi k n/N); N/2+k -> -sum x sin(2 pi k n/N).
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
        // bessel's cos(m (L-1/2) dphi) and sin rows (potsetup.F90:131-139),
        // and the same angles as the quad-folded table of the projection
        std::vector<double> bc((size_t)mmax * N), bs((size_t)mmax * N), Pf;
        for (int M1 = 1; M1 <= mmax; ++M1)
            for (int L = 1; L <= N; ++L) {
                double a = (g.isym == 3 ? g.dphi : 0.5 * g.dphi) * ((double)M1 - 1.0) * (2.0 * (double)L - 1.0);
                bc[(size_t)(M1 - 1) * N + (L - 1)] = std::cos(a);
                bs[(size_t)(M1 - 1) * N + (L - 1)] = std::sin(a);
            }
        proj_coefficients(N, mmax, g.dphi, &Pf, g.isym);
        if ((rc = up(o->Dc, std::move(Dc))) != BH_OK || (rc = up(o->Ds, std::move(Ds))) != BH_OK ||
            (rc = up(o->Ec, std::move(Ec))) != BH_OK || (rc = up(o->Es, std::move(Es))) != BH_OK ||
            (rc = up(o->Pf, std::move(Pf))) != BH_OK || (rc = up(o->besc, std::move(bc))) != BH_OK ||
            (rc = up(o->bess, std::move(bs))) != BH_OK)
            return rc;
    }
    if (N <= DFT_MAX_N) {
        std::vector<float> f32;
        std::vector<double> f64, i64;
        dft_forward_fragments(N, mixed ? &f32 : nullptr, &f64);
        dft_inverse_fragments(N, &i64);
        ALLOC(o->dfwd64, f64.size());
        ALLOC(o->dinv64, i64.size());
        if ((rc = up(o->dfwd64, std::move(f64))) != BH_OK || (rc = up(o->dinv64, std::move(i64))) != BH_OK) return rc;
        if (mixed) {
            ALLOC(o->dfwd32, f32.size());
            if ((rc = upf(o->dfwd32, std::move(f32))) != BH_OK) return rc;
        }
    }
    if ((rc = up(o->alphar, std::vector<double>(h.alphar))) != BH_OK ||
        (rc = up(o->betar, std::vector<double>(h.betar))) != BH_OK ||
        (rc = up(o->elzb, std::vector<double>(h.elzb))) != BH_OK)
        return rc;
    const std::vector<double> &ar = h.betar, &cr = h.alphar;
    if (is_spec64(precision)) {
        // FP64 spectral solve: the sine basis of B_z, lambda_k, mu_k of both
        // schedules, and M_r(m)^T for the N/2 + 1 azimuthal mode numbers
        // (planes m and m + N/2 share one)
        const int NM = N / 2 + 1;
        ALLOC(o->S64, (size_t)nzi * nzi);
        ALLOC(o->lam64, nzi);
        ALLOC(o->mu64[0], nzi);
        ALLOC(o->mu64[1], nzi);
        ALLOC(o->MrT64[0], (size_t)NM * nri * nri);
        ALLOC(o->MrT64[1], (size_t)NM * nri * nri);
        ALLOC(o->brb_d, (size_t)N * (numr + 1));
        std::vector<double> S((size_t)nzi * nzi), lam(nzi);
        for (int a = 0; a < nzi; ++a)
            for (int k = 0; k < nzi; ++k) S[(size_t)a * nzi + k] = sine_basis(a, k, nzi);
        for (int k = 0; k < nzi; ++k) lam[k] = 2.0 * g.gamma * (1.0 - std::cos(pi * (k + 1) / (nzi + 1)));
        std::vector<double> mu[2];
        for (int w = 0; w < 2; ++w) {
            mu[w].resize(nzi);
            for (int k = 0; k < nzi; ++k) {
                double mm = 1.0;
                for (double d : h.dt[w]) mm *= (d - lam[k]) / (d + lam[k]);
                mu[w][k] = mm;
            }
        }
        if ((rc = up(o->brb_d, std::vector<double>(h.brb))) != BH_OK || (rc = up(o->S64, std::move(S))) != BH_OK ||
            (rc = up(o->lam64, std::move(lam))) != BH_OK || (rc = up(o->mu64[0], std::move(mu[0]))) != BH_OK ||
            (rc = up(o->mu64[1], std::move(mu[1]))) != BH_OK)
            return rc;
        const int chunk = 64;
        DevMem scratch;
        double *sv, *su, *sg, *ddt[2];
        const size_t nsc = (size_t)chunk * nri * nri;
        if ((rc = scratch.alloc(&sv, nsc)) != BH_OK || (rc = scratch.alloc(&su, nsc)) != BH_OK ||
            (rc = scratch.alloc(&sg, nsc)) != BH_OK || (rc = scratch.alloc(&ddt[0], h.dt[0].size())) != BH_OK ||
            (rc = scratch.alloc(&ddt[1], h.dt[1].size())) != BH_OK)
            return rc;
        for (int w = 0; w < 2; ++w) {
            if ((rc = upload(ddt[w], h.dt[w], s)) != BH_OK) return rc;
            for (int m0 = 0; m0 < NM; m0 += chunk) {
                MrArgs ma{};
                ma.brb = o->brb_d; ma.alphar = o->alphar; ma.betar = o->betar; ma.dt = ddt[w];
                ma.numr = numr; ma.nri = nri; ma.nsteps = (int)h.dt[w].size();
                ma.mode0 = m0; ma.nmodes = std::min(chunk, NM - m0);
                ma.scratch_v = sv; ma.scratch_u = su; ma.scratch_g = sg;
                ma.Mr64 = o->MrT64[w];
                launch_build_mr(ma, s);
            }
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s));
        return BH_OK;
    }
    if (!mixed) {
        // Thomas factors per pseudo-timestep (tridagr.F90:41-60, tridagz.F90:115-132)
        for (int w = 0; w < 2; ++w) {
            const std::vector<double> &dt = h.dt[w];
            int ns = (int)dt.size();
            ALLOC(o->rc_alpha[w], (size_t)ns * N * nri);
            ALLOC(o->rc_ibet[w], (size_t)ns * N * nri);
            ALLOC(o->rc_gback[w], (size_t)ns * N * nri);
            ALLOC(o->zc_alpha[w], (size_t)ns * nzi);
            ALLOC(o->zc_ibet[w], (size_t)ns * nzi);
            ALLOC(o->zc_gback[w], (size_t)ns * nzi);
            std::vector<double> ra((size_t)ns * N * nri), ri((size_t)ns * N * nri), rg((size_t)ns * N * nri);
            std::vector<double> za((size_t)ns * nzi), zi((size_t)ns * nzi), zg((size_t)ns * nzi);
            std::vector<double> bet(numr + 1), gam(numr + 2);
            for (int st = 0; st < ns; ++st) {
                for (int L = 1; L <= N; ++L) {
                    const double *b = &h.brb[(size_t)(L - 1) * (numr + 1)];
                    bet[2] = b[2] + dt[st];
                    for (int J = 3; J <= numr - 1; ++J) {
                        gam[J] = cr[J - 1] / bet[J - 1];
                        bet[J] = (b[J] + dt[st]) - ar[J] * gam[J];
                    }
                    size_t off = ((size_t)st * N + (L - 1)) * nri;
                    for (int J = 2; J <= numr - 1; ++J) {
                        double ib = 1.0 / bet[J];
                        ra[off + J - 2] = (J == 2) ? 0.0 : -ar[J] * ib;
                        ri[off + J - 2] = ib;
                        rg[off + J - 2] = (J <= numr - 2) ? -gam[J + 1] : 0.0;
                    }
                }
                std::vector<double> bz(g.numz + 1), betz(g.numz + 1), gamz(g.numz + 2);
                const double az = -g.gamma, cz = -g.gamma;
                for (int K = 2; K <= g.numz - 1; ++K) bz[K] = 2.0 * g.gamma + dt[st];
                if (g.isym != 1) bz[2] = g.gamma + dt[st];  // bzb(2) = gamma (potsetup.F90:228)
                betz[2] = bz[2];
                for (int K = 3; K <= g.numz - 1; ++K) {
                    gamz[K] = cz / betz[K - 1];
                    betz[K] = bz[K] - az * gamz[K];
                }
                for (int K = 2; K <= g.numz - 1; ++K) {
                    double ib = 1.0 / betz[K];
                    za[(size_t)st * nzi + K - 2] = (K == 2) ? 0.0 : -az * ib;
                    zi[(size_t)st * nzi + K - 2] = ib;
                    zg[(size_t)st * nzi + K - 2] = (K <= g.numz - 2) ? -gamz[K + 1] : 0.0;
                }
            }
            if ((rc = up(o->rc_alpha[w], std::move(ra))) != BH_OK || (rc = up(o->rc_ibet[w], std::move(ri))) != BH_OK ||
                (rc = up(o->rc_gback[w], std::move(rg))) != BH_OK || (rc = up(o->zc_alpha[w], std::move(za))) != BH_OK ||
                (rc = up(o->zc_ibet[w], std::move(zi))) != BH_OK || (rc = up(o->zc_gback[w], std::move(zg))) != BH_OK)
                return rc;
        }
        HIPCHK(hipStreamSynchronize(s));
        return BH_OK;
    }

    // mixed precision: FP32 forward DFT rows (helmadi's mode order), sine
    // transform, lambda_z, mu_z for both schedules, shifted LU factors, M_r
    const size_t PI = (size_t)N * nri * nzi;  // interior values, all modes
    ALLOC(o->Dc32, (size_t)nh * nh);
    ALLOC(o->Ds32, (size_t)ns_ * ns_);
    ALLOC(o->S32, (size_t)((nzi + 1) / 2) * nzi);
    ALLOC(o->mu32[0], nzi);
    ALLOC(o->mu32[1], nzi);
    // the fused step keeps M_r and the pivots per azimuthal mode number
    // (planes m and m + N/2 share them): N/2 + 1 of each
    o->spec_fused = spectral_fused_ok(nri, nzi);
    const int NM = o->spec_fused ? N / 2 + 1 : N;
    ALLOC(o->Mr[0], (size_t)NM * nri * nri);
    ALLOC(o->Mr[1], (size_t)NM * nri * nri);
    ALLOC(o->tri_ibet, (size_t)NM * nri * nzi);
    if (!o->spec_fused) ALLOC(o->tri_gam, PI);  // the fused step forms g from the pivots
    ALLOC(o->brb_d, (size_t)N * (numr + 1));
    if (o->spec_fused) {
        ALLOC(o->sfb, (size_t)(nzi / 2) * nzi);
        ALLOC(o->sfa, (size_t)(nzi / 2) * nzi);
        if (nzi >= 64) {
            ALLOC(o->sfb16, (size_t)(nzi / 2) * nzi);
            ALLOC(o->sfa16, (size_t)(nzi / 2) * nzi);
        }
    }
#undef ALLOC
    if ((rc = up(o->brb_d, std::vector<double>(h.brb))) != BH_OK) return rc;
    if (o->spec_fused) {
        std::vector<float> fb, fa;
        std::vector<float> fb16, fa16;
        spectral_fragments(nzi, &fb, &fa, &fb16, &fa16);
        if ((rc = upf(o->sfb, std::move(fb))) != BH_OK || (rc = upf(o->sfa, std::move(fa))) != BH_OK) return rc;
        if (o->sfb16 && ((rc = upf(o->sfb16, std::move(fb16))) != BH_OK || (rc = upf(o->sfa16, std::move(fa16))) != BH_OK))
            return rc;
    }
    {
        std::vector<float> Dc32((size_t)nh * nh), Ds32((size_t)ns_ * ns_), S((size_t)((nzi + 1) / 2) * nzi);
        for (int k = 0; k < nh; ++k)
            for (int n = 0; n < nh; ++n)
                Dc32[(size_t)k * nh + n] = (float)std::cos(2.0 * pi * (double)(((long)k * n) % N) / N);
        for (int k = 0; k < ns_; ++k)
            for (int n = 0; n < ns_; ++n)
                Ds32[(size_t)k * ns_ + n] = (float)-std::sin(2.0 * pi * (double)(((long)(k + 1) * (n + 1)) % N) / N);
        // The sine basis of B_z, S[a][k] = sqrt(2/(n+1)) sin(pi (a+1)(k+1)/(n+1)),
        // n = nzi, kept folded: S[n-1-a][k] = (-1)^k S[a][k], so the rows
        // a < ceil(n/2) hold it all.  Spectral storage orders the modes even
        // k first: slot m holds k = spec_mode(m) (zeros of the table exact).
        const int he = (nzi + 1) / 2;
        for (int a = 0; a < he; ++a)
            for (int m = 0; m < nzi; ++m) S[(size_t)a * nzi + m] = (float)sine_basis(a, spec_mode(m, nzi), nzi);
        if ((rc = upf(o->Dc32, std::move(Dc32))) != BH_OK || (rc = upf(o->Ds32, std::move(Ds32))) != BH_OK ||
            (rc = upf(o->S32, std::move(S))) != BH_OK)
            return rc;
        std::vector<double> lamz(nzi);
        for (int k = 0; k < nzi; ++k) lamz[k] = 2.0 * g.gamma * (1.0 - std::cos(pi * (k + 1) / (nzi + 1)));
        for (int w = 0; w < 2; ++w) {
            std::vector<float> mu(nzi);
            for (int m = 0; m < nzi; ++m) {
                const double lk = lamz[spec_mode(m, nzi)];
                double mm = 1.0;
                for (double d : h.dt[w]) mm *= (d - lk) / (d + lk);
                mu[m] = (float)mm;
            }
            if ((rc = upf(o->mu32[w], std::move(mu))) != BH_OK) return rc;
        }
        // LU factors of B_r(L) + lambda_z(k), layout [L][J'][m], k = spec_mode(m)
        std::vector<float> ib((size_t)NM * nri * nzi), gm(o->spec_fused ? 0 : PI);
        for (int L = 0; L < NM; ++L) {
            const double *b = &h.brb[(size_t)L * (numr + 1)];
            for (int m = 0; m < nzi; ++m) {
                const double lk = lamz[spec_mode(m, nzi)];
                size_t base = (size_t)L * nri * nzi + m;
                double bet = b[2] + lk;
                const bool keep_g = !gm.empty();
                ib[base] = (float)(1.0 / bet);
                if (keep_g) gm[base] = 0.f;
                for (int j = 1; j < nri; ++j) {
                    int J = j + 2;
                    double gg = h.alphar[J - 1] / bet;
                    bet = (b[J] + lk) - h.betar[J] * gg;
                    ib[base + (size_t)j * nzi] = (float)(1.0 / bet);
                    if (keep_g) gm[base + (size_t)j * nzi] = (float)gg;
                }
            }
        }
        if ((rc = upf(o->tri_ibet, std::move(ib))) != BH_OK) return rc;
        if (o->tri_gam && (rc = upf(o->tri_gam, std::move(gm))) != BH_OK) return rc;
    }
    // M_r(L) for both schedules, on the device in chunks of modes
    {
        const int chunk = 64;
        DevMem scratch;
        double *sv, *su, *sg, *ddt[2];
        const size_t nsc = (size_t)chunk * nri * nri;
        if ((rc = scratch.alloc(&sv, nsc)) != BH_OK || (rc = scratch.alloc(&su, nsc)) != BH_OK ||
            (rc = scratch.alloc(&sg, nsc)) != BH_OK || (rc = scratch.alloc(&ddt[0], h.dt[0].size())) != BH_OK ||
            (rc = scratch.alloc(&ddt[1], h.dt[1].size())) != BH_OK)
            return rc;
        for (int w = 0; w < 2; ++w) {
            if ((rc = upload(ddt[w], h.dt[w], s)) != BH_OK) return rc;
            for (int m0 = 0; m0 < NM; m0 += chunk) {
                MrArgs ma{};
                ma.brb = o->brb_d; ma.alphar = o->alphar; ma.betar = o->betar; ma.dt = ddt[w];
                ma.numr = numr; ma.nri = nri; ma.nsteps = (int)h.dt[w].size();
                ma.mode0 = m0; ma.nmodes = std::min(chunk, NM - m0);
                ma.scratch_v = sv; ma.scratch_u = su; ma.scratch_g = sg; ma.Mr = o->Mr[w];
                ma.frag = o->spec_fused ? 1 : 0;
                launch_build_mr(ma, s);
            }
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s));
    }
    HIPCHK(hipStreamSynchronize(s));
    return BH_OK;
}

}  // namespace

int get_tables(const HostGeo &h, int precision, int device, std::shared_ptr<const GeoTables> *out, bool *created)
{
    std::lock_guard<std::mutex> lk(g_cache_mu);
    const std::string key = tables_key(h.d, precision, device);
    *created = false;
    auto it = g_tables.find(key);
    if (it != g_tables.end()) {
        if (auto sp = it->second.lock()) {
            *out = sp;
            return BH_OK;
        }
    }
    auto t = std::make_shared<GeoTables>();
    int rc = build_tables(h, table32(precision), t.get());
    if (rc != BH_OK) return rc;
    g_tables[key] = t;
    *out = t;
    *created = true;
    return BH_OK;
}

int get_mode_ops(const HostGeo &h, int precision, int device, std::shared_ptr<const ModeOps> *out, bool *created)
{
    std::lock_guard<std::mutex> lk(g_cache_mu);
    const std::string key = modeops_key(h.d, precision, device);
    *created = false;
    auto it = g_modeops.find(key);
    if (it != g_modeops.end()) {
        if (auto sp = it->second.lock()) {
            *out = sp;
            return BH_OK;
        }
    }
    auto o = std::make_shared<ModeOps>();
    int rc = build_modeops(h, precision, o.get());
    if (rc != BH_OK) return rc;
    g_modeops[key] = o;
    *out = o;
    *created = true;
    return BH_OK;
}

size_t plan_private_bytes(const bh_grid_desc &d, int precision)
{
    const size_t M = (size_t)d.numr * d.numz, N = d.numphi, mm = d.mmax, nri = d.numr - 2, nzi = d.numz - 2;
    const size_t nf = 2 * nri + nzi;
    const size_t NB16 = 16 * ((nri + 15) / 16), W = (nri + 31) / 32;
    size_t b = 8 * (2 * M * N + M * (2 * mm - 1) + mm * 4 * nri + side_splits(d.numz) * mm * 2 * nzi + N * nf +
                    4 * (d.numr + 1) + 6 * N +
                    (tb_blocked(d.numr) ? 0 : mm * nzi * NB16 * 4) +
                    mm * (tb_blocked(d.numr) ? tb_chunks(d.numz) : topsym_splits(d.numz) * W) * NB16 * 4);
    if (is_mixed(precision)) {
        // RH32, r32, D1 + Wf (+ D2 with two or more passes of the fused
        // step); RT, CT, DT + WT only for the separate-kernel step
        const bool fused = spectral_fused_ok((int)nri, (int)nzi);
        b += 4 * ((M + 31) / 32 * 32 * N + (fused ? 1 : 5) * N * nri * nzi +
                  (fused && d.refine_passes >= 2 ? 3 : 2) * M * N);
    }
    else if (is_spec64(precision)) b += 8 * (M * N + 3 * N * nri * nzi);
    else b += 8 * (M * N + N * (2 * tiled_stride((int)nri, (int)nzi) + tiled_stride((int)nzi, (int)nri)));
    return b;
}

size_t plan_shared_bytes(const bh_grid_desc &d, int precision)
{
    return tables_bytes(d, precision) + modeops_bytes(d, precision);
}

size_t plan_new_shared_bytes(const bh_grid_desc &d, int precision, int device)
{
    std::lock_guard<std::mutex> lk(g_cache_mu);
    size_t b = 0;
    auto t = g_tables.find(tables_key(d, precision, device));
    if (t == g_tables.end() || t->second.expired()) b += tables_bytes(d, precision);
    auto o = g_modeops.find(modeops_key(d, precision, device));
    if (o == g_modeops.end() || o->second.expired()) b += modeops_bytes(d, precision);
    return b;
}

}  // namespace bh
