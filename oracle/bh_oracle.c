/*
 * bh_oracle.c -- CPU restatement of bipoly_hydro's self-gravity solve
 * (potential_solver), in plain C99 + OpenMP, FP64 throughout.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP product
 * path is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product library never links
 * or calls it.
 *
 * It restates, for one global (r, z, phi) block (the single-rank collapse of
 * the reference's 2-D MPI decomposition, in which every swap_* transpose and
 * comm* halo exchange is an identity), the arithmetic of:
 *
 *   potential_solver.F90:66-106   source copy + zeroing, nsteps rule, driver
 *   bessel.F90:170-488            boundary potential (projection, contraction,
 *                                 face rebuild, axis ghost, corner zeros)
 *   helmadi.F90:124-405           dt schedule, phi real DFT + mode order,
 *                                 Peaceman-Rachford ADI, inverse DFT, axis BC
 *   tridagr.F90:41-69             Thomas along r
 *   tridagz.F90:115-141           Thomas along z
 *   potsetup.F90:113-252          trig tables, ADI coefficients
 *   tm.F90:101-366, bm.F90:101-172, sm.F90:115-385
 *                                 Green's-function tables (isym = 1, 2, 3)
 *   gammln.F90:14-26              ln Gamma, 6-term Lanczos form
 *   ellf.F90 / elle.F90           complete elliptic K, E via Carlson's RF,
 *                                 RD (rf.F90, rd.F90: semantics only; the
 *                                 series and stopping rule of Carlson 1995)
 *   setup_frac.F90:374-463        grid coordinates
 *
 * All three symmetry settings of the run-time flag isym (setup_frac.F90:313)
 * are restated: 1 = no symmetry (full z range), 2 = equatorial symmetry (the
 * grid starts at the equator, K = 1 is the mirror of K = 2), 3 = equatorial
 * plus pi symmetry (even azimuthal modes only).
 *
 * Array layout follows the reference: Fortran column-major
 * (numr, numz, numphi), r fastest, ghost layers included.  In C:
 *   idx(J, K, L) = (J-1) + numr*((K-1) + numz*(L-1))   with 1-based J,K,L.
 *
 * The phi transform is computed by a radix-2 complex FFT written here from
 * the textbook Cooley-Tukey algorithm; only the mode order and sign
 * convention of realft.F90 are reproduced (helmadi.F90:131-183, 359-392),
 * not its twiddle recurrence.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define BHO_MMAX_DEFAULT 30
#define BHO_MMAX_LIMIT 512 /* mode slots beyond the reference's 30: see bho_qhigh */
#define BHO_HYPR_CUTOFF 1.025
#define BHO_HYPR_UPR_BND 110

typedef struct bho_plan {
    int numr, numz, numphi, isym, mmax, nser; /* nser = hypr_upr_bnd + 1 */
    double hypr_cutoff;
    double pi, piinv, four_pi;
    double dr, dz, dphi, drinv, dzinv, dphiinv, gamma;
    /* global grid, 1-based storage (index 0 unused) */
    double *r_g, *rhf_g, *rinv_g, *rhfinv_g, *zhf_g;
    /* ADI coefficients (potsetup.F90:151-252) */
    double *ar, *cr, *alphar, *betar; /* [numr+1] */
    double *brb;                      /* [numphi][numr+1] */
    double *bzb;                      /* [numz+1] */
    double *elambdazb;                /* [numphi][numr+1] */
    /* bessel trig tables (potsetup.F90:131-139): [mmax][numphi] */
    double *bes_cos, *bes_sin;
    /* Green's tables, interior source points only:
       tab[((m*nout + jout) * nzi + kz) * nri + jr]
       with jr = rindex-2 (0..numr-3), kz = zindex-2 (0..numz-3),
       jout = global target index (0..numr-1 for tmr/bmr, 0..numz-1 smz) */
    double *tmr, *bmr, *smz;
    /* hypergeometric-series coefficients (tm.F90:118-155) */
    double *coefh;  /* [mmax] */
    double *dcoefh; /* [mmax][nser] */
    double *nu;     /* [mmax+1] 1-based */
} bho_plan;

#define IDX(p, J, K, L) ((size_t)((J) - 1) + (size_t)(p)->numr * ((size_t)((K) - 1) + (size_t)(p)->numz * (size_t)((L) - 1)))

/* ------------------------------------------------------------------ */
/* ln Gamma: Lanczos approximation, gamma = 5, six terms, in the form  */
/* published in Numerical Recipes (2nd ed.) sec. 6.1, whose error       */
/* |eps| < 2e-10 is reproduced on purpose: the Green's-function series */
/* coefficients inherit it (gammln.F90:14-26, tm.F90:125-152).          */
/* ------------------------------------------------------------------ */
double bho_gammln(double xx)
{
    static const double c[6] = {76.18009172947146, -86.50532032941677,
                                24.01409824083091, -1.231739572450155,
                                0.1208650973866179e-2, -0.5395239384953e-5};
    const double sqrt2pi = 2.5066282746310005;
    double x = xx, y = xx;
    double tmp = x + 5.5;
    tmp = (x + 0.5) * log(tmp) - tmp;
    double ser = 1.000000000190015;
    for (int j = 0; j < 6; ++j) {
        y = y + 1.0;
        ser = ser + c[j] / y;
    }
    return tmp + log(sqrt2pi * ser / x);
}

/* ------------------------------------------------------------------ */
/* Carlson's symmetric elliptic integrals (B. C. Carlson, Numer.        */
/* Algorithms 10 (1995) 13-26; DLMF 19.36.1-2).  The duplication        */
/* v <- (v + lam) / 4 (lam = sqrt(xy) + sqrt(xz) + sqrt(yz)) shrinks    */
/* the spread of the arguments about their mean A by 4 per step.  It    */
/* stops by Carlson's a-priori rule 4^-m Q < |A_m|, Q = (3 r)^(-1/6)    */
/* (RF) or (r / 4)^(-1/6) (RD) times the initial spread, r = 1e-17; the */
/* deviations X, Y, Z, formed from the initial arguments, enter the     */
/* series in the elementary symmetric functions E2..E5.  The reference  */
/* (rf.F90, rd.F90) stops on the running deviations instead; both are   */
/* accurate to a few units of the last place.                           */
/* ------------------------------------------------------------------ */
static double spread3(double a, double x, double y, double z)
{
    return fmax(fabs(a - x), fmax(fabs(a - y), fabs(a - z)));
}

double bho_rf(double x, double y, double z)
{
    const double x0 = x, y0 = y, a0 = (x + y + z) / 3.0;
    const double q = pow(3.0e-17, -1.0 / 6.0) * spread3(a0, x, y, z);
    double a = a0, f = 1.0; /* f = 4^-m */
    while (f * q >= fabs(a)) {
        const double sx = sqrt(x), sy = sqrt(y), sz = sqrt(z);
        const double lam = sx * sy + sx * sz + sy * sz;
        x = 0.25 * (x + lam);
        y = 0.25 * (y + lam);
        z = 0.25 * (z + lam);
        a = 0.25 * (a + lam);
        f *= 0.25;
    }
    const double X = f * (a0 - x0) / a, Y = f * (a0 - y0) / a, Z = -(X + Y);
    const double e2 = X * Y - Z * Z, e3 = X * Y * Z;
    /* DLMF 19.36.1 through seventh order */
    const double t = 1.0 - e2 / 10.0 + e3 / 14.0 + e2 * e2 / 24.0 - 3.0 * e2 * e3 / 44.0 -
                     5.0 * e2 * e2 * e2 / 208.0 + 3.0 * e3 * e3 / 104.0 + e2 * e2 * e3 / 16.0;
    return t / sqrt(a);
}

double bho_rd(double x, double y, double z)
{
    const double x0 = x, y0 = y, a0 = (x + y + 3.0 * z) / 5.0;
    const double q = pow(0.25e-17, -1.0 / 6.0) * spread3(a0, x, y, z);
    double a = a0, f = 1.0, tail = 0.0; /* tail = sum_m 4^-m / (sqrt(z_m) (z_m + lam_m)) */
    while (f * q >= fabs(a)) {
        const double sx = sqrt(x), sy = sqrt(y), sz = sqrt(z);
        const double lam = sx * sy + sx * sz + sy * sz;
        tail += f / (sz * (z + lam));
        x = 0.25 * (x + lam);
        y = 0.25 * (y + lam);
        z = 0.25 * (z + lam);
        a = 0.25 * (a + lam);
        f *= 0.25;
    }
    const double X = f * (a0 - x0) / a, Y = f * (a0 - y0) / a, Z = -(X + Y) / 3.0;
    const double xy = X * Y, z2 = Z * Z;
    const double e2 = xy - 6.0 * z2, e3 = (3.0 * xy - 8.0 * z2) * Z, e4 = 3.0 * (xy - z2) * z2, e5 = xy * z2 * Z;
    /* DLMF 19.36.2 */
    const double t = 1.0 - 3.0 * e2 / 14.0 + e3 / 6.0 + 9.0 * e2 * e2 / 88.0 - 3.0 * e4 / 22.0 -
                     9.0 * e2 * e3 / 52.0 + 3.0 * e5 / 26.0;
    return f * t / (a * sqrt(a)) + 3.0 * tail;
}

/* complete elliptic integrals at modulus k (ellf.F90:16, elle.F90:16-18) */
double bho_ellk(double k) { return bho_rf(0.0, (1.0 - k) * (1.0 + k), 1.0); }
double bho_elle(double k)
{
    double q = (1.0 - k) * (1.0 + k);
    return bho_rf(0.0, q, 1.0) - k * k * bho_rd(0.0, q, 1.0) * (1.0 / 3.0);
}

/* ------------------------------------------------------------------ */
/* Q_{mu-1/2}(x) for the table's mode slots m = 0..mmax-1, exactly as   */
/* tm.F90:168-194 (isym 1, 2: mu = m) and tm.F90:292-352 (isym 3: even  */
/* mu = 2m) evaluate it: upward recurrence from K, E below the cutoff   */
/* (isym 3: the two-step recurrence from Q_{-1/2} and Q_{3/2}),         */
/* truncated hypergeometric series with exp/log terms at and above it.  */
/* ------------------------------------------------------------------ */
/* ------------------------------------------------------------------ */
/* Mode slots m >= 30 (plans with mmax above the reference's pot.h:3     */
/* value; parity there is with the mathematics, not the reference).      */
/* The reference's upward recurrence multiplies its rounding error by    */
/* about e^{2 mu eta} (chi = cosh eta) and its 111-term series does not  */
/* converge for high orders near chi = 1, so these slots take the        */
/* toroidal functions from the ratios R_mu = Q_{mu-1/2} / Q_{mu-3/2} of  */
/* the minimal solution of the Legendre recurrence, run downwards (the   */
/* stable direction; Gautschi, SIAM Rev. 9 (1967) 24-82):                */
/*   R_mu = (mu - 1/2) / (2 mu chi - (mu + 1/2) R_{mu+1}),               */
/* started at mu = M + 1 with the limit ratio e^{-eta}, M far enough     */
/* above the highest order that the start's error, damped by e^{-2 eta}  */
/* per step, is gone (M - mu_top >= 20 / eta + 16).  Q_{-1/2} = k K(k),   */
/* k^2 = 2 / (chi + 1), normalises the products Q_{mu-1/2} =             */
/* Q_{-1/2} R_1 ... R_mu.  The slots are done in chunks of 16, each      */
/* rerunning the downward pass from M, so that only 16 ratios are held.  */
/* isym = 3: slot m holds mu = 2m, the product of two ratios per slot.   */
/* ------------------------------------------------------------------ */
#define BHO_MREF 30
#define BHO_QCHUNK 16
static void bho_qhigh(const bho_plan *p, double xm, double *q)
{
    const int mmax = p->mmax, fs = p->isym == 3 ? 2 : 1, s0 = BHO_MREF;
    if (mmax <= s0) return;
    const double e1 = xm - 1.0, sh = sqrt(e1 * (xm + 1.0)); /* sinh eta */
    const double eta = log1p(e1 + sh), rinf = 1.0 / (xm + sh);
    const int mu_top = fs * (mmax - 1);
    const double extra = ceil(20.0 / eta);
    const int M = mu_top + 16 + (int)(extra < 1.0e6 ? extra : 1.0e6);
    /* Q_{-1/2} = k K(k) with k'^2 = (chi - 1) / (chi + 1) formed directly
       (not as 1 - k^2), which keeps its digits as chi -> 1 */
    double qprev = sqrt(2.0 / (1.0 + xm)) * bho_rf(0.0, e1 / (xm + 1.0), 1.0); /* then the slot below the chunk */
    for (int slo = s0; slo < mmax; slo += BHO_QCHUNK) {
        const int shi = slo + BHO_QCHUNK < mmax ? slo + BHO_QCHUNK : mmax;
        double sr[BHO_QCHUNK]; /* slot ratios Q(slot s) / Q(slot s - 1) */
        double r = rinf;
        int mu = M;
        for (; mu > fs * (shi - 1); --mu) r = ((double)mu - 0.5) / (2.0 * (double)mu * xm - ((double)mu + 0.5) * r);
        for (int sl = shi - 1; sl >= slo; --sl) {
            double prod = 1.0;
            for (int k = 0; k < fs; ++k, --mu) {
                r = ((double)mu - 0.5) / (2.0 * (double)mu * xm - ((double)mu + 0.5) * r);
                prod *= r;
            }
            sr[sl - slo] = prod;
        }
        if (slo == s0) { /* Q of slot s0 - 1: Q_{-1/2} R_1 ... R_{fs (s0 - 1)} */
            double prod = 1.0;
            for (; mu >= 1; --mu) {
                r = ((double)mu - 0.5) / (2.0 * (double)mu * xm - ((double)mu + 0.5) * r);
                prod *= r;
            }
            qprev *= prod;
        }
        for (int sl = slo; sl < shi; ++sl) {
            qprev *= sr[sl - slo];
            q[sl] = qprev;
        }
    }
}

static void bho_qhalf(const bho_plan *p, double xm, double *q)
{
    /* slots below the reference's mmax as the reference evaluates them */
    const int mmax = p->mmax < BHO_MREF ? p->mmax : BHO_MREF;
    bho_qhigh(p, xm, q);
    if (xm < p->hypr_cutoff) {
        double mum = sqrt(2.0 / (1.0 + xm));
        double lam = sqrt(2.0 * (1.0 + xm));
        double kk = bho_ellk(mum), ee = bho_elle(mum);
        q[0] = kk * mum;
        if (p->isym == 3) {
            const double four_thirds = 4.0 / 3.0, one_third = 1.0 / 3.0;
            if (mmax > 1) q[1] = (four_thirds * xm * xm - one_third) * mum * kk - four_thirds * xm * lam * ee;
            for (int m = 3; m <= mmax; ++m) {
                const double nu = p->nu[m], den = (2.0 * nu - 1.0) * (nu + 2.0) * (nu + 1.0);
                q[m - 1] = q[m - 2] * ((2.0 * nu + 3.0) * (2.0 * nu + 1.0) * xm * xm / ((nu + 2.0) * (nu + 1.0)) -
                                       (2.0 * nu + 3.0) * nu * nu / den - (nu + 1.0) / (nu + 2.0)) -
                           q[m - 3] * (2.0 * nu + 3.0) * (nu - 1.0) * nu / den;
            }
            return;
        }
        if (mmax > 1) q[1] = xm * mum * kk - lam * ee;
        for (int m = 3; m <= mmax; ++m) {
            double nu = p->nu[m];
            q[m - 1] = (2.0 * nu + 1.0) / (nu + 1.0) * xm * q[m - 2] - nu / (nu + 1.0) * q[m - 3];
        }
    } else {
        double l2x = log(2.0 * xm), lx = log(xm);
        for (int m = 1; m <= mmax; ++m) {
            double index = p->isym == 3 ? 2.0 * (double)(m - 1) : (double)(m - 1);
            double coefh_m = exp(p->coefh[m - 1] - (index + 0.5) * l2x);
            double sum = 0.0;
            const double *dc = p->dcoefh + (size_t)(m - 1) * p->nser;
            for (int ir = 0; ir < p->nser; ++ir) sum = sum + exp(dc[ir] - 2.0 * (double)ir * lx);
            q[m - 1] = coefh_m * sum;
        }
    }
}

static void *xcalloc(size_t n, size_t s)
{
    void *v = calloc(n ? n : 1, s);
    return v;
}

void bho_plan_destroy(bho_plan *p)
{
    if (!p) return;
    free(p->r_g); free(p->rhf_g); free(p->rinv_g); free(p->rhfinv_g); free(p->zhf_g);
    free(p->ar); free(p->cr); free(p->alphar); free(p->betar);
    free(p->brb); free(p->bzb); free(p->elambdazb);
    free(p->bes_cos); free(p->bes_sin);
    free(p->tmr); free(p->bmr); free(p->smz);
    free(p->coefh); free(p->dcoefh); free(p->nu);
    free(p);
}

/* mode number of global Fourier index L (potsetup.F90:155-173).  For
   isym = 3 the reference squares the index (potsetup.F90:159-163); that is
   restated as written. */
static int bho_mode_of(int L, int numphi, int isym)
{
    int lstop = numphi / 2 + 1;
    int k = (L <= lstop) ? L - 1 : L - lstop;
    return isym == 3 ? k * k : k;
}

static void bho_build_tables(bho_plan *p);

/* Create a plan: grid, ADI coefficients, trig and Green's tables.
   with_tables = 1 computes the Green's tables; 0 skips them (bessel must
   not be called); 2 allocates them filled with a nonzero constant, every
   page written, for timing the solve's arithmetic and table stream only
   (table generation is one-time setup, potsetup; the results are then
   meaningless).
   Returns NULL on invalid arguments. */
bho_plan *bho_plan_create(int numr, int numz, int numphi, int isym, double dr, int mmax,
                          int with_tables)
{
    if (isym < 1 || isym > 3) return NULL;
    if (numr < 4 || numz < 4 || numphi < 4) return NULL;
    if (numphi & (numphi - 1)) return NULL;
    if (mmax < 1 || mmax > BHO_MMAX_LIMIT) return NULL;
    bho_plan *p = (bho_plan *)xcalloc(1, sizeof(bho_plan));
    p->numr = numr; p->numz = numz; p->numphi = numphi; p->isym = isym; p->mmax = mmax;
    p->nser = BHO_HYPR_UPR_BND + 1;
    p->hypr_cutoff = BHO_HYPR_CUTOFF;
    /* setup_frac.F90:345, 374-379; potsetup.F90:113-114 */
    p->pi = acos(-1.0);
    p->piinv = 1.0 / p->pi;
    p->four_pi = 4.0 * p->pi;
    p->dr = dr;
    p->dz = dr;
    p->dphi = 2.0 * p->pi * (1.0 / (double)numphi);
    p->drinv = 1.0 / p->dr;
    p->dzinv = 1.0 / p->dz;
    p->dphiinv = 1.0 / p->dphi;
    /* global grid: setup_frac.F90:428-463 */
    p->r_g = (double *)xcalloc(numr + 1, sizeof(double));
    p->rhf_g = (double *)xcalloc(numr + 1, sizeof(double));
    p->rinv_g = (double *)xcalloc(numr + 1, sizeof(double));
    p->rhfinv_g = (double *)xcalloc(numr + 1, sizeof(double));
    p->zhf_g = (double *)xcalloc(numz + 1, sizeof(double));
    for (int J = 1; J <= numr; ++J) {
        double x = (double)J;
        p->r_g[J] = (x - 2.0) * p->dr;
        p->rhf_g[J] = p->r_g[J] + 0.5 * p->dr;
        p->rinv_g[J] = (p->r_g[J] != 0.0) ? 1.0 / p->r_g[J] : 0.0;
        p->rhfinv_g[J] = 1.0 / p->rhf_g[J];
    }
    if (isym != 1) { /* setup_frac.F90:452-456: K = 1 is the mirror cell below the equator */
        for (int K = 1; K <= numz; ++K) p->zhf_g[K] = ((double)K - 1.5) * p->dz;
    } else {
        double offset = -((double)(numz / 2) * p->dz);
        for (int K = 1; K <= numz; ++K) p->zhf_g[K] = offset + ((double)K - 0.5) * p->dz;
    }
    /* ADI coefficients: potsetup.F90:151-252 (single block: row_num = 0,
       numr_dd_z = numr, global radial index = J) */
    double gamma = p->dzinv * p->dzinv;
    double drinv2 = p->drinv * p->drinv;
    double dphiinv2 = p->dphiinv * p->dphiinv;
    p->gamma = gamma;
    p->ar = (double *)xcalloc(numr + 1, sizeof(double));
    p->cr = (double *)xcalloc(numr + 1, sizeof(double));
    p->alphar = (double *)xcalloc(numr + 1, sizeof(double));
    p->betar = (double *)xcalloc(numr + 1, sizeof(double));
    for (int J = 2; J <= numr - 1; ++J) {
        p->alphar[J] = -p->r_g[J + 1] * p->rhfinv_g[J] * drinv2;
        p->betar[J] = -p->r_g[J] * p->rhfinv_g[J] * drinv2;
    }
    for (int J = 1; J <= numr; ++J) { p->ar[J] = p->betar[J]; p->cr[J] = p->alphar[J]; }
    p->brb = (double *)xcalloc((size_t)numphi * (numr + 1), sizeof(double));
    p->elambdazb = (double *)xcalloc((size_t)numphi * (numr + 1), sizeof(double));
    for (int L = 1; L <= numphi; ++L) {
        int mode = bho_mode_of(L, numphi, isym);
        double m1mode = (mode % 2 == 0) ? 1.0 : -1.0;
        double elm = cos((double)mode * p->dphi);
        double *brb = p->brb + (size_t)(L - 1) * (numr + 1);
        double *elz = p->elambdazb + (size_t)(L - 1) * (numr + 1);
        for (int J = 3; J <= numr - 1; ++J)
            brb[J] = 2.0 * drinv2 - 2.0 * (elm - 1.0) * dphiinv2 * p->rhfinv_g[J] * p->rhfinv_g[J];
        /* the axis row: potsetup.F90:211-223 and 239-251 (isym = 3 mirrors
           the axis cell instead of rotating it by pi) */
        if (isym == 3)
            brb[2] = -p->alphar[2] - 2.0 * p->betar[2] -
                     2.0 * (elm - 1.0) * dphiinv2 * p->rhfinv_g[2] * p->rhfinv_g[2];
        else
            brb[2] = -p->alphar[2] + (m1mode - 1.0) * p->betar[2] -
                     2.0 * (elm - 1.0) * dphiinv2 * p->rhfinv_g[2] * p->rhfinv_g[2];
        for (int J = 2; J <= numr - 1; ++J)
            elz[J] = -2.0 * drinv2 + 2.0 * (elm - 1.0) * dphiinv2 * p->rhfinv_g[J] * p->rhfinv_g[J];
        if (isym == 3)
            elz[2] = p->alphar[2] + 2.0 * p->betar[2] +
                     2.0 * (elm - 1.0) * dphiinv2 * p->rhfinv_g[2] * p->rhfinv_g[2];
        else
            elz[2] = p->alphar[2] - (m1mode - 1.0) * p->betar[2] +
                     2.0 * (elm - 1.0) * dphiinv2 * p->rhfinv_g[2] * p->rhfinv_g[2];
    }
    p->bzb = (double *)xcalloc(numz + 1, sizeof(double));
    for (int K = 2; K <= numz - 1; ++K) p->bzb[K] = 2.0 * gamma;
    if (isym != 1) p->bzb[2] = gamma; /* potsetup.F90:228 */
    /* bessel trig tables: potsetup.F90:120-139 (isym = 3: even modes) */
    p->bes_cos = (double *)xcalloc((size_t)mmax * numphi, sizeof(double));
    p->bes_sin = (double *)xcalloc((size_t)mmax * numphi, sizeof(double));
    for (int M = 1; M <= mmax; ++M)
        for (int L = 1; L <= numphi; ++L) {
            double a = (isym == 3 ? p->dphi : 0.5 * p->dphi) * ((double)M - 1.0) * (2.0 * (double)L - 1.0);
            p->bes_cos[(size_t)(M - 1) * numphi + (L - 1)] = cos(a);
            p->bes_sin[(size_t)(M - 1) * numphi + (L - 1)] = sin(a);
        }
    /* series coefficients: tm.F90:103-155 */
    p->nu = (double *)xcalloc(mmax + 1, sizeof(double));
    for (int m = 3; m <= mmax; ++m) p->nu[m] = isym == 3 ? 2.0 * (double)m - 4.5 : (double)m - 2.5;
    p->coefh = (double *)xcalloc(mmax, sizeof(double));
    p->dcoefh = (double *)xcalloc((size_t)mmax * p->nser, sizeof(double));
    double gl_half = bho_gammln(0.5);
    for (int m = 1; m <= mmax && m <= BHO_MREF; ++m) {
        double mm = isym == 3 ? 2.0 * ((double)m - 1.0) : (double)m - 1.0;
        double aa = 0.5 * (mm + 1.5), bb = 0.5 * (mm + 0.5), cc = mm + 1.0;
        p->coefh[m - 1] = gl_half + bho_gammln(mm + 0.5) - bho_gammln(aa) - bho_gammln(bb);
        for (int ir = 0; ir < p->nser; ++ir) {
            double ri = (double)ir;
            p->dcoefh[(size_t)(m - 1) * p->nser + ir] =
                bho_gammln(aa + ri) + bho_gammln(bb + ri) - bho_gammln(cc + ri) - bho_gammln(ri + 1.0);
        }
    }
    if (with_tables == 1) bho_build_tables(p);
    if (with_tables == 2) {
        size_t nri = numr - 2, nzi = numz - 2;
        const size_t nt = (size_t)mmax * numr * nzi * nri, ns = (size_t)mmax * numz * nzi * nri;
        p->tmr = (double *)malloc(nt * sizeof(double));
        p->bmr = isym == 1 ? (double *)malloc(nt * sizeof(double)) : NULL;
        p->smz = (double *)malloc(ns * sizeof(double));
        if (p->tmr && (p->bmr || isym != 1) && p->smz) {
#pragma omp parallel for schedule(static)
            for (size_t i = 0; i < nt; ++i) {
                p->tmr[i] = 1e-3;
                if (p->bmr) p->bmr[i] = 1e-3;
            }
#pragma omp parallel for schedule(static)
            for (size_t i = 0; i < ns; ++i) p->smz[i] = 1e-3;
        }
    }
    return p;
}

/* Green's tables: tm.F90:157-363 (top), bm.F90:131-172 (bottom, isym = 1
   only), sm.F90:172-382 (side).  Interior source points only (rlwb..rupb,
   zlwb..zupb of every rank = global 2..numr-1, 2..numz-1).  With isym = 2,
   3 every entry adds the image of the source below the equator:
   coef (Q(chi+) + Q(chi-)), chi+ formed with z_target + z_source. */
static void bho_build_tables(bho_plan *p)
{
    const int numr = p->numr, numz = p->numz, mmax = p->mmax;
    const int nri = numr - 2, nzi = numz - 2;
    const int image = p->isym != 1;
    size_t ntop = (size_t)mmax * numr * nzi * nri;
    size_t nsid = (size_t)mmax * numz * nzi * nri;
    p->tmr = (double *)xcalloc(ntop, sizeof(double));
    p->bmr = image ? NULL : (double *)xcalloc(ntop, sizeof(double));
    p->smz = (double *)xcalloc(nsid, sizeof(double));
    const double zT = p->zhf_g[numz], zBot = p->zhf_g[1];
    const double rB = p->rhf_g[numr], rBinv = 1.0 / rB;
#pragma omp parallel for schedule(dynamic, 1)
    for (int jl = 2; jl <= numr - 1; ++jl) {
        double q[256], qp[256];
        const double rhf = p->rhf_g[jl], rhfinv = p->rhfinv_g[jl];
        for (int kl = 2; kl <= numz - 1; ++kl) {
            const double zhf = p->zhf_g[kl];
            for (int jg = 2; jg <= numr - 1; ++jg) {
                double coef = sqrt(rhf * p->rhfinv_g[jg]) * p->piinv;
                /* top (the sum in each branch's order: tm.F90:164-167, 207-214) */
                double xm = image ? 0.5 * rhfinv * p->rhfinv_g[jg] *
                                        ((zT - zhf) * (zT - zhf) + rhf * rhf + p->rhf_g[jg] * p->rhf_g[jg])
                                  : 0.5 * rhfinv * p->rhfinv_g[jg] *
                                        ((zT - zhf) * (zT - zhf) + p->rhf_g[jg] * p->rhf_g[jg] + rhf * rhf);
                bho_qhalf(p, xm, q);
                if (image) {
                    double xp = 0.5 * rhfinv * p->rhfinv_g[jg] *
                                ((zT + zhf) * (zT + zhf) + rhf * rhf + p->rhf_g[jg] * p->rhf_g[jg]);
                    bho_qhalf(p, xp, qp);
                    for (int m = 0; m < mmax; ++m) q[m] = qp[m] + q[m];
                }
                for (int m = 0; m < mmax; ++m)
                    p->tmr[(((size_t)m * numr + (jg - 1)) * nzi + (kl - 2)) * nri + (jl - 2)] = coef * q[m];
                if (image) continue;
                /* bottom */
                xm = 0.5 * rhfinv * p->rhfinv_g[jg] *
                     ((zBot - zhf) * (zBot - zhf) + p->rhf_g[jg] * p->rhf_g[jg] + rhf * rhf);
                bho_qhalf(p, xm, q);
                for (int m = 0; m < mmax; ++m)
                    p->bmr[(((size_t)m * numr + (jg - 1)) * nzi + (kl - 2)) * nri + (jl - 2)] = coef * q[m];
            }
            double coefs = sqrt(rhf * rBinv) * p->piinv;
            for (int kg = 2; kg <= numz - 1; ++kg) {
                double dzz = p->zhf_g[kg] - zhf;
                double xm = image ? 0.5 * rhfinv * rBinv * (dzz * dzz + rhf * rhf + rB * rB)
                                  : 0.5 * rhfinv * rBinv * (dzz * dzz + rB * rB + rhf * rhf);
                bho_qhalf(p, xm, q);
                if (image) {
                    double szz = p->zhf_g[kg] + zhf;
                    double xp = 0.5 * rhfinv * rBinv * (szz * szz + rhf * rhf + rB * rB);
                    bho_qhalf(p, xp, qp);
                    for (int m = 0; m < mmax; ++m) q[m] = qp[m] + q[m];
                }
                for (int m = 0; m < mmax; ++m)
                    p->smz[(((size_t)m * numz + (kg - 1)) * nzi + (kl - 2)) * nri + (jl - 2)] = coefs * q[m];
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* Real DFT along phi in helmadi's mode order.                          */
/* forward: out[0] = a0, out[k] = a_k (k=1..N/2), out[N/2+k] = -b_k,    */
/*          with a_k = sum x_n cos(2 pi k n/N), b_k = sum x_n sin(...)  */
/*          (realft.F90 isign=+1 + reorder, helmadi.F90:131-156)        */
/* inverse: exact inverse of the forward map (realft isign=-1, x 2/N,   */
/*          helmadi.F90:359-392)                                        */
/* ------------------------------------------------------------------ */
static void fft_c(double *re, double *im, int n, int sign)
{
    /* iterative radix-2 decimation-in-time, textbook form */
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1;
        for (int k = 0; k < half; ++k) {
            double ang = sign * 2.0 * M_PI * (double)k / (double)len;
            double wr = cos(ang), wi = sin(ang);
            for (int i = k; i < n; i += len) {
                int j = i + half;
                double tr = wr * re[j] - wi * im[j];
                double ti = wr * im[j] + wi * re[j];
                re[j] = re[i] - tr; im[j] = im[i] - ti;
                re[i] += tr; im[i] += ti;
            }
        }
    }
}

void bho_dft_forward(int n, const double *x, double *out)
{
    double *re = (double *)malloc(sizeof(double) * n), *im = (double *)malloc(sizeof(double) * n);
    for (int i = 0; i < n; ++i) { re[i] = x[i]; im[i] = 0.0; }
    fft_c(re, im, n, -1); /* X_k = sum x e^{-2 pi i k n/N} = a_k - i b_k */
    out[0] = re[0];
    for (int k = 1; k <= n / 2; ++k) out[k] = re[k];
    for (int k = 1; k < n / 2; ++k) out[n / 2 + k] = im[k]; /* = -b_k */
    free(re); free(im);
}

void bho_dft_inverse(int n, const double *in, double *x)
{
    double *re = (double *)malloc(sizeof(double) * n), *im = (double *)malloc(sizeof(double) * n);
    re[0] = in[0]; im[0] = 0.0;
    re[n / 2] = in[n / 2]; im[n / 2] = 0.0;
    for (int k = 1; k < n / 2; ++k) {
        re[k] = in[k]; im[k] = in[n / 2 + k];
        re[n - k] = in[k]; im[n - k] = -in[n / 2 + k];
    }
    fft_c(re, im, n, +1);
    for (int i = 0; i < n; ++i) x[i] = re[i] / (double)n;
    free(re); free(im);
}

/* transform every (J,K) column of a (numr,numz,numphi) array */
static void dft_columns(const bho_plan *p, const double *in, double *out, int inverse)
{
    const int numr = p->numr, numz = p->numz, n = p->numphi;
    const size_t plane = (size_t)numr * numz;
#pragma omp parallel
    {
        double *col = (double *)malloc(sizeof(double) * n), *res = (double *)malloc(sizeof(double) * n);
#pragma omp for schedule(static)
        for (long jk = 0; jk < (long)plane; ++jk) {
            for (int l = 0; l < n; ++l) col[l] = in[(size_t)l * plane + jk];
            if (inverse) bho_dft_inverse(n, col, res);
            else bho_dft_forward(n, col, res);
            for (int l = 0; l < n; ++l) out[(size_t)l * plane + jk] = res[l];
        }
        free(col); free(res);
    }
}

/* ------------------------------------------------------------------ */
/* bessel: boundary potential written into potp's faces                 */
/* (bessel.F90:170-488, single block: every iam_on_* flag true).        */
/* ------------------------------------------------------------------ */
void bho_bessel(const bho_plan *p, const double *rho, double *potp)
{
    const int numr = p->numr, numz = p->numz, n = p->numphi, mmax = p->mmax;
    const int nri = numr - 2, nzi = numz - 2;
    const size_t npts = (size_t)nri * nzi;
    const int bottom = p->isym == 1;
    /* bessel.F90:170-174 */
    double factor = p->isym == 3 ? -2.0 * p->dr * p->dz * p->dphi : -p->dr * p->dz * p->dphi;
    double *tmpc = (double *)xcalloc((size_t)mmax * npts, sizeof(double));
    double *tmps = (double *)xcalloc((size_t)mmax * npts, sizeof(double));
    /* projection onto the azimuthal modes: bessel.F90:179-185, 228-237 */
#pragma omp parallel for schedule(static)
    for (long pt = 0; pt < (long)npts; ++pt) {
        int jr = (int)(pt % nri), kz = (int)(pt / nri);
        int J = jr + 2, K = kz + 2;
        double s0 = 0.0;
        for (int L = 1; L <= n; ++L) s0 = s0 + rho[IDX(p, J, K, L)];
        tmpc[pt] = s0;
        for (int M = 2; M <= mmax; ++M) {
            double sc = 0.0, ss = 0.0;
            const double *bc = p->bes_cos + (size_t)(M - 1) * n, *bs = p->bes_sin + (size_t)(M - 1) * n;
            for (int L = 1; L <= n; ++L) {
                double r = rho[IDX(p, J, K, L)];
                sc = sc + r * bc[L - 1];
                ss = ss + r * bs[L - 1];
            }
            tmpc[(size_t)(M - 1) * npts + pt] = sc;
            tmps[(size_t)(M - 1) * npts + pt] = ss;
        }
    }
    /* contraction: bessel.F90:186-262; sums indexed [m][target] */
    double *stc = (double *)xcalloc((size_t)mmax * numr, sizeof(double));
    double *sts = (double *)xcalloc((size_t)mmax * numr, sizeof(double));
    double *sbc = (double *)xcalloc((size_t)mmax * numr, sizeof(double));
    double *sbs = (double *)xcalloc((size_t)mmax * numr, sizeof(double));
    double *ssc = (double *)xcalloc((size_t)mmax * numz, sizeof(double));
    double *sss = (double *)xcalloc((size_t)mmax * numz, sizeof(double));
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
    for (int m = 0; m < mmax; ++m)
        for (int jg = 2; jg <= numr - 1; ++jg) {
            const double *tt = p->tmr + ((size_t)m * numr + (jg - 1)) * npts;
            const double *tb = bottom ? p->bmr + ((size_t)m * numr + (jg - 1)) * npts : NULL;
            const double *vc = tmpc + (size_t)m * npts, *vs = tmps + (size_t)m * npts;
            double a = 0, b = 0, c = 0, d = 0;
            for (size_t pt = 0; pt < npts; ++pt) {
                a += tt[pt] * vc[pt]; b += tt[pt] * vs[pt];
            }
            if (bottom)
                for (size_t pt = 0; pt < npts; ++pt) {
                    c += tb[pt] * vc[pt]; d += tb[pt] * vs[pt];
                }
            stc[(size_t)m * numr + jg - 1] = a; sts[(size_t)m * numr + jg - 1] = b;
            sbc[(size_t)m * numr + jg - 1] = c; sbs[(size_t)m * numr + jg - 1] = d;
        }
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
    for (int m = 0; m < mmax; ++m)
        for (int kg = 2; kg <= numz - 1; ++kg) {
            const double *ts = p->smz + ((size_t)m * numz + (kg - 1)) * npts;
            const double *vc = tmpc + (size_t)m * npts, *vs = tmps + (size_t)m * npts;
            double a = 0, b = 0;
            for (size_t pt = 0; pt < npts; ++pt) { a += ts[pt] * vc[pt]; b += ts[pt] * vs[pt]; }
            ssc[(size_t)m * numz + kg - 1] = a; sss[(size_t)m * numz + kg - 1] = b;
        }
    /* face rebuild: bessel.F90:399-488 */
    double *pott = (double *)xcalloc((size_t)numr * n, sizeof(double));
    double *potb = (double *)xcalloc((size_t)numr * n, sizeof(double));
    double *pots = (double *)xcalloc((size_t)numz * n, sizeof(double));
    for (int L = 1; L <= n; ++L) {
        for (int J = 2; J <= numr - 1; ++J) {
            double t = stc[J - 1], b = sbc[J - 1];
            for (int M = 2; M <= mmax; ++M) {
                double bc = p->bes_cos[(size_t)(M - 1) * n + L - 1], bs = p->bes_sin[(size_t)(M - 1) * n + L - 1];
                t = t + 2.0 * (stc[(size_t)(M - 1) * numr + J - 1] * bc) + 2.0 * (sts[(size_t)(M - 1) * numr + J - 1] * bs);
                b = b + 2.0 * (sbc[(size_t)(M - 1) * numr + J - 1] * bc) + 2.0 * (sbs[(size_t)(M - 1) * numr + J - 1] * bs);
            }
            pott[(size_t)(L - 1) * numr + J - 1] = factor * t;
            potb[(size_t)(L - 1) * numr + J - 1] = factor * b;
        }
        for (int K = 2; K <= numz - 1; ++K) {
            double s = ssc[K - 1];
            for (int M = 2; M <= mmax; ++M) {
                double bc = p->bes_cos[(size_t)(M - 1) * n + L - 1], bs = p->bes_sin[(size_t)(M - 1) * n + L - 1];
                s = s + 2.0 * (ssc[(size_t)(M - 1) * numz + K - 1] * bc) + 2.0 * (sss[(size_t)(M - 1) * numz + K - 1] * bs);
            }
            pots[(size_t)(L - 1) * numz + K - 1] = factor * s;
        }
    }
    /* axis ghost by a pi rotation (bessel.F90:418-427, 481-486), a mirror
       for isym = 3; the side face's K = 1 mirrors K = 2 for isym = 2, 3
       (bessel.F90:453-455) */
    if (p->isym == 3) {
        for (int L = 1; L <= n; ++L) pott[(size_t)(L - 1) * numr] = pott[(size_t)(L - 1) * numr + 1];
    } else {
        for (int L = 1; L <= n / 2; ++L) {
            pott[(size_t)(L - 1) * numr] = pott[(size_t)(L - 1 + n / 2) * numr + 1];
            pott[(size_t)(L - 1 + n / 2) * numr] = pott[(size_t)(L - 1) * numr + 1];
            potb[(size_t)(L - 1) * numr] = potb[(size_t)(L - 1 + n / 2) * numr + 1];
            potb[(size_t)(L - 1 + n / 2) * numr] = potb[(size_t)(L - 1) * numr + 1];
        }
    }
    if (!bottom)
        for (int L = 1; L <= n; ++L) pots[(size_t)(L - 1) * numz] = pots[(size_t)(L - 1) * numz + 1];
    /* write faces in the reference order: top, side, bottom (isym = 1) */
    for (int L = 1; L <= n; ++L) {
        for (int J = 1; J <= numr; ++J) potp[IDX(p, J, numz, L)] = pott[(size_t)(L - 1) * numr + J - 1];
        for (int K = 1; K <= numz; ++K) potp[IDX(p, numr, K, L)] = pots[(size_t)(L - 1) * numz + K - 1];
        if (bottom)
            for (int J = 1; J <= numr; ++J) potp[IDX(p, J, 1, L)] = potb[(size_t)(L - 1) * numr + J - 1];
    }
    free(tmpc); free(tmps); free(stc); free(sts); free(sbc); free(sbs); free(ssc); free(sss);
    free(pott); free(potb); free(pots);
}

/* pseudo-timestep schedule (helmadi.F90:124-129) */
void bho_dt_schedule(const bho_plan *p, int nsteps, double *dt)
{
    const int numr = p->numr;
    dt[nsteps - 1] = 4.0 * (p->rinv_g[numr] * p->rinv_g[numr]);
    if (nsteps > 1) {
        double alph = pow(p->r_g[numr] * p->rinv_g[3], 2.0 / ((double)nsteps - 1.0));
        for (int i = nsteps - 2; i >= 0; --i) dt[i] = alph * dt[i + 1];
    }
}

/* One Peaceman-Rachford step on every mode plane (helmadi.F90:195-353).
   ffrho, ffpot: (numr, numz, numphi) in mode order.  work: same size. */
static void bho_adi_step(const bho_plan *p, double dtt, const double *ffrho, double *ffpot, double *known)
{
    const int numr = p->numr, numz = p->numz, n = p->numphi;
    const double gamma = p->gamma, four_pi = p->four_pi;
    const double dtt_mtwogamma = dtt - 2.0 * gamma;
    /* radial half-step: RHS then tridagr, per mode */
#pragma omp parallel
    {
        double *bet = (double *)malloc(sizeof(double) * (numr + 1));
        double *gam = (double *)malloc(sizeof(double) * (numr + 1));
        double *br = (double *)malloc(sizeof(double) * (numr + 1));
#pragma omp for schedule(static)
        for (int L = 1; L <= n; ++L) {
            const double *brb = p->brb + (size_t)(L - 1) * (numr + 1);
            for (int J = 1; J <= numr; ++J) br[J] = brb[J] + dtt;
            for (int K = 2; K <= numz - 1; ++K)
                for (int J = 2; J <= numr - 1; ++J)
                    known[IDX(p, J, K, L)] = -four_pi * ffrho[IDX(p, J, K, L)] +
                                             dtt_mtwogamma * ffpot[IDX(p, J, K, L)] +
                                             gamma * (ffpot[IDX(p, J, K + 1, L)] + ffpot[IDX(p, J, K - 1, L)]);
            if (p->isym != 1) /* K = 2 next to the equator: helmadi.F90:211-238 */
                for (int J = 2; J <= numr - 1; ++J)
                    known[IDX(p, J, 2, L)] = -four_pi * ffrho[IDX(p, J, 2, L)] + (dtt - gamma) * ffpot[IDX(p, J, 2, L)] +
                                             gamma * ffpot[IDX(p, J, 3, L)];
            for (int K = 2; K <= numz - 1; ++K)
                known[IDX(p, numr - 1, K, L)] = known[IDX(p, numr - 1, K, L)] -
                                                p->alphar[numr - 1] * ffpot[IDX(p, numr, K, L)];
            /* tridagr.F90:41-69 */
            bet[2] = br[2];
            for (int K = 2; K <= numz - 1; ++K) ffpot[IDX(p, 2, K, L)] = known[IDX(p, 2, K, L)] / bet[2];
            for (int J = 3; J <= numr - 1; ++J) {
                gam[J] = p->cr[J - 1] / bet[J - 1];
                bet[J] = br[J] - p->ar[J] * gam[J];
                for (int K = 2; K <= numz - 1; ++K)
                    ffpot[IDX(p, J, K, L)] = (known[IDX(p, J, K, L)] - p->ar[J] * ffpot[IDX(p, J - 1, K, L)]) / bet[J];
            }
            for (int K = 2; K <= numz - 1; ++K)
                for (int J = numr - 2; J >= 2; --J)
                    ffpot[IDX(p, J, K, L)] = ffpot[IDX(p, J, K, L)] - gam[J + 1] * ffpot[IDX(p, J + 1, K, L)];
        }
        free(bet); free(gam); free(br);
    }
    /* vertical half-step: RHS then tridagz (helmadi.F90:272-343) */
    double *bz = (double *)malloc(sizeof(double) * (numz + 1));
    double *betz = (double *)malloc(sizeof(double) * (numz + 1));
    double *gamz = (double *)malloc(sizeof(double) * (numz + 1));
    for (int K = 1; K <= numz; ++K) bz[K] = p->bzb[K] + dtt;
    const double az = -gamma, cz = -gamma;
    betz[2] = bz[2];
    for (int K = 3; K <= numz - 1; ++K) {
        gamz[K] = cz / betz[K - 1];
        betz[K] = bz[K] - az * gamz[K];
    }
#pragma omp parallel for schedule(static)
    for (int L = 1; L <= n; ++L) {
        const double *elzb = p->elambdazb + (size_t)(L - 1) * (numr + 1);
        for (int K = 2; K <= numz - 1; ++K) {
            for (int J = 3; J <= numr - 1; ++J) {
                double elz = elzb[J] + dtt;
                known[IDX(p, J, K, L)] = -four_pi * ffrho[IDX(p, J, K, L)] + elz * ffpot[IDX(p, J, K, L)] -
                                         p->alphar[J] * ffpot[IDX(p, J + 1, K, L)] -
                                         p->betar[J] * ffpot[IDX(p, J - 1, K, L)];
            }
            double elz2 = elzb[2] + dtt;
            known[IDX(p, 2, K, L)] = -four_pi * ffrho[IDX(p, 2, K, L)] + elz2 * ffpot[IDX(p, 2, K, L)] -
                                     p->alphar[2] * ffpot[IDX(p, 3, K, L)];
        }
        for (int J = 2; J <= numr - 1; ++J) {
            known[IDX(p, J, numz - 1, L)] = known[IDX(p, J, numz - 1, L)] + gamma * ffpot[IDX(p, J, numz, L)];
            if (p->isym == 1) /* the bottom face: helmadi.F90:334-340 */
                known[IDX(p, J, 2, L)] = known[IDX(p, J, 2, L)] + gamma * ffpot[IDX(p, J, 1, L)];
        }
        for (int J = 2; J <= numr - 1; ++J) ffpot[IDX(p, J, 2, L)] = known[IDX(p, J, 2, L)] / betz[2];
        for (int K = 3; K <= numz - 1; ++K)
            for (int J = 2; J <= numr - 1; ++J)
                ffpot[IDX(p, J, K, L)] = (known[IDX(p, J, K, L)] - az * ffpot[IDX(p, J, K - 1, L)]) / betz[K];
        for (int K = numz - 2; K >= 2; --K)
            for (int J = 2; J <= numr - 1; ++J)
                ffpot[IDX(p, J, K, L)] = ffpot[IDX(p, J, K, L)] - gamz[K + 1] * ffpot[IDX(p, J, K + 1, L)];
    }
    free(bz); free(betz); free(gamz);
}

/* helmadi(nsteps) on rhop / potp (helmadi.F90:121-411) */
void bho_helmadi(const bho_plan *p, int nsteps, double *rhop, double *potp)
{
    const int numr = p->numr, numz = p->numz, n = p->numphi;
    const size_t tot = (size_t)numr * numz * n;
    double *dt = (double *)malloc(sizeof(double) * (nsteps > 0 ? nsteps : 1));
    double *ffrho = (double *)malloc(sizeof(double) * tot);
    double *ffpot = (double *)malloc(sizeof(double) * tot);
    double *known = (double *)xcalloc(tot, sizeof(double));
    bho_dt_schedule(p, nsteps, dt);
    dft_columns(p, rhop, ffrho, 0);
    dft_columns(p, potp, ffpot, 0);
    for (int i = 0; i < nsteps; ++i) bho_adi_step(p, dt[i], ffrho, ffpot, known);
    dft_columns(p, ffpot, potp, 1);
    /* axis: helmadi.F90:396-405 (a mirror for isym = 3) */
    if (p->isym == 3) {
        for (int L = 1; L <= n; ++L)
            for (int K = 1; K <= numz; ++K) potp[IDX(p, 1, K, L)] = potp[IDX(p, 2, K, L)];
    } else {
        for (int L = 1; L <= n / 2; ++L)
            for (int K = 1; K <= numz; ++K) {
                potp[IDX(p, 1, K, L)] = potp[IDX(p, 2, K, L + n / 2)];
                potp[IDX(p, 1, K, L + n / 2)] = potp[IDX(p, 2, K, L)];
            }
    }
    /* equator: helmadi.F90:407-411 */
    if (p->isym != 1)
        for (int L = 1; L <= n; ++L)
            for (int J = 1; J <= numr; ++J) potp[IDX(p, J, 1, L)] = potp[IDX(p, J, 2, L)];
    free(dt); free(ffrho); free(ffpot); free(known);
}

/* potential_solver (potential_solver.F90:58-106).
   potp: in = previous potential (used when tstep > 1), out = new potential.
   Returns nsteps used. */
int bho_potential_solver(const bho_plan *p, const double *rho, double *potp, int tstep)
{
    const int numr = p->numr, numz = p->numz, n = p->numphi;
    const size_t tot = (size_t)numr * numz * n;
    double *rhop = (double *)malloc(sizeof(double) * tot);
    memcpy(rhop, rho, sizeof(double) * tot);
    for (int L = 1; L <= n; ++L) {
        for (int J = 1; J <= numr; ++J) {
            rhop[IDX(p, J, numz - 1, L)] = 0.0; rhop[IDX(p, J, numz, L)] = 0.0;
        }
        for (int K = 1; K <= numz; ++K) {
            rhop[IDX(p, numr - 1, K, L)] = 0.0; rhop[IDX(p, numr, K, L)] = 0.0;
        }
        if (p->isym == 1)
            for (int J = 1; J <= numr; ++J) {
                rhop[IDX(p, J, 1, L)] = 0.0; rhop[IDX(p, J, 2, L)] = 0.0;
            }
    }
    int nsteps;
    if (tstep > 1) nsteps = 5;
    else {
        nsteps = 20;
        memset(potp, 0, sizeof(double) * tot);
    }
    bho_bessel(p, rho, potp);
    bho_helmadi(p, nsteps, rhop, potp);
    free(rhop);
    return nsteps;
}

/* helmadi with an explicit step count (for tests of the ADI alone) */
void bho_helmadi_n(const bho_plan *p, int nsteps, double *rhop, double *potp)
{
    bho_helmadi(p, nsteps, rhop, potp);
}

/* accessors for tests */
int bho_numr(const bho_plan *p) { return p->numr; }
int bho_numz(const bho_plan *p) { return p->numz; }
int bho_numphi(const bho_plan *p) { return p->numphi; }
int bho_mmax(const bho_plan *p) { return p->mmax; }
const double *bho_tmr(const bho_plan *p) { return p->tmr; }
const double *bho_bmr(const bho_plan *p) { return p->bmr; }
const double *bho_smz(const bho_plan *p) { return p->smz; }
const double *bho_coefh(const bho_plan *p) { return p->coefh; }
const double *bho_dcoefh(const bho_plan *p) { return p->dcoefh; }
const double *bho_zhf_g(const bho_plan *p) { return p->zhf_g + 1; }
const double *bho_rhf_g(const bho_plan *p) { return p->rhf_g + 1; }
void bho_qhalf_at(const bho_plan *p, double x, double *q) { bho_qhalf(p, x, q); }
int bho_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void bho_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
