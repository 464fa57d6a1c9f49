/*
 * bh_oracle_source.c -- CPU restatement of bipoly_hydro's momentum source
 * step, the consumer of the potential, in plain C99 + OpenMP, FP64.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP source
 * step is compared against.  Only tests/ may load it (through
 * libbh_oracle.so); the product library never links or calls it.
 *
 * It restates, for one global (r, z, phi) block (the single-rank collapse:
 * comm is the identity), with isym = 1:
 *
 *   source.F90:144-379            phieff, the vertical, radial and
 *                                 azimuthal momentum sources
 *   make_source_temp.F90:113-303  temps, tempa (source_time = 2, and the
 *                                 start-up call of main.F90:315)
 *   source_bc.F90:45-233          boundary conditions of s, t, a
 *   make_source_temp_bc.F90:40-190 boundary conditions of temps, tempa
 *
 * Arrays are global (numr, numz, numphi) Fortran order, ghost layers
 * included.  Grid: r(J) = (J - 2) dr, rhf = r + dr/2, rinv = 1/r (0 on the
 * axis), dz = dr, dphi = 2 pi / numphi (setup_frac.F90:374-463).
 */
#include <math.h>
#include <stdlib.h>

#define IX(J, K, L) ((((size_t)(L) - 1) * numz + ((K) - 1)) * numr + ((J) - 1))

typedef struct {
    int numr, numz, numphi;
    double dr, drinv, dzinv, dphiinv;
    double *r, *rhf, *rinv, *rhfinv; /* 1-based, [numr + 1] */
} bho_sgrid;

static void sgrid_init(bho_sgrid *g, int numr, int numz, int numphi, double dr)
{
    g->numr = numr;
    g->numz = numz;
    g->numphi = numphi;
    g->dr = dr;
    const double dz = dr, dphi = 2.0 * acos(-1.0) * (1.0 / (double)numphi);
    g->drinv = 1.0 / dr;
    g->dzinv = 1.0 / dz;
    g->dphiinv = 1.0 / dphi;
    g->r = calloc(numr + 1, sizeof(double));
    g->rhf = calloc(numr + 1, sizeof(double));
    g->rinv = calloc(numr + 1, sizeof(double));
    g->rhfinv = calloc(numr + 1, sizeof(double));
    for (int J = 1; J <= numr; ++J) {
        g->r[J] = ((double)J - 2.0) * dr;
        g->rhf[J] = g->r[J] + 0.5 * dr;
        g->rinv[J] = g->r[J] != 0.0 ? 1.0 / g->r[J] : 0.0;
        g->rhfinv[J] = 1.0 / g->rhf[J];
    }
}

static void sgrid_free(bho_sgrid *g)
{
    free(g->r);
    free(g->rhf);
    free(g->rinv);
    free(g->rhfinv);
}

/* phieff = pot - 0.5 Omega^2 rhf^2 everywhere (source.F90:149-160) */
static double *make_phieff(const bho_sgrid *g, const double *pot, double omega)
{
    const int numr = g->numr, numz = g->numz, numphi = g->numphi;
    double *pe = malloc((size_t)numr * numz * numphi * sizeof(double));
#pragma omp parallel for
    for (int L = 1; L <= numphi; ++L)
        for (int K = 1; K <= numz; ++K)
            for (int J = 1; J <= numr; ++J)
                pe[IX(J, K, L)] = pot[IX(J, K, L)] - 0.5 * omega * omega * g->rhf[J] * g->rhf[J];
    return pe;
}

/* the axis, bottom, edge and top conditions of a (radial, azimuthal) pair,
   in that order: source_bc.F90 for (s, a) with t, make_source_temp_bc.F90
   for (temps, tempa) (t null) */
static void apply_bc(const bho_sgrid *g, const int bc[3], double *s, double *t, double *a)
{
    const int numr = g->numr, numz = g->numz, numphi = g->numphi, h = numphi / 2;
    const double ninv = 1.0 / (double)numphi;
    /* axis: source_bc.F90:45-93, make_source_temp_bc.F90:40-74 */
    for (int L = 1; L <= numphi; ++L)
        for (int K = 2; K <= numz - 1; ++K) s[IX(2, K, L)] = 0.0;
    for (int K = 2; K <= numz - 1; ++K) {
        double sum = 0.0;
        for (int L = 1; L <= numphi; ++L) sum = sum + a[IX(2, K, L)];
        sum = sum * ninv;
        for (int L = 1; L <= numphi; ++L) a[IX(2, K, L)] = sum;
        if (t) {
            sum = 0.0;
            for (int L = 1; L <= numphi; ++L) sum = sum + t[IX(2, K, L)];
            sum = sum * ninv;
            for (int L = 1; L <= numphi; ++L) t[IX(2, K, L)] = sum;
        }
    }
    for (int L = 1; L <= h; ++L)
        for (int K = 2; K <= numz - 1; ++K) {
            s[IX(1, K, L)] = -s[IX(3, K, L + h)];
            s[IX(1, K, L + h)] = -s[IX(3, K, L)];
        }
    if (t)
        for (int L = 1; L <= h; ++L)
            for (int K = 2; K <= numz - 1; ++K) {
                t[IX(1, K, L)] = t[IX(2, K, L + h)];
                t[IX(1, K, L + h)] = t[IX(2, K, L)];
            }
    for (int L = 1; L <= h; ++L)
        for (int K = 2; K <= numz - 1; ++K) {
            a[IX(1, K, L)] = a[IX(2, K, L + h)];
            a[IX(1, K, L + h)] = a[IX(2, K, L)];
        }
    /* bottom (isym = 1): source_bc.F90:95-148, make_source_temp_bc.F90:76-110 */
    for (int L = 1; L <= numphi; ++L)
        for (int J = 2; J <= numr - 1; ++J) {
            if (bc[0] == 1) {
                s[IX(J, 1, L)] = s[IX(J, 2, L)];
                if (t) {
                    t[IX(J, 1, L)] = -t[IX(J, 3, L)];
                    t[IX(J, 2, L)] = 0.0;
                }
                a[IX(J, 1, L)] = a[IX(J, 2, L)];
            } else if (bc[0] == 2) {
                s[IX(J, 1, L)] = s[IX(J, 2, L)];
                a[IX(J, 1, L)] = a[IX(J, 2, L)];
                if (t) t[IX(J, 1, L)] = t[IX(J, 2, L)];
            } else {
                s[IX(J, 1, L)] = 0.0;
                a[IX(J, 1, L)] = 0.0;
                if (t) {
                    t[IX(J, 1, L)] = 0.0;
                    t[IX(J, 2, L)] = 0.5 * (fabs(t[IX(J, 2, L)]) - t[IX(J, 2, L)]);
                }
            }
        }
    /* edge: source_bc.F90:165-196, make_source_temp_bc.F90:112-143 */
    for (int L = 1; L <= numphi; ++L)
        for (int K = 2; K <= numz - 1; ++K) {
            if (bc[1] == 1) {
                s[IX(numr, K, L)] = -s[IX(numr - 2, K, L)];
                s[IX(numr - 1, K, L)] = 0.0;
                a[IX(numr, K, L)] = a[IX(numr - 1, K, L)];
                if (t) t[IX(numr, K, L)] = t[IX(numr - 1, K, L)];
            } else if (bc[1] == 2) {
                s[IX(numr, K, L)] = s[IX(numr - 1, K, L)];
                if (t) t[IX(numr, K, L)] = t[IX(numr - 1, K, L)];
                a[IX(numr, K, L)] = a[IX(numr - 1, K, L)];
            } else {
                s[IX(numr, K, L)] = 0.0;
                s[IX(numr - 1, K, L)] = 0.5 * (fabs(s[IX(numr - 1, K, L)]) + s[IX(numr - 1, K, L)]);
                a[IX(numr, K, L)] = 0.0;
                if (t) t[IX(numr, K, L)] = 0.0;
            }
        }
    /* top: source_bc.F90:198-230, make_source_temp_bc.F90:145-176 */
    for (int L = 1; L <= numphi; ++L)
        for (int J = 2; J <= numr - 1; ++J) {
            if (bc[2] == 1) {
                s[IX(J, numz, L)] = s[IX(J, numz - 1, L)];
                a[IX(J, numz, L)] = a[IX(J, numz - 1, L)];
                if (t) {
                    t[IX(J, numz, L)] = -t[IX(J, numz - 2, L)];
                    t[IX(J, numz - 1, L)] = 0.0;
                }
            } else if (bc[2] == 2) {
                s[IX(J, numz, L)] = s[IX(J, numz - 1, L)];
                if (t) t[IX(J, numz, L)] = t[IX(J, numz - 1, L)];
                a[IX(J, numz, L)] = a[IX(J, numz - 1, L)];
            } else {
                s[IX(J, numz, L)] = 0.0;
                a[IX(J, numz, L)] = 0.0;
                if (t) {
                    t[IX(J, numz, L)] = 0.0;
                    t[IX(J, numz - 1, L)] = 0.5 * (fabs(t[IX(J, numz - 1, L)]) + t[IX(J, numz - 1, L)]);
                }
            }
        }
}

/* radial momentum update from the azimuthal momentum at the radial faces
   (source.F90:225-284 with sa = s, src = tempa; make_source_temp.F90:139-208
   with sa = temps, src = a): out = sin - dt (...) */
static void radial_update(const bho_sgrid *g, int com, double dt, double omega, double R_com, const double *cac,
                          const double *cvc, const double *sin_, const double *src, const double *rho, const double *p,
                          const double *pe, double *out)
{
    const int numr = g->numr, numz = g->numz, numphi = g->numphi;
    const double dr = g->dr;
#pragma omp parallel for
    for (int L = 1; L <= numphi; ++L) {
        const int Lp = L < numphi ? L + 1 : 1;
        for (int K = 2; K <= numz - 1; ++K)
            for (int J = 2; J <= numr - 1; ++J) {
                const double rj = g->r[J], ri = g->rinv[J];
                const double rh = 0.5 * ri * (rho[IX(J, K, L)] * (rj + 0.25 * dr) + rho[IX(J - 1, K, L)] * (rj - 0.25 * dr));
                const double tp = 0.25 * ri * ((src[IX(J, K, L)] + src[IX(J, K, Lp)]) * (rj + 0.25 * dr) +
                                               (src[IX(J - 1, K, L)] + src[IX(J - 1, K, Lp)]) * (rj - 0.25 * dr));
                double v;
                if (rh > 0.0) {
                    double f = g->drinv * ((p[IX(J, K, L)] - p[IX(J - 1, K, L)]) + rh * (pe[IX(J, K, L)] - pe[IX(J - 1, K, L)])) -
                               tp * tp * ri * ri * ri / rh - 2.0 * omega * tp * ri;
                    if (com) f = f + rh * (cac[3 * (L - 1) + 0] - omega * omega * R_com - 2 * omega * cvc[3 * (L - 1) + 1]);
                    v = sin_[IX(J, K, L)] - dt * f;
                } else {
                    v = 0.0;
                }
                out[IX(J, K, L)] = v;
            }
    }
}

/* azimuthal momentum update from the radial momentum at the azimuthal
   faces (source.F90:286-375 with src = temps; make_source_temp.F90:210-297
   with src = s) */
static void azimuthal_update(const bho_sgrid *g, int com, double dt, double omega, const double *cac,
                             const double *cvc, const double *ain, const double *src, const double *rho,
                             const double *p, const double *pe, double *out)
{
    const int numr = g->numr, numz = g->numz, numphi = g->numphi;
    const double dr = g->dr;
#pragma omp parallel for
    for (int L = 1; L <= numphi; ++L) {
        const int Lm = L > 1 ? L - 1 : numphi;
        for (int K = 2; K <= numz - 1; ++K)
            for (int J = 2; J <= numr - 1; ++J) {
                const double rh = 0.5 * (rho[IX(J, K, L)] + rho[IX(J, K, Lm)]);
                const double hf = g->rhf[J];
                const double tp = 0.25 * g->rhfinv[J] *
                                  ((src[IX(J + 1, K, L)] + src[IX(J + 1, K, Lm)]) * (hf + 0.25 * dr) +
                                   (src[IX(J, K, L)] + src[IX(J, K, Lm)]) * (hf - 0.25 * dr));
                double f = g->dphiinv * ((p[IX(J, K, L)] - p[IX(J, K, Lm)]) + rh * (pe[IX(J, K, L)] - pe[IX(J, K, Lm)])) +
                           2.0 * omega * hf * tp;
                if (com) f = f + rh * hf * (cac[3 * (L - 1) + 1] + 2 * omega * cvc[3 * (L - 1) + 0]);
                out[IX(J, K, L)] = ain[IX(J, K, L)] - dt * f;
            }
    }
}

/* make_source_temp: temps and tempa from s, a (make_source_temp.F90:113-303) */
void bho_make_source_temp(int numr, int numz, int numphi, double dr, int tstep, const int *bc, double dt,
                          double omega, double R_com, const double *cac, const double *cvc, const double *s,
                          const double *a, double *temps, double *tempa, const double *rho, const double *p,
                          const double *pot)
{
    bho_sgrid g;
    sgrid_init(&g, numr, numz, numphi, dr);
    double *pe = make_phieff(&g, pot, omega);
    const int com = tstep > 1000;
    radial_update(&g, com, dt, omega, R_com, cac, cvc, s, a, rho, p, pe, temps);
    azimuthal_update(&g, com, dt, omega, cac, cvc, a, s, rho, p, pe, tempa);
    apply_bc(&g, bc, temps, NULL, tempa);
    free(pe);
    sgrid_free(&g);
}

/* source(source_time): source.F90:144-379 */
void bho_source(int numr, int numz, int numphi, double dr, int source_time, int tstep, const int *bc, double dt,
                double omega, double R_com, const double *cac, const double *cvc, double *s, double *t, double *a,
                double *temps, double *tempa, const double *rho, const double *p, const double *pot)
{
    bho_sgrid g;
    sgrid_init(&g, numr, numz, numphi, dr);
    double *pe = make_phieff(&g, pot, omega);
    const int com = tstep > 1000;
    /* vertical momentum (source.F90:172-203) */
#pragma omp parallel for
    for (int L = 1; L <= numphi; ++L)
        for (int K = 2; K <= numz - 1; ++K)
            for (int J = 2; J <= numr - 1; ++J) {
                const double rh = 0.5 * (rho[IX(J, K, L)] + rho[IX(J, K - 1, L)]);
                const double grad = (p[IX(J, K, L)] - p[IX(J, K - 1, L)]) + rh * (pe[IX(J, K, L)] - pe[IX(J, K - 1, L)]);
                if (!com) t[IX(J, K, L)] = t[IX(J, K, L)] - dt * g.dzinv * grad;
                else t[IX(J, K, L)] = t[IX(J, K, L)] - dt * (g.dzinv * grad + rh * cac[3 * (L - 1) + 2]);
            }
    free(pe);
    sgrid_free(&g);
    if (source_time == 2)
        bho_make_source_temp(numr, numz, numphi, dr, tstep, bc, dt, omega, R_com, cac, cvc, s, a, temps, tempa, rho, p,
                             pot);
    sgrid_init(&g, numr, numz, numphi, dr);
    pe = make_phieff(&g, pot, omega);
    const size_t n = (size_t)numr * numz * numphi;
    double *s2 = malloc(n * sizeof(double)), *a2 = malloc(n * sizeof(double));
    for (size_t i = 0; i < n; ++i) {
        s2[i] = s[i];
        a2[i] = a[i];
    }
    radial_update(&g, com, dt, omega, R_com, cac, cvc, s, tempa, rho, p, pe, s2);
    azimuthal_update(&g, com, dt, omega, cac, cvc, a, temps, rho, p, pe, a2);
    for (size_t i = 0; i < n; ++i) {
        s[i] = s2[i];
        a[i] = a2[i];
    }
    apply_bc(&g, bc, s, t, a);
    free(s2);
    free(a2);
    free(pe);
    sgrid_free(&g);
}
