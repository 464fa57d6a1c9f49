/*
 * bh_potential.h -- C ABI of the MI355X self-gravity solve (libbh_potential.so).
 *
 * Drop-in for bipoly_hydro's potential_solver path.  The reference exposes
 * that path as two argument-less Fortran subroutines working on COMMON
 * blocks; each entry point below names the reference interface it replaces.
 * A Fortran driver binds these through ISO_C_BINDING (bh_potential.f90 in
 * the package's fortran/ folder).
 *
 * Arrays are the reference's: Fortran column-major (numr, numz, numphi),
 * r fastest, including the ghost/boundary layers J = 1, J = numr, K = 1,
 * K = numz (potential_solver.F90:18-22, runhydro.h:1-30 with one rank).
 *
 * Every function returns BH_OK (0) or a negative BH_E* code; the message of
 * the last failure on the calling thread is returned by bh_last_error().
 * Nothing here exits or aborts.
 */
#ifndef BH_POTENTIAL_H
#define BH_POTENTIAL_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BH_OK 0
#define BH_E_INVALID -1  /* bad descriptor or argument */
#define BH_E_NOMEM -2    /* device or host allocation failed / over budget */
#define BH_E_DEVICE -3   /* HIP runtime error */
#define BH_E_NONFINITE -4 /* rho holds NaN or Inf */
#define BH_E_STATE -5    /* call not valid in the plan's current state */

#define BH_PREC_FP64 0  /* plain path: FP64 throughout */
#define BH_PREC_MIXED 1 /* FP32 sweeps / DFT, FP64 tables, residual, refinement */

/* Geometry and solver constants.  Replaces the compile-time parameters of
 * runhydro.h:1-30 and pot.h:3-7 (numr, numz, numphi; mmax = 30,
 * hypr_cutoff = 1.025, hypr_upr_bnd = 110) and the run-time flag isym
 * (setup_frac.F90:313).  Fields left 0 take the reference's value. */
typedef struct bh_grid_desc {
    int numr;          /* radial extent incl. axis ghost and outer boundary  */
    int numz;          /* vertical extent incl. bottom and top boundaries    */
    int numphi;        /* azimuthal zones, a power of two                    */
    int isym;          /* 1 = no symmetry (the only value accepted)          */
    double dr;         /* dr = dz; 0 -> 1/(130 - 3) (setup_frac.F90:374-376) */
    int mmax;          /* boundary azimuthal modes; 0 -> 30                  */
    int nsteps_cold;   /* ADI steps when tstep <= 1; 0 -> 20                 */
    int nsteps_warm;   /* ADI steps when tstep > 1;  0 -> 5                  */
} bh_grid_desc;

typedef struct bh_plan bh_plan;

/* Build a plan for one geometry on the current HIP device: coefficient
 * arrays, pseudo-timestep schedules, DFT matrices and the Green's-function
 * boundary tables, all resident in device memory.  Replaces potsetup
 * (potsetup.F90:6-255, called from setup_frac.F90:689).  hbm_budget = 0
 * means no limit; otherwise creation fails with BH_E_NOMEM when the plan
 * would need more device memory than that. */
int bh_plan_create(const bh_grid_desc *desc, int precision, size_t hbm_budget, bh_plan **out);

/* Release every device and host resource of the plan. */
int bh_plan_destroy(bh_plan *plan);

/* One self-gravity solve from host arrays.  Replaces potential_solver
 * (potential_solver.F90:6-109; callers main.F90:395-397 and
 * setup_frac.F90:690).  rho: density, pot: receives the potential, both
 * (numr, numz, numphi) host arrays.  tstep <= 1 is a cold start (20 ADI
 * steps from zero), tstep > 1 a warm start (5 steps) from the potential the
 * plan kept from its previous solve, as potp persists in the reference
 * (potential_solver.F90:80-94).  Synchronous. */
int bh_solve(bh_plan *plan, const double *rho, double *pot, int tstep);

/* The same with device pointers, enqueued on `stream` (a hipStream_t; NULL
 * = the null stream).  Returns once the work is enqueued. */
int bh_solve_device(bh_plan *plan, const double *d_rho, double *d_pot, int tstep, void *stream);

/* Replace the plan's warm-start potential (the reference's potp) with a
 * potential given as a (numr, numz, numphi) array, e.g. one read back from a
 * continuation file (setup_frac.F90:625-666).  on_device selects whether pot
 * is a device or a host pointer.  Synchronous. */
int bh_set_potential(bh_plan *plan, const double *pot, int on_device);

/* Device memory the plan holds, in bytes. */
size_t bh_plan_device_bytes(const bh_plan *plan);

/* Stage timings of the last solve in milliseconds (HIP events), written to
 * ms[0..n-1] in the order: boundary, forward DFT, ADI, inverse DFT, total.
 * Timing is recorded only after bh_plan_set_timing(plan, 1). */
int bh_plan_set_timing(bh_plan *plan, int on);
int bh_last_timings(const bh_plan *plan, double *ms, int n);
/* Wait for the last timed solve enqueued by bh_solve_device and read its
 * stage timings (bh_solve does this by itself). */
int bh_sync_timing(bh_plan *plan);

/* Message of the last failure on this thread ("" if none). */
const char *bh_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* BH_POTENTIAL_H */
