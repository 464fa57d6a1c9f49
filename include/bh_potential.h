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
/* As BH_PREC_MIXED, but the top/bottom Green's table is stored in FP32 (its
 * products and sums stay FP64): half the bytes of the boundary contraction,
 * about 1e-10 more error.  An opt-in variant; BH_PREC_MIXED keeps every
 * Green's table in FP64. */
#define BH_PREC_MIXED_TABLE32 2
/* The truncated Peaceman-Rachford iteration in its exact spectral form (the
 * mixed solve's), every product and sweep in FP64 with no refinement: the
 * same potential as BH_PREC_FP64 to rounding, by another sequence of
 * kernels (sine transforms along z, one dense product per azimuthal mode
 * along r).  isym = 1 only. */
#define BH_PREC_SPECTRAL64 3

/* Geometry and solver constants.  Replaces the compile-time parameters of
 * runhydro.h:1-30 and pot.h:3-7 (numr, numz, numphi; mmax = 30,
 * hypr_cutoff = 1.025, hypr_upr_bnd = 110) and the run-time flag isym
 * (setup_frac.F90:313).  Fields left 0 take the reference's value. */
typedef struct bh_grid_desc {
    int numr;          /* radial extent incl. axis ghost and outer boundary  */
    int numz;          /* vertical extent incl. bottom and top boundaries    */
    int numphi;        /* azimuthal zones, a power of two                    */
    int isym;          /* 1 = no symmetry; 2 = equatorial, 3 = equatorial + pi
                          (the grid starts at the equator: K = 1 mirrors
                          K = 2); 2 and 3 need BH_PREC_FP64; 0 -> 1        */
    double dr;         /* dr = dz; 0 -> 1/(130 - 3) (setup_frac.F90:374-376) */
    int mmax;          /* boundary azimuthal modes, 1..512; 0 -> 30.  Slots
                          past the reference's 30 take Q_{m-1/2} from the
                          downward ratios of the Legendre recurrence (the
                          reference's own upward recurrence and 111-term
                          series fail there); slots below 30 are the
                          reference's                                        */
    int nsteps_cold;   /* ADI steps when tstep <= 1; 0 -> 20                 */
    int nsteps_warm;   /* ADI steps when tstep > 1;  0 -> 5                  */
    int refine_passes; /* mixed plans: FP64 residual + refinement passes after
                          the first FP32 solve, 1..4; 0 -> 2.  FP64 plans
                          ignore it.                                         */
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

/* ---- The source step: the consumer of the potential ----------------------
 * The hydro code's momentum source step, which reads the potential the
 * solve left in HBM, so that it need not be downloaded each step: replaces
 * source(source_time) (source.F90:6-382, with make_source_temp and the
 * boundary routines source_bc.F90 and make_source_temp_bc.F90; comm is the
 * identity on one global block) and the start-up make_source_temp
 * (make_source_temp.F90:6-306, main.F90:315).  isym = 1.  All fields are
 * device (numr, numz, numphi) arrays of the plan's grid, ghost layers
 * included (the COMMONs /kinematic/, /source_temp/, /thermo/, /poisson/). */
typedef struct bh_hydro_fields {
    double *s, *t, *a;       /* radial, vertical, angular momentum densities: updated */
    double *temps, *tempa;   /* written by make_source_temp (source_time 2), read */
    const double *rho, *p;   /* density, pressure */
    const double *pot;       /* the potential, e.g. bh_solve_device's d_pot */
} bh_hydro_fields;
typedef struct bh_source_params {
    double dt;                   /* the (half) step, /timestep/ dt */
    double omega_frame;          /* frame rotation, /rotframe/ */
    int tstep;                   /* > 1000 adds the centre-of-mass acceleration terms */
    int boundary_condition[3];   /* bottom, edge, top: 1 wall, 2 free, other Dirichlet */
    double R_com;                /* /centerofmass/, read when tstep > 1000 */
    const double *cylin_a_com;   /* host (3, numphi), read when tstep > 1000 */
    const double *cylin_v_com;   /* host (3, numphi), read when tstep > 1000 */
} bh_source_params;
/* source(source_time), source_time 1 or 2 (2 first forms temps, tempa).
 * Enqueued on stream; returns once enqueued (with tstep > 1000 it first
 * waits for the stream's earlier work, then copies the centre-of-mass
 * arrays). */
int bh_source_device(bh_plan *plan, int source_time, const bh_hydro_fields *fields,
                     const bh_source_params *params, void *stream);
/* make_source_temp alone (t is not used). */
int bh_make_source_temp_device(bh_plan *plan, const bh_hydro_fields *fields, const bh_source_params *params,
                               void *stream);

/* ---- Continuation files -------------------------------------------------
 * The hydro code's restart files, read and written on the host in the
 * reference's own format: fort.12.<tstep>.<rank>, one per MPI rank, each one
 * unformatted sequential record (4-byte length markers) of that rank's
 * (numr_dd, numz_dd, numphi) blocks, ghost layers included, of the nine
 * fields s, t, a, rho, tau, pot, tempa, frac1, frac2; and fort.13.<tstep>,
 * the root rank's scalars.  Written by ritecont_frac_new.F90:117-158, read
 * by setup_frac.F90:625-653.  Fields here are global (numr, numz, numphi)
 * arrays, Fortran order; rank = column + numr_procs * row. */
#define BH_CONT_NFIELDS 9
typedef struct bh_cont_layout {
    int numr, numz, numphi;
    int numr_procs, numz_procs;  /* the writer's process grid */
} bh_cont_layout;
/* Writes <dir>/fort.12.<tstep>.<rank> for every rank, the step formatted as
 * the reference does (i5, i6 or i7: a blank-padded width below 10000);
 * fields[i] NULL writes zeros.  Replaces ritecont_frac_new.F90:117-152. */
int bh_cont_write(const char *dir, int tstep, const bh_cont_layout *layout, const double *const *fields);
/* Reads <prefix><rank> for every rank (prefix as setup_frac.F90:283's
 * conts_template, e.g. "input/conts/fort.12.12345.") into the global
 * fields; fields[i] NULL skips that field.  Replaces setup_frac.F90:625-641. */
int bh_cont_read(const char *prefix, const bh_cont_layout *layout, double *const *fields);
typedef struct bh_cont_scalars {
    double time;
    int frnum;
    double phi_com, R_com;
    double com[9], v_com[3], a_com[3];
    double *cylin_a_com, *cylin_v_com;  /* caller's arrays of 3 * numphi */
    double delt[3];
} bh_cont_scalars;
/* fort.13.<tstep> in <dir> (ritecont_frac_new.F90:154-158) and its reader
 * (setup_frac.F90:645-653). */
int bh_cont_write_scalars(const char *dir, int tstep, int numphi, const bh_cont_scalars *scalars);
int bh_cont_read_scalars(const char *path, int numphi, bh_cont_scalars *scalars);
/* Reads pot from the fort.12 files of a numr_procs x numz_procs run and
 * makes it the plan's warm-start potential (the restart path of
 * setup_frac.F90:625-641 followed by the next potential_solver call's use
 * of potp).  Synchronous. */
int bh_set_potential_from_cont(bh_plan *plan, const char *prefix, int numr_procs, int numz_procs);

/* As bh_set_potential from a host array, but allocating nothing: the field
 * is streamed to the device in column chunks through a fixed 32 MiB staging
 * array that the library holds on every device it runs on until the
 * process ends, where bh_set_potential allocates a staging buffer of the
 * whole field (numr*numz*numphi doubles) on first use.  No counterpart in
 * the reference; for callers that cannot allocate device memory (tight
 * restart or serving loops, harnesses that forbid hipMalloc).  Plans of
 * any grid size are accepted up to numphi = 65536 (one 64-column chunk of
 * all planes must fit); calls that overlap on one GPU share the array and
 * race.  The plan's device must be current.  Synchronous. */
int bh_set_potential_noalloc(bh_plan *plan, const double *pot);

/* Device memory the plan holds, in bytes. */
size_t bh_plan_device_bytes(const bh_plan *plan);

/* Stage timings in milliseconds (HIP events on the solve's stream), written
 * to ms[0..n-1] in the order: boundary, forward DFT, ADI, inverse DFT,
 * total, and the top/bottom boundary contraction kernel alone.  They are the
 * mean over the solves timed since bh_plan_set_timing(plan, 1) or since the
 * last bh_sync_timing; timing is recorded only while switched on. */
int bh_plan_set_timing(bh_plan *plan, int on);
int bh_last_timings(const bh_plan *plan, double *ms, int n);
/* Wait for the timed solves enqueued by bh_solve_device and average their
 * stage timings (bh_solve does this by itself after each solve). */
int bh_sync_timing(bh_plan *plan);

/* Message of the last failure on this thread ("" if none). */
const char *bh_last_error(void);

/* ---- Batches of independent grids --------------------------------------
 * No counterpart in the reference, which solves one grid per run.  A batch
 * holds one plan per grid (grids may differ in size; plans of equal size
 * share their Green's tables and transform matrices) and solves all grids
 * together: step k of every grid's solve runs in one persistent grouped
 * launch per kernel kind, over the tiles of all grids. */
typedef struct bh_batch bh_batch;

#define BH_BATCH_GRAPH 1    /* capture the launch sequence as a hipGraph and replay it */
#define BH_BATCH_AUTOTUNE 2 /* search the launch shapes over the first solves: one tuning
                               axis after another (below), each candidate of the axis
                               timed in one solve at every step it applies to, the
                               fastest kept per step; no choice changes any result */

/* Create a batch of ngrids grids (descs[i] as for bh_plan_create).
 * hbm_budget = 0 means no limit; otherwise creation fails with BH_E_NOMEM
 * when the batch would need more device memory than that. */
int bh_batch_create(const bh_grid_desc *descs, int ngrids, int precision, size_t hbm_budget, int flags,
                    bh_batch **out);
int bh_batch_destroy(bh_batch *batch);

/* One solve of every grid: d_rho[i], d_pot[i] are grid i's device arrays,
 * tstep[i] its cold/warm choice as in bh_solve (tstep = NULL: all cold).
 * Enqueued on `stream`; returns once enqueued (autotuning solves wait for
 * their own timing).  Each grid keeps its potential for warm starts. */
int bh_batch_solve_device(bh_batch *batch, const double *const *d_rho, double *const *d_pot, const int *tstep,
                          void *stream);
/* The same from host arrays; synchronous. */
int bh_batch_solve(bh_batch *batch, const double *const *rho, double *const *pot, const int *tstep);

size_t bh_batch_device_bytes(const bh_batch *batch);

typedef struct bh_batch_stats {
    int ngrids;
    int launches;    /* grouped launches per solve */
    int graph;       /* 1 when solves replay a captured hipGraph */
    int tuning_left; /* autotuning solves still to run (-1 before the first solve) */
    double last_ms;  /* device time of the last timed solve (ms) */
} bh_batch_stats;
/* Statistics; waits for the last timed solve. */
int bh_batch_get_stats(bh_batch *batch, bh_batch_stats *stats);
int bh_batch_set_timing(bh_batch *batch, int on);
/* Autotuning axes: the FP64 and separate-kernel GEMM tile shape (0..4),
 * line-sweep lines per workgroup (8, 16), residual rows per tile (2, 4, 8),
 * persistent workgroups per CU of a grouped launch (0: one per tile) and
 * its tile order (0: XCD-chunked, 1: linear). */
#define BH_TUNE_TILE 0
#define BH_TUNE_NL 1
#define BH_TUNE_ROWS 2
#define BH_TUNE_GRID 3
#define BH_TUNE_ORDER 4
/* The choice made on `axis` for each step of the solve (-1: the axis does
 * not apply there, or was not tuned); writes up to n entries into choice,
 * and into *gain_ms (may be NULL) the time the axis's choices saved over
 * its first (default) candidate, summed over the steps, in the solve that
 * timed them.  Returns the step count. */
int bh_batch_tuning(const bh_batch *batch, int axis, int *choice, int n, double *gain_ms);
/* Grid i's plan (owned by the batch: do not destroy), e.g. for
 * bh_set_potential. */
int bh_batch_plan(bh_batch *batch, int i, bh_plan **plan);

/* ---- Serving --------------------------------------------------------------
 * No counterpart in the reference.  A server runs requests of different grid
 * sizes, each a self-consistent-field (SCF) sequence of dependent solves: a
 * cold first solve, then warm solves from the potential the request's plan
 * keeps on the device, on a density that changes slowly from solve to solve
 * and is generated on the device.  A worker thread runs continuous batching:
 * every loop iteration takes the next solve of each active request and runs
 * them together as grouped launches; requests that arrive meanwhile are
 * admitted, first come first served, as soon as a slot and device memory
 * (hbm_budget) are free.  Plans of finished requests are kept per grid size
 * for reuse and released when memory is needed. */
typedef struct bh_server bh_server;

/* Binary of two bipolytropes (core of index n_core out to core_frac of the
 * radius, envelope of index n_env; density and pressure continuous), centred
 * on the x axis in the equatorial plane about their centre of mass (masses
 * taken as rho_c R^3).  Solve i uses the separation
 * separation * (1 + drift * i).  Zero fields take: rho_c 1, n_core 3,
 * n_env 1.5, core_frac 0.3. */
typedef struct bh_binary_model {
    double rho_c[2];      /* central densities */
    double radius[2];     /* radii, in cells */
    double separation;    /* centre distance at the first solve, in cells */
    double drift;         /* relative change of the separation per solve */
    double n_core, n_env, core_frac;
    double noise;         /* noise amplitude inside the stars, relative to rho_c */
    unsigned long long seed;
} bh_binary_model;

typedef struct bh_request {
    bh_grid_desc grid;
    int nsolves;            /* length of the SCF sequence */
    bh_binary_model model;
    double *pot_out;        /* optional host (numr, numz, numphi) array: the last potential */
    double *solve_ms;       /* optional host array [nsolves]: per-solve latencies (ms) */
    const double *start_pot; /* optional host (numr, numz, numphi) potential to resume from
                                (e.g. read by bh_cont_read): the first solve is then warm
                                from it; read during bh_server_submit */
} bh_request;

typedef struct bh_request_result {
    long long id;
    int status;             /* BH_OK or a BH_E_* code */
    int nsolves;            /* solves completed */
    double submit_s, admit_s, done_s;  /* seconds since the server started */
} bh_request_result;

typedef struct bh_server_config {
    size_t hbm_budget;  /* device memory the server may hold; 0 -> 90% of what is free */
    int max_active;     /* requests on the device at once; 0 -> 64 */
    int precision;      /* a BH_PREC_* value */
} bh_server_config;

typedef struct bh_server_stats {
    long long submitted, completed, solves, iterations;
    int active, waiting, peak_active;
    size_t bytes_in_use, peak_bytes;
    double elapsed_s;        /* first submission to last completion */
    double p50_ms, p99_ms, mean_ms;  /* per-solve latency: completion minus the
                                        completion of the request's previous solve
                                        (minus submission for the first solve) */
} bh_server_stats;

int bh_server_create(const bh_server_config *config, bh_server **out);
/* Queue a request (thread-safe); *id identifies it.  Requests that could
 * not fit in hbm_budget even alone are refused with BH_E_NOMEM. */
int bh_server_submit(bh_server *server, const bh_request *request, long long *id);
/* Wait for request id to finish; fills *result (may be NULL).  Each id can
 * be waited for once: a second call for it, or one while another thread
 * waits for it, returns BH_E_STATE at once. */
int bh_server_wait(bh_server *server, long long id, bh_request_result *result);
/* Wait until every submitted request has finished. */
int bh_server_drain(bh_server *server);
int bh_server_get_stats(bh_server *server, bh_server_stats *stats);
/* Finish the queued requests, stop the worker, release everything. */
int bh_server_destroy(bh_server *server);

#ifdef __cplusplus
}
#endif

#endif /* BH_POTENTIAL_H */
