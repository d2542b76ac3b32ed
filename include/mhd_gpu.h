/*
 * mhd_gpu.h -- C ABI of the MI355X (gfx950) RK4 Hall-MHD step.
 *
 * Replaces the argument-less Fortran subroutines of t3dsdsc.f that talk
 * through COMMON blocks:
 *   setdt   (t3dsdsc.f:4694-4810)  -> mhd_compute_dt
 *   stepon  (t3dsdsc.f:1063-1185)  -> mhd_step
 *   initia's xresd capture (t3dsdsc.f:816-834) -> mhd_init_residual
 *   current (t3dsdsc.f:2452-2500)  -> mhd_get_current
 *   diagn   (t3dsdsc.f:1189-1235), diverB (2504-2535) -> mhd_diagnostics
 * and the data the reference keeps in COMMON:
 *   /var21/ x (mhd02.for:4)                    -> mhd_set_state / mhd_get_state
 *   /var11/ rhoi,bxi,bzi,pri, /wrk/ byi       -> mhd_set_equilibrium
 *   /var15/ xx,yy,zz,dx,dy,dz (mhd01.for:70)   -> mhd_create arguments
 *   /cst/ eta0,gamma,cfl (mhd01.for:55-56), /logic/ hall,periody -> mhd_config
 *
 * Layout of every host array: the reference's column-major padded layout,
 * a(mx+4, my, mz+4[, m]) with x fastest (C order a[m][k][j][i]).  Physical
 * points are Fortran indices 3..mx+2 in x and z; the two outer columns per
 * side are never read.  y has no ghosts when non-periodic; when periodic the
 * caller supplies my = ny + 4 rows (two ghost rows per side).
 *
 * Ownership: host pointers are borrowed for the call only.  The handle owns
 * all device memory and one HIP stream; calls return after the work they
 * started has finished, except mhd_step/mhd_advance which return once
 * queued work is complete as well (synchronous API).
 * Errors: int status, 0 = OK, < 0 = error; text via mhd_last_error.
 * Threading: one host thread per handle.
 */
#ifndef MHD_GPU_H
#define MHD_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mhd_config {
    int mx, my, mz;        /* mhd01.for:3-4 (my = full y extent)            */
    int periodic_y;        /* mhd03.for:15 periody                          */
    int hall;              /* mhd03.for:21 hall                             */
    int whistler_dt;       /* add pi*cht*|B|/(rho*dx) to the CFL speeds     */
    double di;             /* Hall coefficient cht (uniform)                */
    double eta0;           /* mhd03.for:61 (foreta, t3dsdsc.f:6507)          */
    double fmu0, frho0, fpr0; /* mhd03.for:62-64 (formu, t3dsdsc.f:6524-6528) */
    double gamma;          /* mhd03.for:60                                  */
    double cfl;            /* mhd03.for:71                                  */
    /* magnetosphere mode (the reference's own default case); any of the four
     * switches selects it; it needs non-periodic y and no viscosity or
     * diffusion.  Zero: the Harris configuration. */
    int magnetosphere;     /* rr0 = rr1 = |(x,y,z)|: inner-sphere mask rr1 <= rcr at
                            * the start of every bndry* (t3dsdsc.f:1692-1707, 2003-2017,
                            * 2549-2559, 2709-2719, 2881-2891), setdt over rr1 >= 1 (4712) */
    int inflow;            /* driven x-left inflow (t3dsdsc.f:1711-1772) instead of the
                            * mirrored x-right rule; face data via mhd_set_inflow_face */
    int radial_hall;       /* cht = 0.5 di (tanh((rr0-1.1)/8) - tanh((rr0-2)/10)), di = chall
                            * (t3dsdsc.f:1398-1399); 0: cht = di */
    int partial_e;         /* totalE = .false.: E from B - B_i plus the curl(v x B0) dipole
                            * terms (t3dsdsc.f:1370-1386, 1460-1548); 0: totalE */
    double rcr;            /* mhd03.for rcr: mask radius */
} mhd_config;

typedef struct mhd_diag {
    /* diagn (t3dsdsc.f:1189-1235): current extrema on the plane jy=my/2+1 */
    double cxmax, cxmin, cymax, cymin, czmax, czmin;
    /* energy integrals over unique physical points (trapezoid weights) */
    double e_kin, e_mag, e_th, e_mag_pert;
    /* diverB + bndrydiv (t3dsdsc.f:2504-2535, 3031-3712) */
    double divb_max, divb_rms;
} mhd_diag;

typedef struct mhd_handle mhd_handle;

enum {
    MHD_OK = 0,
    MHD_ERR_ARG = -1,
    MHD_ERR_HIP = -2,
    MHD_ERR_NOMEM = -3,
    MHD_ERR_DT_COLLAPSE = -4,   /* t3dsdsc.f:142-154 */
    MHD_ERR_BLOWUP = -5         /* t3dsdsc.f:1232    */
};

/* geometry arrays: xx, dx of length mx+4; yy, dy of length my; zz, dz of
 * length mz+4 (input: gridpntnew2 + message1, t3dsdsc.f:248-313) */
int mhd_create(const mhd_config *cfg, const double *xx, const double *yy, const double *zz,
               const double *dx, const double *dy, const double *dz, mhd_handle **out);
/* rhoi, bxi, byi, bzi, pri: each (mx+4, my, mz+4) (initia, t3dsdsc.f:806-814) */
int mhd_set_equilibrium(mhd_handle *h, const double *rhoi, const double *bxi,
                        const double *byi, const double *bzi, const double *pri);
/* x(mx+4, my, mz+4, 8) (mhd02.for:1, /var21/) */
int mhd_set_state(mhd_handle *h, const double *x);
int mhd_get_state(mhd_handle *h, double *x);
/* The Harris-sheet initial state of the benchmark workloads, built on the
 * device from the handle's mesh: it takes the place of mhd_set_equilibrium
 * + mhd_set_state for these cases without 3-D host arrays (no counterpart
 * in the reference, whose initia, t3dsdsc.f:612-850, builds the
 * magnetosphere).  Equilibrium Bx = tanh(z/lam), By = guide, Bz = 0,
 * p = beta_inf/2 + sech^2(z/lam)/2, rho = p/temp; state = equilibrium plus
 * dBx = -d(dAy)/dz, dBz = d(dAy)/dx of dAy = amp exp(-((x-x_offset)^2+z^2)/width^2)
 * and rho v_c = rho noise U(-1,1) with U from the splitmix64 finaliser of
 * (seed, ((c mz + k) ny + j) mx + i) over the unique physical points
 * (0-based i < mx, j < ny unique rows, k < mz, component c < 3).  Image
 * rows of periodic y repeat their unique rows. */
typedef struct mhd_harris {
    double lam, beta_inf, guide, temp, amp, width, noise, x_offset;
    unsigned long long seed;
} mhd_harris;
int mhd_init_harris(mhd_handle *h, const mhd_harris *p);
/* xresd = RHS momentum at t=0 with eta=mu=0 (t3dsdsc.f:816-834) */
int mhd_init_residual(mhd_handle *h);
/* setdt (t3dsdsc.f:4694-4810) */
int mhd_compute_dt(mhd_handle *h, double *dt);
/* mhd_compute_dt without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  The same kernels and result as
 * mhd_compute_dt; the 1024 block partial minima go to a fixed 32 KB device
 * array (4096 doubles) of the library instead of the handle's reduction
 * buffer.  Each GPU has one such array, shared by every handle on it and
 * holding the last call's partials: calls that overlap on one GPU race.
 * Returns MHD_ERR_ARG if setdt needed more partials than the array holds. */
int mhd_compute_dt_noalloc(mhd_handle *h, double *dt);
/* stepon (t3dsdsc.f:1063-1165): one RK4 step of size dt from time */
int mhd_step(mhd_handle *h, double time, double dt);
/* the inflow face data of magnetosphere mode, bxinfc/byinfc/bzinfc(my, mz+4)
 * (common /var17/, read from fcinp/bin* and scaled by -50 at t3dsdsc.f:728-768);
 * zero until set */
int mhd_set_inflow_face(mhd_handle *h, const double *bxinfc, const double *byinfc, const double *bzinfc);
/* the time loop body of program main (t3dsdsc.f:139-176) without I/O:
 * nsteps x (setdt; stepon; time += dt) with dt kept on the device.
 * dt0 > 0 arms the collapse guard dt < 0.1*dt0 (t3dsdsc.f:142): the loop
 * stops at the collapsing setdt as the reference does (a device flag; the
 * steps queued after it leave the state and time alone) and the call
 * returns MHD_ERR_DT_COLLAPSE with the state and *time of before that step
 * and *dt_last = the collapsing dt.  Otherwise *time is advanced and
 * *dt_last holds the last dt.  *elapsed_ms is the HIP-event time of the
 * loop on the handle's stream. */
int mhd_advance(mhd_handle *h, int nsteps, double dt0, double *time, double *dt_last,
                double *elapsed_ms);
/* current (t3dsdsc.f:2452-2500) of the current state, cur(mx+4,my,mz+4,3) */
int mhd_get_current(mhd_handle *h, double *cur);
/* diagn + energy + div B of the current state (returns MHD_ERR_BLOWUP
 * when |max Jy| >= 1e7, t3dsdsc.f:1232, after filling *d) */
int mhd_diagnostics(mhd_handle *h, mhd_diag *d);
/* diverB + bndrydiv field of the current state, (mx+4, my, mz+4) */
int mhd_get_divb(mhd_handle *h, double *divb);

/* stage-level access for parity checks of the RHS (right, t3dsdsc.f:1297-1673):
 * mhd_rhs runs one right(t) on the current state (t == 0 skips the final
 * smoothing as t3dsdsc.f:1669 does); mhd_get_xdif reads the xdif array
 * (8 fields) the last mhd_rhs / mhd_init_residual left behind (a step's
 * last sweep updates the state in place and leaves none: MHD_ERR_ARG);
 * xresd is the frozen momentum residual. */
int mhd_rhs(mhd_handle *h, double t);
/* right(t, ddt, km) with the stage arguments the driven inflow reads
 * (ddt the stage's time step, km = 1: the quotient subtracts xfold, 2: xm);
 * mhd_rhs(h, t) is mhd_rhs_stage(h, t, 1e-9, 1) */
int mhd_rhs_stage(mhd_handle *h, double t, double ddt, int km);
int mhd_get_xdif(mhd_handle *h, double *xdif);
int mhd_get_residual(mhd_handle *h, double *xresd);
int mhd_set_residual(mhd_handle *h, const double *xresd);

/* profiling: per-kernel-family HIP-event times of the last mhd_advance
 * when enabled (names are stable strings; returns the count written) */
int mhd_set_profiling(mhd_handle *h, int on);
int mhd_kernel_times(mhd_handle *h, int max, const char **names, double *ms, int *calls);

/* measured read+write rate of a plain device copy of `bytes` (no handle;
 * the reference point for achieved-bandwidth figures) */
int mhd_copy_rate(long bytes, int iters, double *bytes_per_s);

const char *mhd_last_error(const mhd_handle *h);
void mhd_destroy(mhd_handle *h);

/* Ensemble of independent runs (no counterpart in the reference, which runs
 * one configuration per job; the batched form of the time loop above).
 * The runs are handles the caller created and initialised (equilibrium,
 * state, residual) as for mhd_advance; they stay owned by the caller and
 * must outlive the ensemble.  mhd_ensemble_advance advances every run
 * nsteps x (setdt; stepon; time += dt) with its own device dt, all runs
 * concurrently: each run's step pair is a captured HIP graph replayed on
 * the run's stream.  dt0[r] arms run r's collapse guard (a collapsing run
 * stops there, as in mhd_advance); time[r] is advanced in place; *elapsed_ms is the wall time of the whole batch
 * (HIP events).  Returns the first failing run's status (its text via
 * mhd_ensemble_error); mhd_ensemble_status gives every run's.  While
 * grouped, a run's kernels march longer z chunks (the runs fill the device
 * together); results do not change, and the run's own chunking returns at
 * mhd_ensemble_destroy. */
typedef struct mhd_ensemble mhd_ensemble;
int mhd_ensemble_create(mhd_handle *const *runs, int n, mhd_ensemble **out);
int mhd_ensemble_advance(mhd_ensemble *e, int nsteps, const double *dt0, double *time,
                         double *elapsed_ms);
int mhd_ensemble_status(const mhd_ensemble *e, int *status);
const char *mhd_ensemble_error(const mhd_ensemble *e);
void mhd_ensemble_destroy(mhd_ensemble *e);
/* Tile-shape autotuner of the ensemble (no counterpart in the reference):
 * times `trial` steps of the whole batch (after two untimed steps that
 * capture the graphs) under each z-chunk policy -- 0 long chunks, 1 each
 * run's own, 2 16/16, 3 8/8 planes (tile / smoothing kernels) -- restores
 * every run's state after each trial, and keeps the fastest.  ms[4] (may
 * be NULL) receives the trial times, *best (may be NULL) the choice.  The
 * runs' states and the caller's times are unchanged on return. */
int mhd_ensemble_tune(mhd_ensemble *e, int trial, const double *dt0, const double *time, int *best,
                      double *ms);
/* Tile-shape autotuner of the ensemble (no counterpart in the reference):
 * a coordinate search over the shape of every run's launches -- the replay
 * form, the z chunk of the smoothing kernels (zkc 16..64 planes), of the
 * tile kernels (kc 8..64), the rows of the face windows (12 or 16), the
 * right-part-2 kernels, then the other replay forms -- timing `trial` batch steps per shape (after two
 * untimed steps that capture the graphs) and restoring the states after
 * each.  The fastest shape stays in force.  trials[max_trials] (may be NULL
 * when max_trials is 0) receives every shape tried with its time, *ntrials
 * their count, *best the choice.
 *
 * Replay forms of a step pair (every form gives each run the results of
 * mhd_advance on that run alone, bit for bit):
 *   MHD_FORM_RUNS         each run's own captured graph on its own stream;
 *   MHD_FORM_BATCH_GRAPH  one graph holding every run's pair, a stream each;
 *   MHD_FORM_KERNELS      batched kernels: each sweep of the pair is one
 *                         launch over the blocks of the runs of a wave (a
 *                         merged block table and a table of the runs'
 *                         arguments), a wave's launches captured into one
 *                         graph, the waves replayed concurrently on their own
 *                         streams.  A wave holds runs with the same launch
 *                         sequence up to wave_cells cells.  Runs with
 *                         viscosity (fmu0, frho0, fpr0) have no batched
 *                         form: mhd_ensemble_advance returns MHD_ERR_ARG. */
#define MHD_FORM_RUNS 0
#define MHD_FORM_BATCH_GRAPH 1
#define MHD_FORM_KERNELS 2
typedef struct mhd_tile_shape {
    int kc;           /* z planes per chunk of the tile kernels (curl, RHS)    */
    int zkc;          /* interior z planes per chunk of the smoothing kernels  */
    int face_rows;    /* rows of a face window's block (12 or 16)              */
    int form;         /* replay form, MHD_FORM_*                               */
    int rhs_rows;     /* right part 2: 0 the tile kernel, 1 windows of 16     */
                      /* rows, 2 windows of 8 rows, -1 each run's own choice  */
    int wave_cells;   /* MHD_FORM_KERNELS: most physical cells of the runs    */
                      /* one wave of batched launches covers (0: no limit)    */
    double ms;        /* the trial's time                                      */
} mhd_tile_shape;
int mhd_ensemble_tune_shapes(mhd_ensemble *e, int trial, const double *dt0, const double *time, int max_trials,
                             mhd_tile_shape *trials, int *ntrials, mhd_tile_shape *best);
/* put a shape in force for every run (the ms field is ignored); results do
 * not depend on it */
int mhd_ensemble_set_shape(mhd_ensemble *e, const mhd_tile_shape *shape);

/* Discrete adjoint of the step (no counterpart in the reference, which has
 * no derivatives; the reverse mode of stepon/right, t3dsdsc.f:1063-1673).
 * From the current state x_0 (with its residual already captured), runs
 * nsteps steps with the frozen dt sequence dts[nsteps] (setdt is not
 * differentiated), evaluates the objective at x_N and returns
 *   *value    objective 0: magnetic energy 1/2 sum |B|^2 dV (trapezoid
 *             weights over unique physical points); objective 1:
 *             reconnected flux 1/2 sum_x |Bz~| dx averaged over y, Bz~ the
 *             mean of the two z planes nearest z = 0;
 *   grad_x    d value / d x_0, (mx+4, my, mz+4, 8), image rows 0 (the
 *             sensitivity of a periodic row is carried by its unique copy);
 *   *grad_eta d value / d eta0;
 *   ms[2]     HIP-event times of the forward pass and of the reverse sweep
 *             (recomputation included).
 * Reverse sweep: states are kept every `every` steps (0: ceil(sqrt(nsteps)));
 * each segment is replayed recording stage states and currents, then
 * traversed backwards.  every < 0: the forward pass records the stage states
 * and currents of every step (44 fields per step of device memory) and
 * nothing is recomputed.  flags & MHD_ADJ_RESIDUAL: x_0's residual capture
 * (mhd_init_residual) is differentiated too; otherwise xresd is a constant.
 * Any of value, grad_x,
 * grad_eta, ms may be NULL.  On return the state is x_N. */
enum { MHD_ADJ_RESIDUAL = 1 };
/* stage-level access for parity checks of the adjoint: the vector-Jacobian
 * product of one right(t) at the current state (smooth = t != 0; init: the
 * residual capture, eta = 0, no final smoothing): given xdif_bar (8 fields)
 * returns x_bar (8 fields), xresd_bar (3 fields, may be NULL) and *eta_bar
 * (may be NULL), host layout, image rows 0. */
int mhd_adjoint_rhs(mhd_handle *h, const double *xdif_bar, double *x_bar, double *xresd_bar, double *eta_bar,
                    int init, int smooth);
int mhd_adjoint_run(mhd_handle *h, int nsteps, const double *dts, int objective, int every, int flags,
                    double *value, double *grad_x, double *grad_eta, double *ms);

#ifdef __cplusplus
}
#endif
#endif
