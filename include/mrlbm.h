/*
 * mrlbm.h -- C-ABI of the MI355X (gfx950) D2Q9 moments-based regularized LBM time step.
 *
 * The reference (sakthi-lbm/mrlbm_2d) has no callable interface: its time step is the
 * inline body of `program lbm_2d` (2d_mrlbm.f90:229-555) talking through program-scope
 * arrays (2d_mrlbm.f90:3-73).  This header is the boundary placed at that loop body.
 * Each entry point names the reference code it replaces.  All arrays are host arrays in
 * the reference's column-major order (x fastest; field(nx,ny), f(nx,ny,0:8)); the
 * library copies synchronously and keeps no host pointers.  Every call returns 0 on
 * success or a negative MRLBM_E_* code; mrlbm_last_error() explains the last failure.
 * One host thread per handle; calls on one handle are not re-entrant.
 */
#ifndef MRLBM_H
#define MRLBM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MRLBM_ABI_VERSION 2

enum {
    MRLBM_OK = 0,
    MRLBM_E_ARG = -1,        /* bad argument or unsupported configuration */
    MRLBM_E_HIP = -2,        /* HIP runtime failure */
    MRLBM_E_GEOMETRY = -3,   /* cylinder too close to an edge, too large, or unsupported label */
    MRLBM_E_NONFINITE = -4,  /* NaN/Inf detected by mrlbm_check_finite */
    MRLBM_E_NOMEM = -5,
    MRLBM_E_STATE = -6       /* call out of order (e.g. nothing recorded) */
};

enum { MRLBM_FP64 = 0, MRLBM_FP32_DEV = 1 };                        /* precision form */
enum { MRLBM_MASS_STRONG = 0, MRLBM_MASS_EQ = 1, MRLBM_MASS_WEAK = 2 }; /* curved-wall density rule */
enum { MRLBM_INLET_UNIFORM = 0, MRLBM_INLET_PARABOLIC = 1 };
enum { MRLBM_LABEL_REFERENCE = 0, MRLBM_LABEL_FLUID = 1 };          /* policy for labels 0/6/9/15 */

/* Run configuration.  Replaces the namelist-derived globals of 2d_mrlbm.f90:83-178
 * (&DomainSize/&Cylinder/&Numbers/&Controls/&LogicalControls, read at :2222-2275). */
typedef struct mrlbm_params {
    int32_t nx, ny;               /* lattice size (2d_mrlbm.f90:86-92 derives it from ncy, L_*) */
    int32_t x_periodic, y_periodic;
    int32_t channel_with_cylinder;
    double x_c, y_c, r_c;         /* 1-based lattice coordinates of the centre, nominal radius
                                     (2d_mrlbm.f90:946-948) */
    double Re;                    /* nu = uo*2*r_cyl/Re, or uo*(ny-1)/Re without cylinder (:133-134) */
    double uo;                    /* inlet axial velocity uinlet (:123) */
    double omega;                 /* > 0 overrides the value derived from Re */
    int32_t inlet_profile;        /* MRLBM_INLET_*; uniform is the reference (:258) */
    int32_t mass_bc;              /* MRLBM_MASS_*: strong (:1635), eq (:1873), weak (:1767) */
    int32_t rotated_coordinate;   /* 0: unrotated BC with Cramer solve (:1496-1633, :1974) */
    int32_t vel_interp, mom_interp;
    int32_t incomp_bc, incomp_bulk;
    int32_t precision;            /* MRLBM_FP64 or MRLBM_FP32_DEV */
    int32_t label_policy;         /* MRLBM_LABEL_REFERENCE reproduces the reference (NaN on
                                     labels the reference leaves unset, :2008) */
    int32_t strip_rows;           /* tuning: rows per bulk strip (0 = auto) */
    int32_t kernel;               /* bulk kernel: 0 auto (FP32: 304, FP64: 8), 1 one node per thread
                                     (checking), 2 row loop over strip_rows rows, 4/8/16 unrolled strip
                                     of that many rows with 62 written columns per wave, 304/308/316
                                     aligned 64-column waves with an unrolled strip of kernel-300 rows */
    int32_t block_order;          /* bulk blocks: 0 plain, 1 XCD-aware (vertical neighbours share an L2) */
    int32_t time_blocking;        /* 0: one bulk pass per step; 2..16 (even): two steps per bulk pass
                                     over the plain interior (strips of that many rows), the edges
                                     and the cylinder band in two ordinary sub-steps; -1 (default)
                                     auto: 10 rows FP32, 12 rows FP64.  Same bits either way */
    int32_t steps_per_pass;       /* with time_blocking on: 0 (default) or 4: four steps per bulk pass
                                     over the plain interior where four consecutive steps allow it
                                     (grids of at least 32 x 32; no plot step among them), else pairs;
                                     2: pairs only.  Same bits either way */
    int32_t quad_rows;            /* four-step strip height: 0 (default) auto (FP32 20 rows, FP64 40,
                                     each while it leaves 4096 waves, then 20, then FP32 12 / FP64
                                     8), else FP32 12, 16, 20, 24, 28 or 32, FP64 8, 16, 20, 24, 32
                                     or 40.  Same bits either way */
    int32_t reserved[2];
} mrlbm_params;

/* What one call to mrlbm_step records (2d_mrlbm.f90:355-502). */
typedef struct mrlbm_step_opts {
    int32_t post_process;         /* &LogicalControls post_process */
    int32_t statsbegin, statsend; /* window for forces, surface p, inlet-rho mean (:357, :467) */
    int32_t iplot;                /* solid-node u is zeroed before collision on iter%iplot==0 (:371-384) */
    int32_t record_surface_p;     /* compute the surface-pressure feed each window step (:479) */
    int32_t stats_interval;       /* >0: accumulate mean/RMS u fields every n window steps */
    int32_t nprobe;               /* 0..4 probes (1-based i,j) recorded every window step (:471-476) */
    int32_t probe_i[4], probe_j[4];
    int32_t reserved[6];
} mrlbm_step_opts;

typedef struct mrlbm_info {
    int32_t nx, ny, nbct, precision;
    double r_cyl;                 /* largest boundary-node radius (2d_mrlbm.f90:1038-1050) */
    double nu, omega, x_c, y_c, r_c;
    int64_t bytes_per_step;       /* algorithmic HBM bytes of one bulk step (12 words/node) */
    int32_t box[4];               /* 0-based x0, y0, width, height of the cylinder box */
    int32_t n_stencil;            /* unique stencil nodes of the boundary workgroup */
    int32_t levels;               /* dependency levels of the in-place velocity interpolation */
    int32_t pair_rows;            /* strip height of two-step passes, 0 = none */
    int32_t quad_rows;            /* strip height of four-step passes, 0 = none */
} mrlbm_info;

typedef struct mrlbm_handle mrlbm_handle;

/* Setup: replaces allocate_memory/get_coord/initialize and the rotated tables
 * (2d_mrlbm.f90:94-214, 931-1235, 2005-2057, 2277-2332).  The state starts at rest
 * (rho = 1, u = 0: f = w_k, 2d_mrlbm.f90:1220-1235). */
int mrlbm_create(const mrlbm_params *p, mrlbm_handle **out);
int mrlbm_destroy(mrlbm_handle *h);
/* Last error message of h, or of the last failed mrlbm_create when h is NULL. */
const char *mrlbm_last_error(const mrlbm_handle *h);
int mrlbm_abi_version(void);
void mrlbm_default_params(mrlbm_params *p);
void mrlbm_default_step_opts(mrlbm_step_opts *o);

int mrlbm_get_info(const mrlbm_handle *h, mrlbm_info *out);
/* Boundary list in the reference's order (i outer, j inner; 2d_mrlbm.f90:1007-1033):
 * 1-based bou_i/bou_j, quadrant label, theta (radians in [0,2pi), :1132-1133) and the
 * theta-sorted permutation idx (1-based, :1137-1152).  Any pointer may be NULL. */
int mrlbm_get_boundary(const mrlbm_handle *h, int32_t *bou_i, int32_t *bou_j, int32_t *label,
                       double *theta, int32_t *idx);
/* Node classes (1 = fluid, 2 = solid, 3 = boundary) as int8, (nx,ny) column-major. */
int mrlbm_get_node_class(const mrlbm_handle *h, int8_t *cls);

/* State in: post-stream populations f(nx,ny,0:8), as read from restart.dat
 * (read_restart_file, 2d_mrlbm.f90:2204-2220).  The next step starts from exactly these
 * populations, and out-of-domain (stale) populations keep their values from f. */
int mrlbm_set_populations(mrlbm_handle *h, const double *f);
/* State in: post-collision moments; f_k = reconstruction of (rho, u, m) (2d_mrlbm.f90:445-457).
 * mxx/myy/mxy may be NULL for u_a*u_b (the second-order equilibrium). */
int mrlbm_set_moments(mrlbm_handle *h, const double *rho, const double *ux, const double *uy,
                      const double *mxx, const double *myy, const double *mxy);

/* Advance nsteps iterations numbered iter_first..iter_first+nsteps-1 (the body of the
 * main loop, 2d_mrlbm.f90:229-535).  opts may be NULL (no recording). */
int mrlbm_step(mrlbm_handle *h, int32_t iter_first, int32_t nsteps, const mrlbm_step_opts *opts);
int mrlbm_synchronize(mrlbm_handle *h);

/* Post-BC rho, ux, uy of the last step (the arrays written to plot3d and restart,
 * 2d_mrlbm.f90:2178-2180, :2198).  Any pointer may be NULL. */
int mrlbm_get_fields(const mrlbm_handle *h, double *rho, double *ux, double *uy);
/* All six post-collision moments (rho, ux, uy, mxx, myy, mxy). */
int mrlbm_get_moments(const mrlbm_handle *h, double *rho, double *ux, double *uy,
                      double *mxx, double *myy, double *mxy);
/* Post-stream populations f(nx,ny,0:8) for restart.dat (write_restart_file, :2186-2202). */
int mrlbm_get_populations(const mrlbm_handle *h, double *f);
/* The same populations, written to the same f, without allocating anything (no counterpart in
 * the reference; for callers that cannot allocate device memory: code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  mrlbm_get_populations allocates and frees
 * 72 B per node on every call; this variant stages bands of rows through a fixed 8 MiB array
 * per GPU that the library holds until the process ends.  Limit: nx <= 116508 (one row of nine
 * FP64 populations must fit the array), else MRLBM_E_ARG.  Every call on one GPU shares that
 * GPU's array: calls that overlap in time on one GPU (from several host threads) race. */
int mrlbm_get_populations_noalloc(mrlbm_handle *h, double *f);

/* Force series recorded since the last call (evaluate_forces + sums, :481-494):
 * up to max entries of (iter, Fd, Fl); *count gets the number returned. */
int mrlbm_get_forces(mrlbm_handle *h, int32_t *iters, double *fd, double *fl, int32_t max,
                     int32_t *count);
/* Surface pressure of the last recorded step, theta-sorted, raw and smoothed
 * (write_statistics_out, :593-617).  Arrays of nbct. */
int mrlbm_get_surface_p(const mrlbm_handle *h, int32_t *iter, double *ps, double *pfit);
/* Surface-pressure series recorded since the last call: iters[n], ps[n*nbct], pfit[n*nbct]. */
int mrlbm_get_surface_p_series(mrlbm_handle *h, int32_t *iters, double *ps, double *pfit, int32_t max,
                               int32_t *count);
/* Running mean of the inlet density (:357-360). */
int mrlbm_get_rho_inf_mean(const mrlbm_handle *h, double *value);
/* Probe series since the last call: per entry iter, then rho, ux, uy at each probe. */
int mrlbm_get_probes(mrlbm_handle *h, int32_t *iters, double *vals, int32_t max, int32_t *count);
/* Time-averaged fields (new in this library; the reference's iplotstats is unused, :6):
 * mean ux, uy and rms of their fluctuations over the samples taken so far. */
int mrlbm_get_stats(const mrlbm_handle *h, double *mean_ux, double *mean_uy, double *rms_ux,
                    double *rms_uy, int64_t *nsamples);
int mrlbm_reset_stats(mrlbm_handle *h);
/* max(ux) over all nodes including solid ones, as logged every 100 steps (:546); NaN values
 * are ignored (mrlbm_check_finite reports them).  Reduced on the device. */
int mrlbm_max_ux(const mrlbm_handle *h, double *value);
/* MRLBM_E_NONFINITE when any stored moment is NaN/Inf (the reference never checks).  Reduced
 * on the device: one read of the state, 16 bytes back to the host. */
int mrlbm_check_finite(const mrlbm_handle *h);

/* Brinkman solid fraction chi(nx,ny), column-major (no counterpart in the reference): before
 * the collision of every non-edge, non-wall node the velocity is scaled, u <- (1 - chi) u.
 * The physical range is [0,1); any finite value below 1 is accepted (finite-difference
 * checks step below 0).  NULL removes the field.  While a field is set the bulk runs the
 * one-node-per-thread kernel. */
int mrlbm_set_brinkman(mrlbm_handle *h, const double *chi);

/* Discrete adjoint of the step for the time-averaged drag (no counterpart in the
 * reference; objective built on evaluate_forces, 2d_mrlbm.f90:558-591, and the Cd
 * normalisation of 2d_statistics.f90:29-32). */
typedef struct mrlbm_adjoint_opts {
    int32_t iter_first;           /* iteration number of the first forward step */
    int32_t nsteps;               /* forward steps from the current state */
    int32_t window;               /* J = mean Cd over the last `window` steps */
    int32_t checkpoint_interval;  /* K: moment fields are checkpointed every K steps and the
                                     K steps in between are recomputed in the reverse sweep */
    int32_t serial_recompute;     /* 0 (default): when device memory holds a second segment
                                     buffer, segment s-1 is recomputed on a second stream while
                                     segment s is reversed; 1: one buffer, one stream.  Same
                                     result either way */
    int32_t reserved[3];
} mrlbm_adjoint_opts;

typedef struct mrlbm_adjoint_result {
    double J;                     /* mean of Fx / (0.5 * 1 * uref^2 * 2 r_cyl) over the window,
                                     uref = 2/3 uo (parabolic inlet) or uo (uniform) */
    double dJ_duo;                /* dJ/d(uo) at fixed omega (inlet profile and Cd scale) */
    double forward_ms;            /* checkpointed forward run, wall clock */
    double adjoint_ms;            /* reverse sweep (recomputation + adjoint steps), wall clock */
    double reserved[4];
} mrlbm_adjoint_result;

void mrlbm_default_adjoint_opts(mrlbm_adjoint_opts *o);
/* Runs the forward steps from the current state (the handle ends in the final state),
 * then the adjoint sweep back to the start.  dJ_dchi (nx*ny, column-major, may be NULL)
 * receives dJ/dchi; nodes on the edges and the curved wall, where chi does not act, get 0.
 * Requires a cylinder, walls (no periodic axis), incomp_bulk = 0, and a state set by
 * mrlbm_set_moments or reached by stepping (not mrlbm_set_populations).  Uses a zero
 * Brinkman field when none is set. */
int mrlbm_adjoint_drag(mrlbm_handle *h, const mrlbm_adjoint_opts *opts, double *dJ_dchi,
                       mrlbm_adjoint_result *result);

/* Batched runs (no counterpart in the reference, which runs one case per process): many
 * independent runs of uneven size advance together, one launch per kernel per step for
 * all runs, the steps captured in HIP graphs.  Every run is an ordinary handle (set and
 * read its state with the single-run calls above) that must not be stepped on its own
 * while it belongs to the batch. */
typedef struct mrlbm_batch mrlbm_batch;

typedef struct mrlbm_batch_opts {
    int32_t kernel;               /* bulk variant of every run (as mrlbm_params::kernel, not 1);
                                     0: autotune over the candidate variants */
    int32_t persistent;           /* bulk grid: 1 resident blocks looping over the tile table,
                                     0 one block per tile, -1 autotune */
    int32_t graph_steps;          /* steps per captured graph (even; 0 = 50) */
    int32_t tune_steps;           /* timed steps per autotune candidate (0 = 20) */
    int32_t time_blocking;        /* -1 autotune, 0 one step per pass, 2..10 (even) two steps per pass
                                     with strips of that many rows (as mrlbm_params) */
    int32_t steps_per_pass;       /* 0: autotune over one, two and four steps per bulk pass;
                                     2: at most two; 4: four (strip height autotuned; pairs or single
                                     steps for the rest of a graph) */
    int32_t reserved[2];
} mrlbm_batch_opts;

typedef struct mrlbm_batch_info {
    int32_t nruns, kernel, persistent, grid, ntiles, graph_steps;
    int64_t nodes;                /* sum of nx*ny over the runs */
    int32_t time_blocking;        /* chosen two-step strip height, 0 = one step per pass */
    int32_t ntune;                /* autotune candidates timed */
    int32_t tune_kernel[16], tune_persistent[16], tune_tb[16];
    double tune_ms[16];           /* mean step time of each candidate */
    int32_t quad_rows;            /* chosen four-step strip height, 0 = no four-step passes */
    int32_t tune_quad[16];        /* four-step strip height of each candidate, 0 = none */
} mrlbm_batch_info;

void mrlbm_default_batch_opts(mrlbm_batch_opts *o);
/* Creates nruns handles from runs[] (each must have a cylinder-free or cylinder geometry,
 * no Brinkman field) and the batch around them; autotunes on a saved copy of the states. */
int mrlbm_batch_create(const mrlbm_params *runs, int32_t nruns, const mrlbm_batch_opts *opts,
                       mrlbm_batch **out);
int mrlbm_batch_destroy(mrlbm_batch *b);
const char *mrlbm_batch_last_error(const mrlbm_batch *b);
/* Handle of run r (owned by the batch), or NULL. */
mrlbm_handle *mrlbm_batch_run(mrlbm_batch *b, int32_t r);
/* Re-runs the autotuner on the current states (they are restored afterwards). */
int mrlbm_batch_autotune(mrlbm_batch *b);
/* Advance every run nsteps; record_forces != 0 adds each step's Fd, Fl, Fl^2 to the runs'
 * sums.  Synchronous. */
int mrlbm_batch_step(mrlbm_batch *b, int32_t nsteps, int32_t record_forces);
/* Per-run sums since the last reset: fd[r], fl[r], fl2[r], count[r] (any may be NULL). */
int mrlbm_batch_forces(mrlbm_batch *b, double *fd_sum, double *fl_sum, double *fl2_sum, int64_t *count);
int mrlbm_batch_reset_forces(mrlbm_batch *b);
int mrlbm_batch_get_info(const mrlbm_batch *b, mrlbm_batch_info *info);

/* Timing mode: on = 0 off; on = n >= 1 takes the start and end times of every n-th bulk
 * launch of each mrlbm_step call (the first included) into a HIP event pair given to the
 * launch itself (hipExtLaunchKernel).  mrlbm_last_step_ms: mean time per step of the last
 * mrlbm_step call between HIP events on the handle's stream (0 when timing is off). */
int mrlbm_set_timing(mrlbm_handle *h, int32_t on);
int mrlbm_last_step_ms(const mrlbm_handle *h, double *ms);
/* Mean duration (ms) of one bulk-kernel launch over the last timed mrlbm_step call (a
 * two- or four-step launch covers two or four iterations), from the start/end event pairs of the timed
 * launches (at most 4096); *launches = number of launches timed.
 * New: no counterpart in the reference (its loop prints MLUPS only, 2d_mrlbm.f90:540-553). */
int mrlbm_last_bulk_ms(const mrlbm_handle *h, double *ms_per_launch, int32_t *launches);
/* Bulk launches of the last mrlbm_step call: all of them, and those timed in timing mode.
 * Four-step launches advance 4 steps, two-step launches 2, single-step launches 1; timed_steps counts the steps
 * the timed launches advanced, so timed_steps * bytes per step / timed_ms is exact whatever
 * mix of launches ran.  New: no counterpart in the reference. */
typedef struct mrlbm_timing {
    double step_ms;               /* mean time per step of the call (HIP events on the handle's stream) */
    double timed_ms;              /* summed durations of the timed bulk launches */
    double bulk_mean_ms, bulk_median_ms, bulk_min_ms, bulk_max_ms; /* per timed launch */
    int32_t bulk_launches;        /* bulk launches of the call */
    int32_t bulk_steps;           /* steps those launches advanced */
    int32_t timed_launches;
    int32_t timed_steps;
    int32_t reserved[4];
} mrlbm_timing;
int mrlbm_last_timing(const mrlbm_handle *h, mrlbm_timing *out);

#ifdef __cplusplus
}
#endif
#endif
