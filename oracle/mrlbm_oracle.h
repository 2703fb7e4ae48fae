/*
 * mrlbm_oracle.h -- CPU restatement of the reference D2Q9 moments-based regularized
 * LBM time step (sakthi-lbm/mrlbm_2d, 2d_mrlbm.f90 and 2d_statistics.f90).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * Nothing in the mrlbm_2d_amd package links, imports or calls it.
 *
 * Layout follows the reference: populations f(nx,ny,0:8) column-major, 0-based here:
 *   node n = i + nx*j (i in [0,nx), j in [0,ny)), population (n, k) at n + N*k, N = nx*ny.
 * Iteration numbers keep the reference's 1-based convention (a fresh run's first
 * step is iter = 2, 2d_mrlbm.f90:7,229).
 */
#ifndef MRLBM_ORACLE_H
#define MRLBM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* The reference's namelist groups (2d_mrlbm.f90:2227-2233), minus the unused ones. */
typedef struct orc_params {
    int nx, ny;                               /* &DomainSize (only used without cylinder) */
    int ncy, L_front, L_back, L_top, L_bot;   /* &Cylinder */
    double Re;                                /* &Numbers */
    double uo;                                /* &Controls */
    int iplot, statsbegin, statsend;
    int post_process, x_periodic, y_periodic; /* &LogicalControls */
    int channel_with_cylinder, channel_with_square;
    int incomp_bc, incomp_bulk, vel_interp, mom_interp, rotated_coordinate;
    int strong_mass_bc, eq_mass_bc, weak_mass_bc;
    /* Extensions beyond the reference (its cases keep all of these 0):
     * geometry_explicit: grid nx, ny and the cylinder's 1-based centre (x_c, y_c) and nominal
     *   radius r_c given directly, instead of derived from ncy and L_* (M:86-92, M:946-948);
     * inlet_parabolic: inlet u(j) = 4 uo s (1 - s), s = (j - 1)/(ny - 1), in place of the
     *   uniform uo (M:258), and nu taken on the mean velocity 2/3 uo (M:133-134);
     * label_fluid: boundary nodes whose quadrant label the reference leaves without masks
     *   (0, 6, 9, 15; M:2008-2057) are bulk fluid instead of boundary nodes. */
    int geometry_explicit;
    double x_c, y_c, r_c;
    int inlet_parabolic, label_fluid;
} orc_params;

typedef struct orc_state orc_state;

orc_state *orc_create(const orc_params *p);
void orc_destroy(orc_state *s);

int orc_nx(const orc_state *s);
int orc_ny(const orc_state *s);
int orc_nbct(const orc_state *s);
double orc_r_cyl(const orc_state *s);
double orc_omega(const orc_state *s);
double orc_nu(const orc_state *s);
void orc_weights(const orc_state *s, double *w9);

/* Geometry of the compacted boundary list (1-based i,j as in the reference). */
void orc_get_geometry(const orc_state *s, int *isfluid, int *issolid, int *isbound);
void orc_get_boundary(const orc_state *s, int *bou_i, int *bou_j, int *label, double *delta_uk,
                      int *ip2, int *jp2, int *ip3, int *jp3, int *ip4, int *jp4,
                      double *theta_c, int *idx);

/* State: f (N*9), ux, uy, rho (N each) -- the restart record (2d_mrlbm.f90:2198). */
void orc_set_state(orc_state *s, const double *f, const double *ux, const double *uy, const double *rho);
void orc_get_f(const orc_state *s, double *f);
void orc_get_fout(const orc_state *s, double *fout);
/* Post-BC moments of the last step: rho, ux, uy and the post-collision mxx, myy, mxy. */
void orc_get_moments(const orc_state *s, double *rho, double *ux, double *uy,
                     double *mxx, double *myy, double *mxy);

/* Brinkman solid fraction chi (N values, node i + nx*j; NULL removes it).  Extension, no
 * counterpart in the reference: after the edge BCs, every node off the domain's outer ring
 * and off the curved-wall list has its velocity scaled, u <- (1 - chi) u, before the
 * plot-step zeroing and the collision. */
void orc_set_brinkman(orc_state *s, const double *chi);

/* One iteration of the main loop body (2d_mrlbm.f90:229-535) numbered `iter`. */
void orc_step(orc_state *s, int iter);
void orc_run(orc_state *s, int iter_first, int nsteps);

/* Outputs recorded by the last step inside [statsbegin, statsend]. */
void orc_get_forces(const orc_state *s, double *fd, double *fl);
void orc_get_surface_p(const orc_state *s, double *ps, double *pfit);
double orc_rho_inf_mean(const orc_state *s);
/* OpenMP threads the per-node sweeps use (1 without OpenMP). */
int orc_num_threads(void);

/* 2d_statistics.f90 post-processor restated: peaks of the lift series, St, mean Cd and
 * rms Cl (2d_statistics.f90:90-131, 183-214).  Returns 0, or -1 when fewer than two peaks. */
int orc_statistics(int nstep, const double *fd, const double *fl, double uo, double rho_infty,
                   double r_cy, int *cycle_start, int *period, int *n_cycle,
                   double *st, double *cd_mean, double *cl_rms);

/* Nadaraya-Watson smoothing (2d_mrlbm.f90:722-741) and radial quadratic fit
 * (2d_mrlbm.f90:743-766), exported so tests can pin them on closed forms. */
void orc_gaussian_smoothing(int n, const double *theta, const double *p, double *out);
double orc_quadratic_interpolation(double x1, double y1, double p1, double x2, double y2, double p2,
                                   double x3, double y3, double p3, double xs, double ys);
void orc_solve_system(const double *A9, const double *b3, double *x3);

#ifdef __cplusplus
}
#endif
#endif
