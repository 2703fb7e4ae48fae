/*
 * mrlbm_oracle.c -- CPU restatement of the reference time step of sakthi-lbm/mrlbm_2d.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never the product.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Every expression keeps the reference's operand order, because Fortran evaluates
 * left to right and the FP64 path is compared to it at 1e-12.  Build with
 * -ffp-contract=off and without -march=native (no FMA), as the reference is built.
 * Citations: "M" = 2d_mrlbm.f90, "S" = 2d_statistics.f90 (file:line).
 *
 * Parity pins (tests/test_oracle_*.py): the restatement is compared with fixtures
 * written by the reference itself (built with flang from its own sources by
 * oracle/Makefile into oracle/_ref/) for every BC flag combination the reference
 * supports, plus closed forms (rest-state growth c^n, Hermite identities, push/pull
 * streaming on a tiny grid, Cramer's rule vs numpy, quadratic fit on exact
 * quadratics, Nadaraya-Watson brute force, peak finding on a synthetic sinusoid).
 */
#include "mrlbm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define Q 9
static const int EX[Q] = {0, 1, 0, -1, 0, 1, -1, -1, 1}; /* M:130 */
static const int EY[Q] = {0, 0, 1, 0, -1, 1, 1, -1, -1}; /* M:131 */

struct orc_state {
    orc_params p;
    int nx, ny, N;
    int nfront, nback, ntop, nbot;
    double w[Q], Hxx[Q], Hyy[Q], Hxy[Q];
    double cs2, nu, omega, uinlet, vinlet;
    double x_c, y_c, r_c, r_cyl, delx;
    /* fields */
    double *f, *fout, *ftmp;
    double *rho, *pres, *ux, *uy, *mxx, *myy, *mxy, *mxxp, *myyp, *mxx_s, *myy_s, *mxy_s;
    double *cth, *sth, *s2th, *c2th;
    int *isfluid, *issolid, *isbound;
    /* boundary list */
    int nbct;
    int *bou_i, *bou_j, *label; /* 0-based i,j */
    int *ip2, *jp2, *ip3, *jp3, *ip4, *jp4; /* [4*k + c], 0-based */
    double *xref1, *yref1, *xref2, *yref2, *xref3, *yref3, *xref4, *yref4, *delta;
    double *theta_c, *thetas;
    int *idx; /* 0-based sorted order */
    double *cxp, *cyp, *Hxxp, *Hyyp, *Hxyp; /* [9*k + q] rotated tables at boundary nodes */
    double *uxp_b, *uyp_b, *Fd, *Fl, *p_cy, *ps_cy, *p_fit;
    double *uin; /* inlet velocity per row (uniform uo, or the parabolic extension) */
    double *chi; /* Brinkman solid fraction (extension), or NULL */
    int Incs[16][Q], Outs[16][Q];
    /* recorded outputs */
    double F_drag_net, F_lift_net, rho_inf_mean;
};

#define IDX(i, j) ((i) + s->nx * (j))
#define FK(a, n, k) ((a)[(size_t)(n) + (size_t)s->N * (size_t)(k)])

static double *dalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }
static int *ialloc(size_t n) { return (int *)calloc(n ? n : 1, sizeof(int)); }

/* Incoming / outgoing population masks per quadrant label (M:2005-2057). */
static void set_masks(orc_state *s)
{
    static const int inc[16][Q] = {
        [1] = {1, 1, 0, 0, 1, 0, 0, 0, 1},  [2] = {1, 0, 0, 1, 1, 0, 0, 1, 0},
        [3] = {1, 1, 0, 1, 1, 0, 0, 1, 1},  [4] = {1, 1, 1, 0, 0, 1, 0, 0, 0},
        [5] = {1, 1, 1, 0, 1, 1, 0, 0, 1},  [7] = {1, 1, 1, 1, 1, 1, 0, 1, 1},
        [8] = {1, 0, 1, 1, 0, 0, 1, 0, 0},  [10] = {1, 0, 1, 1, 1, 0, 1, 1, 0},
        [11] = {1, 1, 1, 1, 1, 0, 1, 1, 1}, [12] = {1, 1, 1, 1, 0, 1, 1, 0, 0},
        [13] = {1, 1, 1, 1, 1, 1, 1, 0, 1}, [14] = {1, 1, 1, 1, 1, 1, 1, 1, 0},
    };
    static const int out[16][Q] = {
        [1] = {1, 0, 1, 1, 0, 0, 1, 0, 0},  [2] = {1, 1, 1, 0, 0, 1, 0, 0, 0},
        [3] = {1, 1, 1, 1, 0, 1, 1, 0, 0},  [4] = {1, 0, 0, 1, 1, 0, 0, 1, 0},
        [5] = {1, 0, 1, 1, 1, 0, 1, 1, 0},  [7] = {1, 1, 1, 1, 1, 1, 1, 1, 0},
        [8] = {1, 1, 0, 0, 1, 0, 0, 0, 1},  [10] = {1, 1, 1, 0, 1, 1, 0, 0, 1},
        [11] = {1, 1, 1, 1, 1, 1, 1, 0, 1}, [12] = {1, 1, 0, 1, 1, 0, 0, 1, 1},
        [13] = {1, 1, 1, 1, 1, 0, 1, 1, 1}, [14] = {1, 1, 1, 1, 1, 1, 0, 1, 1},
    };
    memcpy(s->Incs, inc, sizeof inc);
    memcpy(s->Outs, out, sizeof out);
}

/* get_coord (M:931-1157): node classes, boundary list, labels, r_cyl, reference points. */
static void get_coord(orc_state *s)
{
    const int nx = s->nx, ny = s->ny;
    const double epsilon = (double)0.5f;
    if (s->p.geometry_explicit) { /* extension: the centre and radius as given */
        s->x_c = s->p.x_c;
        s->y_c = s->p.y_c;
        s->r_c = s->p.r_c;
    } else {
        s->x_c = (double)(float)(s->nfront + (s->p.ncy / 2) + 1); /* M:946 */
        s->y_c = (double)(float)(s->nbot + s->p.ncy / 2);         /* M:947 */
        s->r_c = (double)(float)(s->p.ncy / 2);                   /* M:948 */
    }
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) { /* M:958-966 */
            const int n = IDX(i, j);
            const double x = (double)(i + 1), y = (double)(j + 1);
            const double radi = sqrt(((x - s->x_c) * (x - s->x_c)) + ((y - s->y_c) * (y - s->y_c)));
            s->cth[n] = (x - s->x_c) / radi;
            s->sth[n] = (y - s->y_c) / radi;
            s->s2th[n] = (2.0 * s->sth[n]) * s->cth[n];
            s->c2th[n] = (s->cth[n] * s->cth[n]) - (s->sth[n] * s->sth[n]);
        }
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) { /* M:969-980 */
            const int n = IDX(i, j);
            const double x = (double)(i + 1), y = (double)(j + 1);
            const double d = sqrt((x - s->x_c) * (x - s->x_c) + (y - s->y_c) * (y - s->y_c));
            if (fabs(d - s->r_c) <= epsilon)
                s->isbound[n] = 1;
            else if (fabs(d) <= s->r_c)
                s->issolid[n] = 1;
            else
                s->isfluid[n] = 1;
        }
    /* promote non-fluid 8-neighbours of fluid nodes (M:982-1002); 0-based i in [1,nx-2] */
    static const int NB[Q][2] = {{0, 0}, {-1, 0}, {0, -1}, {1, 0}, {0, 1}, {-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
    for (int i = 1; i < nx - 1; i++)
        for (int j = 1; j < ny - 1; j++) {
            if (s->isfluid[IDX(i, j)] != 1) continue;
            for (int l = 0; l < Q; l++) {
                const int m = IDX(i + NB[l][0], j + NB[l][1]);
                if (s->isfluid[m] == 0 && s->isbound[m] == 0) {
                    s->isbound[m] = 1;
                    s->issolid[m] = 0;
                }
            }
        }
    int nbct = 0;
    for (int n = 0; n < s->N; n++) nbct += s->isbound[n];
    s->nbct = nbct;
    const size_t B = (size_t)nbct;
    s->bou_i = ialloc(B); s->bou_j = ialloc(B); s->label = ialloc(B);
    s->ip2 = ialloc(4 * B); s->jp2 = ialloc(4 * B); s->ip3 = ialloc(4 * B);
    s->jp3 = ialloc(4 * B); s->ip4 = ialloc(4 * B); s->jp4 = ialloc(4 * B);
    s->xref1 = dalloc(B); s->yref1 = dalloc(B); s->xref2 = dalloc(B); s->yref2 = dalloc(B);
    s->xref3 = dalloc(B); s->yref3 = dalloc(B); s->xref4 = dalloc(B); s->yref4 = dalloc(B);
    s->delta = dalloc(B); s->theta_c = dalloc(B); s->thetas = dalloc(B); s->idx = ialloc(B);
    s->cxp = dalloc(9 * B); s->cyp = dalloc(9 * B);
    s->Hxxp = dalloc(9 * B); s->Hyyp = dalloc(9 * B); s->Hxyp = dalloc(9 * B);
    s->uxp_b = dalloc(B); s->uyp_b = dalloc(B); s->Fd = dalloc(B); s->Fl = dalloc(B);
    s->p_cy = dalloc(B); s->ps_cy = dalloc(B); s->p_fit = dalloc(B);

    int cnt = 0; /* M:1007-1033: i outer, j inner */
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            if (s->isbound[IDX(i, j)] != 1) continue;
            s->bou_i[cnt] = i;
            s->bou_j[cnt] = j;
#define FL(a, b) (s->isfluid[IDX((a), (b))])
            int f0 = 1, f1 = 1, f2 = 1, f3 = 1;
            if (FL(i - 1, j) == 0 && FL(i - 1, j + 1) == 0 && FL(i, j + 1) == 0) f0 = 0;
            if (FL(i + 1, j) == 0 && FL(i + 1, j + 1) == 0 && FL(i, j + 1) == 0) f1 = 0;
            if (FL(i - 1, j) == 0 && FL(i - 1, j - 1) == 0 && FL(i, j - 1) == 0) f2 = 0;
            if (FL(i + 1, j) == 0 && FL(i + 1, j - 1) == 0 && FL(i, j - 1) == 0) f3 = 0;
#undef FL
            s->label[cnt] = f0 * 1 + f1 * 2 + f2 * 4 + f3 * 8;
            cnt++;
        }
    /* r_cyl = largest boundary-node radius (M:1035-1050) */
    double max_radii = 0.0;
    for (int k = 0; k < nbct; k++) {
        const double x = (double)(s->bou_i[k] + 1), y = (double)(s->bou_j[k] + 1);
        const double r = sqrt(((x - s->x_c) * (x - s->x_c)) + ((y - s->y_c) * (y - s->y_c)));
        if (k == 0 || r > max_radii) max_radii = r;
    }
    s->r_cyl = max_radii;
    if (s->p.label_fluid) {
        /* extension: nodes of labels 0/6/9/15 (no masks in the reference, M:2008) leave the
           list and are fluid; r_cyl and the other nodes' labels stay as computed above */
        int kept = 0;
        for (int k = 0; k < nbct; k++) {
            const int l = s->label[k];
            if (l == 0 || l == 6 || l == 9 || l == 15) {
                const int n = IDX(s->bou_i[k], s->bou_j[k]);
                s->isbound[n] = 0;
                s->issolid[n] = 0;
                s->isfluid[n] = 1;
                continue;
            }
            s->bou_i[kept] = s->bou_i[k];
            s->bou_j[kept] = s->bou_j[k];
            s->label[kept] = l;
            kept++;
        }
        nbct = kept;
        s->nbct = kept;
    }

    const double delx = sqrt(2.0); /* M:1055 */
    s->delx = delx;
    const double PI = 4.0 * atan(1.0);
    const double tiny = (double)1.0e-12f;
    for (int k = 0; k < nbct; k++) { /* M:1056-1134 */
        const double x = (double)(s->bou_i[k] + 1), y = (double)(s->bou_j[k] + 1);
        const double rr = sqrt((x - s->x_c) * (x - s->x_c) + (y - s->y_c) * (y - s->y_c));
        double unx, uny, xb, yb;
        if (rr > tiny) {
            unx = (x - s->x_c) / rr;
            uny = (y - s->y_c) / rr;
            xb = s->x_c + (s->r_cyl * unx);
            yb = s->y_c + (s->r_cyl * uny);
        } else {
            unx = 0.0; uny = 0.0; xb = s->x_c; yb = s->y_c;
        }
        s->xref1[k] = xb; s->yref1[k] = yb;
        s->xref2[k] = xb + (delx * unx); s->yref2[k] = yb + (delx * uny);
        s->xref3[k] = xb + ((2.0 * delx) * unx); s->yref3[k] = yb + ((2.0 * delx) * uny);
        s->xref4[k] = xb + ((3.0 * delx) * unx); s->yref4[k] = yb + ((3.0 * delx) * uny);
        const double dx = x - xb, dy = y - yb;
        s->delta[k] = sqrt(dx * dx + dy * dy);
        const int fx2 = (int)floor(s->xref2[k]), fy2 = (int)floor(s->yref2[k]);
        const int fx3 = (int)floor(s->xref3[k]), fy3 = (int)floor(s->yref3[k]);
        const int fx4 = (int)floor(s->xref4[k]), fy4 = (int)floor(s->yref4[k]);
        /* corners 1..4 = (fx,fy) (fx+1,fy) (fx,fy+1) (fx+1,fy+1); stored 0-based */
        const int ci[4] = {0, 1, 0, 1}, cj[4] = {0, 0, 1, 1};
        for (int c = 0; c < 4; c++) {
            s->ip2[4 * k + c] = fx2 + ci[c] - 1; s->jp2[4 * k + c] = fy2 + cj[c] - 1;
            s->ip3[4 * k + c] = fx3 + ci[c] - 1; s->jp3[4 * k + c] = fy3 + cj[c] - 1;
            s->ip4[4 * k + c] = fx4 + ci[c] - 1; s->jp4[4 * k + c] = fy4 + cj[c] - 1;
        }
        double th = atan2(yb - s->y_c, xb - s->x_c);
        if (th < 0) th = th + 2.0 * PI;
        s->theta_c[k] = th;
    }
    for (int i = 0; i < nbct; i++) s->idx[i] = i; /* exchange sort (M:1137-1152) */
    for (int i = 0; i < nbct - 1; i++)
        for (int j = i + 1; j < nbct; j++)
            if (s->theta_c[s->idx[i]] > s->theta_c[s->idx[j]]) {
                const int t = s->idx[i];
                s->idx[i] = s->idx[j];
                s->idx[j] = t;
            }
    for (int i = 0; i < nbct; i++) s->thetas[i] = s->theta_c[s->idx[i]];
}

orc_state *orc_create(const orc_params *p)
{
    orc_state *s = (orc_state *)calloc(1, sizeof(orc_state));
    s->p = *p;
    if (p->channel_with_cylinder && !p->geometry_explicit) { /* M:86-92 */
        s->nfront = p->L_front * p->ncy;
        s->nback = p->L_back * p->ncy;
        s->ntop = p->L_top * p->ncy;
        s->nbot = p->L_bot * p->ncy;
        s->nx = s->nfront + p->ncy + s->nback;
        s->ny = s->ntop + p->ncy + s->nbot;
    } else {
        s->nx = p->nx;
        s->ny = p->ny;
    }
    s->N = s->nx * s->ny;
    const size_t N = (size_t)s->N;
    s->f = dalloc(9 * N); s->fout = dalloc(9 * N); s->ftmp = dalloc(9 * N);
    s->rho = dalloc(N); s->pres = dalloc(N); s->ux = dalloc(N); s->uy = dalloc(N);
    s->mxx = dalloc(N); s->myy = dalloc(N); s->mxy = dalloc(N);
    s->mxxp = dalloc(N); s->myyp = dalloc(N);
    s->mxx_s = dalloc(N); s->myy_s = dalloc(N); s->mxy_s = dalloc(N);
    s->cth = dalloc(N); s->sth = dalloc(N); s->s2th = dalloc(N); s->c2th = dalloc(N);
    s->isfluid = ialloc(N); s->issolid = ialloc(N); s->isbound = ialloc(N);
    if (p->channel_with_cylinder) get_coord(s);

    s->uinlet = p->uo; /* M:123-125 */
    s->vinlet = 0.0;
    s->uin = dalloc((size_t)s->ny);
    for (int j = 0; j < s->ny; j++) {
        if (p->inlet_parabolic) { /* extension: u_max = uo at mid-channel, 0 at both walls */
            const double sj = (double)j / (double)(s->ny - 1);
            s->uin[j] = ((4.0 * p->uo) * sj) * (1.0 - sj);
        } else {
            s->uin[j] = s->uinlet; /* M:258 */
        }
    }
    /* Re is defined on the inlet's mean velocity: uo, or 2/3 uo for the parabola */
    const double uref = p->inlet_parabolic ? (2.0 / 3.0) * p->uo : s->uinlet;
    const float wf[Q] = {16.0f / 36.0f, 4.0f / 36.0f, 4.0f / 36.0f, 4.0f / 36.0f, 4.0f / 36.0f,
                         1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f}; /* M:129, REAL(4) */
    for (int k = 0; k < Q; k++) s->w[k] = (double)wf[k];
    s->nu = uref * (double)(s->ny - 1) / p->Re; /* M:133: (uinlet*(ny-1))/Re */
    if (p->channel_with_cylinder) s->nu = uref * (2.0 * s->r_cyl) / p->Re; /* M:134 */
    s->omega = 1.0 / ((3.0 * s->nu) + (1.0 / 2.0)); /* M:135 (3.0 is REAL(4), exact) */
    s->cs2 = 1.0 / 3.0;
    for (int k = 0; k < Q; k++) { /* M:171-178 */
        s->Hxx[k] = (double)(EX[k] * EX[k]) - s->cs2;
        s->Hyy[k] = (double)(EY[k] * EY[k]) - s->cs2;
        s->Hxy[k] = (double)(EX[k] * EY[k]);
    }
    for (int b = 0; b < s->nbct; b++) { /* rotated tables at boundary nodes (M:183-197) */
        const int n = IDX(s->bou_i[b], s->bou_j[b]);
        const double c = s->cth[n], sn = s->sth[n], s2 = s->s2th[n], c2 = s->c2th[n];
        for (int k = 0; k < Q; k++) {
            s->cxp[9 * b + k] = (double)EX[k] * c + (double)EY[k] * sn;
            s->cyp[9 * b + k] = -((double)EX[k] * sn) + (double)EY[k] * c;
            s->Hxxp[9 * b + k] = ((s->Hxx[k] * (c * c)) + (s->Hyy[k] * (sn * sn))) + (s->Hxy[k] * s2);
            s->Hyyp[9 * b + k] = ((s->Hxx[k] * (sn * sn)) + (s->Hyy[k] * (c * c))) - (s->Hxy[k] * s2);
            s->Hxyp[9 * b + k] = (((s->Hyy[k] - s->Hxx[k]) * 0.50) * s2) + (s->Hxy[k] * c2);
        }
    }
    for (int n = 0; n < s->N; n++) { /* initialize (M:1220-1235) */
        s->rho[n] = 1.0;
        s->ux[n] = 0.0;
        s->uy[n] = 0.0;
        for (int k = 0; k < Q; k++) FK(s->f, n, k) = s->w[k] * s->rho[n];
    }
    set_masks(s);
    return s;
}

void orc_destroy(orc_state *s)
{
    if (!s) return;
    void *ptrs[] = {s->f, s->fout, s->ftmp, s->rho, s->pres, s->ux, s->uy, s->mxx, s->myy, s->mxy,
                    s->mxxp, s->myyp, s->mxx_s, s->myy_s, s->mxy_s, s->cth, s->sth, s->s2th, s->c2th,
                    s->isfluid, s->issolid, s->isbound, s->bou_i, s->bou_j, s->label, s->ip2, s->jp2,
                    s->ip3, s->jp3, s->ip4, s->jp4, s->xref1, s->yref1, s->xref2, s->yref2, s->xref3,
                    s->yref3, s->xref4, s->yref4, s->delta, s->theta_c, s->thetas, s->idx, s->cxp,
                    s->cyp, s->Hxxp, s->Hyyp, s->Hxyp, s->uxp_b, s->uyp_b, s->Fd, s->Fl, s->p_cy,
                    s->ps_cy, s->p_fit, s->uin, s->chi};
    for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; i++) free(ptrs[i]);
    free(s);
}

int orc_nx(const orc_state *s) { return s->nx; }
int orc_ny(const orc_state *s) { return s->ny; }
int orc_nbct(const orc_state *s) { return s->nbct; }
double orc_r_cyl(const orc_state *s) { return s->r_cyl; }
double orc_omega(const orc_state *s) { return s->omega; }
double orc_nu(const orc_state *s) { return s->nu; }
void orc_weights(const orc_state *s, double *w9) { memcpy(w9, s->w, sizeof s->w); }

void orc_get_geometry(const orc_state *s, int *isfluid, int *issolid, int *isbound)
{
    memcpy(isfluid, s->isfluid, sizeof(int) * (size_t)s->N);
    memcpy(issolid, s->issolid, sizeof(int) * (size_t)s->N);
    memcpy(isbound, s->isbound, sizeof(int) * (size_t)s->N);
}

void orc_get_boundary(const orc_state *s, int *bou_i, int *bou_j, int *label, double *delta_uk,
                      int *ip2, int *jp2, int *ip3, int *jp3, int *ip4, int *jp4,
                      double *theta_c, int *idx)
{
    for (int k = 0; k < s->nbct; k++) {
        bou_i[k] = s->bou_i[k] + 1;
        bou_j[k] = s->bou_j[k] + 1;
        label[k] = s->label[k];
        delta_uk[k] = s->delta[k];
        theta_c[k] = s->theta_c[k];
        idx[k] = s->idx[k] + 1;
        for (int c = 0; c < 4; c++) {
            ip2[4 * k + c] = s->ip2[4 * k + c] + 1; jp2[4 * k + c] = s->jp2[4 * k + c] + 1;
            ip3[4 * k + c] = s->ip3[4 * k + c] + 1; jp3[4 * k + c] = s->jp3[4 * k + c] + 1;
            ip4[4 * k + c] = s->ip4[4 * k + c] + 1; jp4[4 * k + c] = s->jp4[4 * k + c] + 1;
        }
    }
}

void orc_set_state(orc_state *s, const double *f, const double *ux, const double *uy, const double *rho)
{
    memcpy(s->f, f, sizeof(double) * 9 * (size_t)s->N);
    if (ux) memcpy(s->ux, ux, sizeof(double) * (size_t)s->N);
    if (uy) memcpy(s->uy, uy, sizeof(double) * (size_t)s->N);
    if (rho) memcpy(s->rho, rho, sizeof(double) * (size_t)s->N);
}

void orc_get_f(const orc_state *s, double *f) { memcpy(f, s->f, sizeof(double) * 9 * (size_t)s->N); }
void orc_get_fout(const orc_state *s, double *f) { memcpy(f, s->fout, sizeof(double) * 9 * (size_t)s->N); }

void orc_get_moments(const orc_state *s, double *rho, double *ux, double *uy,
                     double *mxx, double *myy, double *mxy)
{
    const size_t b = sizeof(double) * (size_t)s->N;
    if (rho) memcpy(rho, s->rho, b);
    if (ux) memcpy(ux, s->ux, b);
    if (uy) memcpy(uy, s->uy, b);
    const int bulk_s = s->p.incomp_bulk;
    if (mxx) memcpy(mxx, bulk_s ? s->mxx_s : s->mxx, b);
    if (myy) memcpy(myy, bulk_s ? s->myy_s : s->myy, b);
    if (mxy) memcpy(mxy, bulk_s ? s->mxy_s : s->mxy, b);
}

/* Bilinear value of field v at a reference point with corners c[0..3] (M:783-806). */
static double bilinear(const orc_state *s, const double *v, const int *ip, const int *jp,
                       double xr, double yr)
{
    const double xd = (xr - (double)(ip[0] + 1)) / 1.0;
    const double yd = (yr - (double)(jp[0] + 1)) / 1.0;
    const double v0 = (1.0 - xd) * v[IDX(ip[0], jp[0])] + xd * v[IDX(ip[1], jp[1])];
    const double v1 = (1.0 - xd) * v[IDX(ip[2], jp[2])] + xd * v[IDX(ip[3], jp[3])];
    return (1.0 - yd) * v0 + yd * v1;
}

static void interp_terms(double dlt, double delx, double *t1, double *t2, double *t3)
{ /* M:809-811 */
    *t1 = ((((2.0 * delx) * delx) - (dlt * dlt)) + ((3.0 * dlt) * delx)) / (2.0 * (delx * delx));
    *t2 = (dlt * (dlt - 2.0 * delx)) / (delx * delx);
    *t3 = (dlt * (dlt - delx)) / (2.0 * (delx * delx));
}

/* velocity_interpolation_boundary_nodes (M:768-819): in place, in list order. */
static void velocity_interpolation(orc_state *s)
{
    for (int k = 0; k < s->nbct; k++) {
        const int n = IDX(s->bou_i[k], s->bou_j[k]);
        const double ux2 = bilinear(s, s->ux, &s->ip2[4 * k], &s->jp2[4 * k], s->xref2[k], s->yref2[k]);
        const double uy2 = bilinear(s, s->uy, &s->ip2[4 * k], &s->jp2[4 * k], s->xref2[k], s->yref2[k]);
        const double ux3 = bilinear(s, s->ux, &s->ip3[4 * k], &s->jp3[4 * k], s->xref3[k], s->yref3[k]);
        const double uy3 = bilinear(s, s->uy, &s->ip3[4 * k], &s->jp3[4 * k], s->xref3[k], s->yref3[k]);
        double t1, t2, t3;
        interp_terms(s->delta[k], s->delx, &t1, &t2, &t3);
        s->ux[n] = (t1 * 0.0 + t2 * ux2) - t3 * ux3;
        s->uy[n] = (t1 * 0.0 + t2 * uy2) - t3 * uy3;
    }
}

/* Rotated Mxx' and Myy' at one stencil node (M:840-869). */
static void rotated_m(const orc_state *s, int n, double *mxxp, double *myyp)
{
    const double c = s->cth[n], sn = s->sth[n], s2 = s->s2th[n];
    *mxxp = ((s->mxx[n] * (c * c)) + (s->myy[n] * (sn * sn))) + (s->mxy[n] * s2);
    *myyp = ((s->mxx[n] * (sn * sn)) + (s->myy[n] * (c * c))) - (s->mxy[n] * s2);
}

static void rotated_bilinear(const orc_state *s, const int *ip, const int *jp, double xr, double yr,
                             double *mx, double *my)
{
    const double xd = (xr - (double)(ip[0] + 1)) / 1.0;
    const double yd = (yr - (double)(jp[0] + 1)) / 1.0;
    double a[4], b[4];
    for (int c = 0; c < 4; c++) rotated_m(s, IDX(ip[c], jp[c]), &a[c], &b[c]);
    const double mxx0 = (1.0 - xd) * a[0] + xd * a[1];
    const double myy0 = (1.0 - xd) * b[0] + xd * b[1];
    const double mxx1 = (1.0 - xd) * a[2] + xd * a[3];
    const double myy1 = (1.0 - xd) * b[2] + xd * b[3];
    *mx = (1.0 - yd) * mxx0 + yd * mxx1;
    *my = (1.0 - yd) * myy0 + yd * myy1;
}

/* moments_interpolation_boundary_nodes (M:821-928). */
static void moments_interpolation(orc_state *s)
{
    for (int k = 0; k < s->nbct; k++) {
        const int n = IDX(s->bou_i[k], s->bou_j[k]);
        double mx2, my2, mx3, my3, t1, t2, t3;
        rotated_bilinear(s, &s->ip2[4 * k], &s->jp2[4 * k], s->xref2[k], s->yref2[k], &mx2, &my2);
        rotated_bilinear(s, &s->ip3[4 * k], &s->jp3[4 * k], s->xref3[k], s->yref3[k], &mx3, &my3);
        interp_terms(s->delta[k], s->delx, &t1, &t2, &t3);
        s->mxxp[n] = (t1 * 0.0 + t2 * mx2) - t3 * mx3;
        s->myyp[n] = (t1 * 0.0 + t2 * my2) - t3 * my3;
    }
}

/* Rotated moment-based wall BC (M:1635-1765, variants M:1767-1871 weak, M:1873-1972 rhoeq).
 * mode 0 = strong mass, 1 = equilibrium mass, 2 = weak mass. */
static void boundary_rotation(orc_state *s, int b, double uxp, double uyp, int mode)
{
    const int n = IDX(s->bou_i[b], s->bou_j[b]);
    const int *Is = s->Incs[s->label[b]], *Os = s->Outs[s->label[b]];
    const double *cxp = &s->cxp[9 * b], *cyp = &s->cyp[9 * b];
    const double *Hxx = &s->Hxxp[9 * b], *Hyy = &s->Hyyp[9 * b], *Hxy = &s->Hxyp[9 * b];
    const double om = s->omega;
    double A[Q], E[Q], B11[Q], B22[Q], B12[Q];
    for (int k = 0; k < Q; k++) {
        A[k] = s->w[k] * ((1.0 + (3.0 * uxp) * cxp[k]) + (3.0 * uyp) * cyp[k]);
        E[k] = s->w[k] * (((((1.0 + (3.0 * uxp) * cxp[k]) + (3.0 * uyp) * cyp[k])
                            + (4.50 * (uxp * uxp)) * Hxx[k]) + (4.50 * (uyp * uyp)) * Hyy[k])
                          + (9.0 * (uxp * uyp)) * Hxy[k]);
        B11[k] = (4.50 * s->w[k]) * Hxx[k];
        B22[k] = (4.50 * s->w[k]) * Hyy[k];
        B12[k] = (4.50 * s->w[k]) * Hxy[k];
    }
    double rhoI = 0, mxxI = 0, myyI = 0, mxyI = 0, Ap = 0, Ep = 0, B11p = 0, B22p = 0, B12p = 0;
    double Dxx = 0, Dyy = 0, Dxy = 0;
    double F11xx = 0, F22xx = 0, F12xx = 0, F11yy = 0, F22yy = 0, F12yy = 0, F11xy = 0, F22xy = 0, F12xy = 0;
    for (int k = 0; k < Q; k++) {
        if (mode == 0 && Os[k] == 1) {
            Ap = Ap + A[k];
            Ep = Ep + E[k];
            B11p = B11p + B11[k];
            B22p = B22p + B22[k];
            B12p = B12p + B12[k];
        }
        if (Is[k] == 1) {
            const double fk = FK(s->f, n, k);
            rhoI = rhoI + fk;
            mxxI = mxxI + fk * Hxx[k];
            myyI = myyI + fk * Hyy[k];
            mxyI = mxyI + fk * Hxy[k];
            if (mode == 2) { /* weak: A', B' over incoming (M:1808-1814) */
                Ap = Ap + A[k];
                B11p = B11p + B11[k];
                B22p = B22p + B22[k];
                B12p = B12p + B12[k];
            }
            if (mode == 1) Ep = Ep + E[k]; /* rhoeq: E' over incoming (M:1915) */
            Dxx = Dxx + A[k] * Hxx[k];
            Dyy = Dyy + A[k] * Hyy[k];
            Dxy = Dxy + A[k] * Hxy[k];
            F11xx = F11xx + B11[k] * Hxx[k];
            F22xx = F22xx + B22[k] * Hxx[k];
            F12xx = F12xx + B12[k] * Hxx[k];
            F11yy = F11yy + B11[k] * Hyy[k];
            F22yy = F22yy + B22[k] * Hyy[k];
            F12yy = F12yy + B12[k] * Hyy[k];
            F11xy = F11xy + B11[k] * Hxy[k];
            F22xy = F22xy + B22[k] * Hxy[k];
            F12xy = F12xy + B12[k] * Hxy[k];
        }
    }
    (void)F11xx; (void)F22xx; (void)F12xx; (void)F11yy; (void)F22yy; (void)F12yy;
    const double inv = 1.0 / rhoI;
    mxxI = mxxI * inv;
    myyI = myyI * inv;
    mxyI = mxyI * inv;
    double L11xy, L22xy, L12xy, Rxy, rho_b;
    const double Mxx = s->mxxp[n], Myy = s->myyp[n];
    double Mxy;
    if (mode == 0) {
        const double G = (1.0 - om) * Ap + om * Ep;
        const double J11 = (1.0 - om) * B11p, J22 = (1.0 - om) * B22p, J12 = (1.0 - om) * B12p;
        L11xy = J11 * mxyI - F11xy;
        L22xy = J22 * mxyI - F22xy;
        L12xy = 2.0 * (J12 * mxyI - F12xy);
        Rxy = Dxy - G * mxyI;
        Mxy = ((Rxy - Mxx * L11xy) - Myy * L22xy) / L12xy;
        const double den = ((1.0 - om) * Ap + (1.0 - om) * (((Mxx * B11p) + (Myy * B22p)) + ((2.0 * Mxy) * B12p))) + om * Ep;
        rho_b = rhoI / den;
    } else if (mode == 2) {
        L11xy = B11p * mxyI - F11xy;
        L22xy = B22p * mxyI - F22xy;
        L12xy = 2.0 * (B12p * mxyI - F12xy);
        Rxy = Dxy - Ap * mxyI;
        Mxy = ((Rxy - Mxx * L11xy) - Myy * L22xy) / L12xy;
        const double den = Ap + (((Mxx * B11p) + (Myy * B22p)) + ((2.0 * Mxy) * B12p));
        rho_b = rhoI / den;
    } else {
        L11xy = F11xy;
        L22xy = F22xy;
        L12xy = 2.0 * F12xy;
        Rxy = Ep * mxyI - Dxy;
        Mxy = ((Rxy - Mxx * L11xy) - Myy * L22xy) / L12xy;
        rho_b = rhoI / Ep;
    }
    const double c = s->cth[n], sn = s->sth[n], s2 = s->s2th[n], c2 = s->c2th[n];
    s->rho[n] = rho_b;
    s->mxx[n] = ((Mxx * (c * c)) + (Myy * (sn * sn))) - (Mxy * s2);
    s->myy[n] = ((Mxx * (sn * sn)) + (Myy * (c * c))) + (Mxy * s2);
    s->mxy[n] = (((Mxx - Myy) * 0.50) * s2) + (Mxy * c2);
}

void orc_solve_system(const double *A9, const double *b3, double *x3)
{ /* solve_system (M:1974-2003); A9 row-major A(1..3,1..3) */
    const double a1 = A9[0], b1 = A9[1], c1 = A9[2], d1 = b3[0];
    const double a2 = A9[3], b2 = A9[4], c2 = A9[5], d2 = b3[1];
    const double a3 = A9[6], b3_ = A9[7], c3 = A9[8], d3 = b3[2];
    const double den = (((((a3 * b2) * c1 - (a2 * b3_) * c1) - (a3 * b1) * c2) + (a1 * b3_) * c2)
                        + (a2 * b1) * c3) - (a1 * b2) * c3;
    x3[0] = -((((((-b3_ * c2) * d1 + (b2 * c3) * d1) + (b3_ * c1) * d2) - (b1 * c3) * d2)
                - (b2 * c1) * d3) + (b1 * c2) * d3) / den;
    x3[1] = -((((((a3 * c2) * d1 - (a2 * c3) * d1) - (a3 * c1) * d2) + (a1 * c3) * d2)
               + (a2 * c1) * d3) - (a1 * c2) * d3) / den;
    x3[2] = -((((((-a3 * b2) * d1 + (a2 * b3_) * d1) + (a3 * b1) * d2) - (a1 * b3_) * d2)
                - (a2 * b1) * d3) + (a1 * b2) * d3) / den;
}

/* Unrotated moment-based wall BC with a 3x3 solve (M:1496-1633). */
static void boundary_unrotated(orc_state *s, int b)
{
    const int n = IDX(s->bou_i[b], s->bou_j[b]);
    const int *Is = s->Incs[s->label[b]], *Os = s->Outs[s->label[b]];
    const double uxp = s->ux[n], uyp = s->uy[n], om = s->omega;
    double A[Q], E[Q], B11[Q], B22[Q], B12[Q];
    for (int k = 0; k < Q; k++) {
        A[k] = s->w[k] * ((1.0 + (3.0 * uxp) * (double)EX[k]) + (3.0 * uyp) * (double)EY[k]);
        E[k] = s->w[k] * (((((1.0 + (3.0 * uxp) * (double)EX[k]) + (3.0 * uyp) * (double)EY[k])
                            + (4.50 * (uxp * uxp)) * s->Hxx[k]) + (4.50 * (uyp * uyp)) * s->Hyy[k])
                          + (9.0 * (uxp * uyp)) * s->Hxy[k]);
        B11[k] = (4.50 * s->w[k]) * s->Hxx[k];
        B22[k] = (4.50 * s->w[k]) * s->Hyy[k];
        B12[k] = (4.50 * s->w[k]) * s->Hxy[k];
    }
    double rhoI = 0, mxxI = 0, myyI = 0, mxyI = 0, Ap = 0, Ep = 0, B11p = 0, B22p = 0, B12p = 0;
    double Dxx = 0, Dyy = 0, Dxy = 0;
    double F11xx = 0, F22xx = 0, F12xx = 0, F11yy = 0, F22yy = 0, F12yy = 0, F11xy = 0, F22xy = 0, F12xy = 0;
    for (int k = 0; k < Q; k++) {
        if (Os[k] == 1) {
            Ap = Ap + A[k]; Ep = Ep + E[k];
            B11p = B11p + B11[k]; B22p = B22p + B22[k]; B12p = B12p + B12[k];
        }
        if (Is[k] == 1) {
            const double fk = FK(s->f, n, k);
            rhoI = rhoI + fk;
            mxxI = mxxI + fk * s->Hxx[k];
            myyI = myyI + fk * s->Hyy[k];
            mxyI = mxyI + fk * s->Hxy[k];
            Dxx = Dxx + A[k] * s->Hxx[k]; Dyy = Dyy + A[k] * s->Hyy[k]; Dxy = Dxy + A[k] * s->Hxy[k];
            F11xx = F11xx + B11[k] * s->Hxx[k]; F22xx = F22xx + B22[k] * s->Hxx[k]; F12xx = F12xx + B12[k] * s->Hxx[k];
            F11yy = F11yy + B11[k] * s->Hyy[k]; F22yy = F22yy + B22[k] * s->Hyy[k]; F12yy = F12yy + B12[k] * s->Hyy[k];
            F11xy = F11xy + B11[k] * s->Hxy[k]; F22xy = F22xy + B22[k] * s->Hxy[k]; F12xy = F12xy + B12[k] * s->Hxy[k];
        }
    }
    const double inv = 1.0 / rhoI;
    mxxI = mxxI * inv; myyI = myyI * inv; mxyI = mxyI * inv;
    const double G = (1.0 - om) * Ap + om * Ep;
    const double J11 = (1.0 - om) * B11p, J22 = (1.0 - om) * B22p, J12 = (1.0 - om) * B12p;
    const double A9[9] = {
        J11 * mxxI - F11xx, J22 * mxxI - F22xx, 2.0 * (J12 * mxxI - F12xx),
        J11 * myyI - F11yy, J22 * myyI - F22yy, 2.0 * (J12 * myyI - F12yy),
        J11 * mxyI - F11xy, J22 * mxyI - F22xy, 2.0 * (J12 * mxyI - F12xy)};
    const double b3[3] = {Dxx - G * mxxI, Dyy - G * myyI, Dxy - G * mxyI};
    double x[3];
    orc_solve_system(A9, b3, x);
    const double den = ((1.0 - om) * Ap + (1.0 - om) * (((x[0] * B11p) + (x[1] * B22p)) + ((2.0 * x[2]) * B12p))) + om * Ep;
    s->rho[n] = rhoI / den;
    s->mxx[n] = x[0];
    s->myy[n] = x[1];
    s->mxy[n] = x[2];
}

/* Regularized edge BCs apply_bc (M:1237-1366) and apply_bc_incomp (M:1368-1494).
 * side: 0 outlet (i = nx), 1 inlet (i = 1), 2 top wall (j = ny), 3 bottom wall (j = 1). */
static void apply_edge_bc(orc_state *s, int i, int j, int side, int incomp)
{
    static const int IS[4][Q] = {{1, 1, 1, 0, 1, 1, 0, 0, 1}, {1, 0, 1, 1, 1, 0, 1, 1, 0},
                                 {1, 1, 1, 1, 0, 1, 1, 0, 0}, {1, 1, 0, 1, 1, 0, 0, 1, 1}};
    const int n = IDX(i, j);
    double rhoI = 0, mxxI = 0, myyI = 0, mxyI = 0;
    for (int k = 0; k < Q; k++)
        if (IS[side][k] == 1) {
            const double fk = FK(s->f, n, k);
            rhoI = rhoI + fk;
            mxxI = mxxI + fk * s->Hxx[k];
            myyI = myyI + fk * s->Hyy[k];
            mxyI = mxyI + fk * s->Hxy[k];
        }
    const double inv = 1.0 / rhoI;
    mxxI = mxxI * inv;
    myyI = myyI * inv;
    mxyI = mxyI * inv;
    const double om = s->omega, ux = s->ux[n], uy = s->uy[n];
    double rp;
    if (!incomp) {
        switch (side) {
        case 0:
            rp = s->rho[n];
            s->mxx[n] = ((rp + (9.0 * rhoI) * mxxI) - (3.0 * rp) * ux) / (6.0 * rp);
            s->myy[n] = ((6.0 * rhoI) * myyI) / (5.0 * rp);
            s->mxy[n] = ((6.0 * rhoI) * mxyI - rp * uy) / (3.0 * rp);
            break;
        case 1:
            s->rho[n] = ((4.0 * rhoI) + (3.0 * rhoI) * mxxI) / (3.0 - 3.0 * ux);
            rp = s->rho[n];
            s->mxx[n] = ((rp + (9.0 * rhoI) * mxxI) + (3.0 * rp) * ux) / (6.0 * rp);
            s->myy[n] = ((6.0 * rhoI) * myyI) / (5.0 * rp);
            s->mxy[n] = ((2.0 * rhoI) * mxyI) / rp;
            break;
        default:
            s->rho[n] = ((3.0 * rhoI) * (4.0 + (3.0 * (1.0 - om)) * myyI)) / (9.0 + om);
            rp = s->rho[n];
            s->mxx[n] = ((6.0 * rhoI) * mxxI) / (5.0 * rp);
            s->myy[n] = (rp + (9.0 * rhoI) * myyI) / (6.0 * rp);
            s->mxy[n] = ((2.0 * rhoI) * mxyI) / rp;
            break;
        }
    } else {
        switch (side) {
        case 0:
            rp = s->rho[n];
            s->mxx[n] = ux * ux;
            s->myy[n] = uy * uy;
            s->mxy[n] = ((6.0 * rhoI) * mxyI - rp * uy) / (3.0 * rp);
            break;
        case 1:
            s->rho[n] = (6.0 * rhoI) / ((5.0 - 3.0 * ux) - 3.0 * (ux * ux));
            rp = s->rho[n];
            s->mxx[n] = ux * ux;
            s->myy[n] = 0.0;
            s->mxy[n] = ((2.0 * rhoI) * mxyI) / rp;
            break;
        default:
            s->rho[n] = (6.0 * rhoI) / 5.0;
            s->mxx[n] = 0.0;
            s->myy[n] = 0.0;
            s->mxy[n] = (5.0 * mxyI) / 3.0;
            break;
        }
    }
}

/* evaluate_forces (M:558-591): momentum exchange at one boundary node. */
static void evaluate_forces(const orc_state *s, int b, double *Fx, double *Fy)
{
    const int n = IDX(s->bou_i[b], s->bou_j[b]);
    const int *Is = s->Incs[s->label[b]], *Os = s->Outs[s->label[b]];
    double fxi = 0, fyi = 0, fxo = 0, fyo = 0;
    for (int k = 0; k < Q; k++) {
        if (Os[k] == 1) {
            fxo = fxo + FK(s->fout, n, k) * (double)EX[k];
            fyo = fyo + FK(s->fout, n, k) * (double)EY[k];
        }
        if (Is[k] == 1) {
            fxi = fxi + FK(s->f, n, k) * (double)EX[k];
            fyi = fyi + FK(s->f, n, k) * (double)EY[k];
        }
    }
    *Fx = fxi - fxo;
    *Fy = fyi - fyo;
}

double orc_quadratic_interpolation(double x1, double y1, double p1, double x2, double y2, double p2,
                                   double x3, double y3, double p3, double xs, double ys)
{ /* M:743-766 */
    const double r1 = sqrt((x1 * x1) + (y1 * y1));
    const double r2 = sqrt((x2 * x2) + (y2 * y2));
    const double r3 = sqrt((x3 * x3) + (y3 * y3));
    const double rs = sqrt((xs * xs) + (ys * ys));
    const double den = ((r1 - r2) * (r1 - r3)) * (r2 - r3);
    const double a0 = (((((r1 * r3) * p2) * (r3 - r1)) + ((r2 * r2) * ((r3 * p1) - (r1 * p3))))
                       + (r2 * (((r1 * r1) * p3) - ((r3 * r3) * p1)))) / den;
    const double a1 = ((((r3 * r3) * (p1 - p2)) + ((r1 * r1) * (p2 - p3))) + ((r2 * r2) * (p3 - p1))) / den;
    const double a2 = (((r3 * (p2 - p1)) + (r2 * (p1 - p3))) + (r1 * (p3 - p2))) / den;
    return (a0 + (a1 * rs)) + (a2 * (rs * rs));
}

void orc_gaussian_smoothing(int n, const double *theta, const double *p, double *out)
{ /* M:722-741 */
    const double sigma = 0.1;
    for (int i = 0; i < n; i++) {
        double sw = 0.0, sv = 0.0;
        for (int j = 0; j < n; j++) {
            const double d = theta[i] - theta[j];
            const double wgt = exp((-(d * d)) / (2.0 * (sigma * sigma)));
            sw = sw + wgt;
            sv = sv + wgt * p[j];
        }
        out[i] = sv / sw;
    }
}

/* write_statistics_out (M:593-617) without the file append: extrapolated surface
 * pressure (M:620-676), theta-sorted (M:602-604) and smoothed (M:607). */
static void surface_pressure(orc_state *s)
{
    for (int k = 0; k < s->nbct; k++) {
        const double v2 = bilinear(s, s->pres, &s->ip2[4 * k], &s->jp2[4 * k], s->xref2[k], s->yref2[k]);
        const double v3 = bilinear(s, s->pres, &s->ip3[4 * k], &s->jp3[4 * k], s->xref3[k], s->yref3[k]);
        const double v4 = bilinear(s, s->pres, &s->ip4[4 * k], &s->jp4[4 * k], s->xref4[k], s->yref4[k]);
        s->p_cy[k] = orc_quadratic_interpolation(
            s->xref2[k] - s->x_c, s->yref2[k] - s->y_c, v2, s->xref3[k] - s->x_c, s->yref3[k] - s->y_c, v3,
            s->xref4[k] - s->x_c, s->yref4[k] - s->y_c, v4, s->xref1[k] - s->x_c, s->yref1[k] - s->y_c);
    }
    for (int k = 0; k < s->nbct; k++) s->ps_cy[k] = s->p_cy[s->idx[k]];
    orc_gaussian_smoothing(s->nbct, s->thetas, s->ps_cy, s->p_fit);
}

void orc_step(orc_state *s, int iter)
{
    const orc_params *P = &s->p;
    const int nx = s->nx, ny = s->ny, N = s->N;
    const double om = s->omega;
    /* moments (M:231-245); the per-node sweeps run on OpenMP threads like the reference's
       (each node's arithmetic is unchanged, so results do not depend on the thread count) */
#pragma omp parallel for schedule(static)
    for (int n = 0; n < N; n++) {
        double f[Q];
        for (int k = 0; k < Q; k++) f[k] = FK(s->f, n, k);
        double r = 0.0, sx = 0.0, sy = 0.0, sxx = 0.0, syy = 0.0, sxy = 0.0;
        for (int k = 0; k < Q; k++) {
            r = r + f[k];
            sx = sx + f[k] * (double)EX[k];
            sy = sy + f[k] * (double)EY[k];
            sxx = sxx + f[k] * s->Hxx[k];
            syy = syy + f[k] * s->Hyy[k];
            sxy = sxy + f[k] * s->Hxy[k];
        }
        s->rho[n] = r;
        const double inv = 1.0 / r;
        s->pres[n] = r * s->cs2;
        s->ux[n] = sx * inv;
        s->uy[n] = sy * inv;
        s->mxx[n] = sxx * inv;
        s->myy[n] = syy * inv;
        s->mxy[n] = sxy * inv;
    }
    /* macroscopic Dirichlet (M:250-272) */
    if (!P->x_periodic)
        for (int j = 0; j < ny; j++) {
            s->rho[IDX(nx - 1, j)] = s->rho[IDX(nx - 2, j)];
            s->ux[IDX(nx - 1, j)] = s->ux[IDX(nx - 2, j)];
            s->uy[IDX(nx - 1, j)] = s->uy[IDX(nx - 2, j)];
        }
    if (!P->x_periodic)
        for (int j = 0; j < ny; j++) {
            s->ux[IDX(0, j)] = s->uin[j];
            s->uy[IDX(0, j)] = s->vinlet;
        }
    if (!P->y_periodic)
        for (int i = 0; i < nx; i++) {
            s->ux[IDX(i, ny - 1)] = 0.0;
            s->uy[IDX(i, ny - 1)] = 0.0;
            s->ux[IDX(i, 0)] = 0.0;
            s->uy[IDX(i, 0)] = 0.0;
        }
    /* curved wall (M:280-311) */
    if (P->channel_with_cylinder && !P->rotated_coordinate) {
        velocity_interpolation(s);
        for (int b = 0; b < s->nbct; b++) boundary_unrotated(s, b);
    }
    if (P->channel_with_cylinder && P->rotated_coordinate) {
        if (P->vel_interp) velocity_interpolation(s);
        if (P->mom_interp) moments_interpolation(s);
        for (int b = 0; b < s->nbct; b++) {
            const int n = IDX(s->bou_i[b], s->bou_j[b]);
            s->uxp_b[b] = s->ux[n] * s->cth[n] + s->uy[n] * s->sth[n];
            s->uyp_b[b] = s->uy[n] * s->cth[n] - s->ux[n] * s->sth[n];
            if (P->strong_mass_bc) boundary_rotation(s, b, s->uxp_b[b], s->uyp_b[b], 0);
            if (P->eq_mass_bc) boundary_rotation(s, b, s->uxp_b[b], s->uyp_b[b], 1);
            if (P->weak_mass_bc) boundary_rotation(s, b, s->uxp_b[b], s->uyp_b[b], 2);
        }
    }
    /* edge BCs, last writer wins at corners (M:313-351) */
    if (!P->x_periodic)
        for (int j = 0; j < ny; j++) apply_edge_bc(s, nx - 1, j, 0, P->incomp_bc);
    if (!P->x_periodic)
        for (int j = 0; j < ny; j++) apply_edge_bc(s, 0, j, 1, P->incomp_bc);
    if (!P->y_periodic)
        for (int i = 0; i < nx; i++) apply_edge_bc(s, i, ny - 1, 2, P->incomp_bc);
    if (!P->y_periodic)
        for (int i = 0; i < nx; i++) apply_edge_bc(s, i, 0, 3, P->incomp_bc);
    /* Brinkman penalization (extension): off the outer ring and off the wall list */
    if (s->chi)
#pragma omp parallel for schedule(static)
        for (int j = 1; j < ny - 1; j++)
            for (int i = 1; i < nx - 1; i++) {
                const int n = IDX(i, j);
                if (s->isbound[n] == 1) continue;
                const double keep = 1.0 - s->chi[n];
                s->ux[n] = s->ux[n] * keep;
                s->uy[n] = s->uy[n] * keep;
            }
    /* post_process: inlet density mean, solid-u zeroing on plot steps (M:355-387) */
    if (P->post_process) {
        if (iter >= P->statsbegin && iter <= P->statsend) {
            const double c = (double)(float)(iter - P->statsbegin);
            double sr = 0.0;
            for (int j = 0; j < ny; j++) sr = sr + s->rho[IDX(0, j)];
            const double rinf = sr / (double)ny;
            s->rho_inf_mean = ((c * s->rho_inf_mean) + rinf) / (c + 1.0);
        }
        if (iter % P->iplot == 0 && (P->channel_with_cylinder || P->channel_with_square))
#pragma omp parallel for schedule(static)
            for (int n = 0; n < N; n++)
                if (s->issolid[n] == 1) {
                    s->ux[n] = 0.0;
                    s->uy[n] = 0.0;
                }
    }
    /* collision (M:407-415) */
#pragma omp parallel for schedule(static)
    for (int n = 0; n < N; n++) {
        const double ux = s->ux[n], uy = s->uy[n];
        s->mxx[n] = (1.0 - om) * s->mxx[n] + (om * ux) * ux;
        s->myy[n] = (1.0 - om) * s->myy[n] + (om * uy) * uy;
        s->mxy[n] = (1.0 - om) * s->mxy[n] + (om * ux) * uy;
    }
    const double *Mxx = s->mxx, *Myy = s->myy, *Mxy = s->mxy;
    if (P->incomp_bulk) { /* trace projection (M:419-428) */
#pragma omp parallel for schedule(static)
        for (int n = 0; n < N; n++) {
            const double ux = s->ux[n], uy = s->uy[n];
            const double tr = 0.50 * (((s->mxx[n] + s->myy[n]) - ux * ux) - uy * uy);
            s->mxx_s[n] = s->mxx[n] - tr;
            s->myy_s[n] = s->myy[n] - tr;
            s->mxy_s[n] = s->mxy[n];
        }
        Mxx = s->mxx_s; Myy = s->myy_s; Mxy = s->mxy_s;
    }
    /* regularized reconstruction (M:445-457) */
#pragma omp parallel for schedule(static)
    for (int n = 0; n < N; n++) {
        const double ux = s->ux[n], uy = s->uy[n], r = s->rho[n];
        for (int k = 0; k < Q; k++) {
            const double cu = (double)EX[k] * ux + (double)EY[k] * uy;
            FK(s->fout, n, k) = (s->w[k] * r) * ((((1.0 + (3.0 * cu)) + (4.50 * Mxx[n]) * s->Hxx[k])
                                                 + (4.50 * Myy[n]) * s->Hyy[k]) + (9.00 * Mxy[n]) * s->Hxy[k]);
        }
    }
    /* statistics window: surface pressure and forces (M:461-502) */
    if (P->post_process && iter >= P->statsbegin && iter <= P->statsend && P->channel_with_cylinder) {
        surface_pressure(s);
        double sd = 0.0, sl = 0.0;
        for (int b = 0; b < s->nbct; b++) {
            evaluate_forces(s, b, &s->Fd[b], &s->Fl[b]);
            sd = sd + s->Fd[b];
            sl = sl + s->Fl[b];
        }
        s->F_drag_net = sd;
        s->F_lift_net = sl;
    }
    /* streaming: copy, push, shifted periodic wraps (M:508-535) */
#pragma omp parallel for schedule(static)
    for (long long q = 0; q < 9LL * N; q++) s->ftmp[q] = s->fout[q];
#pragma omp parallel for schedule(static) /* every target is written by one source */
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < Q; k++) {
                const int ti = i + EX[k], tj = j + EY[k];
                if (ti >= 0 && ti < nx && tj >= 0 && tj < ny) FK(s->f, IDX(ti, tj), k) = FK(s->ftmp, IDX(i, j), k);
            }
    if (P->x_periodic)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < Q; k++) {
                if (EX[k] > 0) FK(s->f, IDX(0, j), k) = FK(s->ftmp, IDX(nx - 1, j), k);
                if (EX[k] < 0) FK(s->f, IDX(nx - 1, j), k) = FK(s->ftmp, IDX(0, j), k);
            }
    if (P->y_periodic)
        for (int i = 0; i < nx; i++)
            for (int k = 0; k < Q; k++) {
                if (EY[k] > 0) FK(s->f, IDX(i, 0), k) = FK(s->ftmp, IDX(i, ny - 1), k);
                if (EY[k] < 0) FK(s->f, IDX(i, ny - 1), k) = FK(s->ftmp, IDX(i, 0), k);
            }
}

void orc_set_brinkman(orc_state *s, const double *chi)
{
    if (!chi) {
        free(s->chi);
        s->chi = NULL;
        return;
    }
    if (!s->chi) s->chi = dalloc((size_t)s->N);
    memcpy(s->chi, chi, sizeof(double) * (size_t)s->N);
}

void orc_run(orc_state *s, int iter_first, int nsteps)
{
    for (int t = 0; t < nsteps; t++) orc_step(s, iter_first + t);
}

void orc_get_forces(const orc_state *s, double *fd, double *fl)
{
    *fd = s->F_drag_net;
    *fl = s->F_lift_net;
}

void orc_get_surface_p(const orc_state *s, double *ps, double *pfit)
{
    memcpy(ps, s->ps_cy, sizeof(double) * (size_t)s->nbct);
    memcpy(pfit, s->p_fit, sizeof(double) * (size_t)s->nbct);
}

double orc_rho_inf_mean(const orc_state *s) { return s->rho_inf_mean; }

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int orc_statistics(int nstep, const double *fd, const double *fl, double uo, double rho_infty,
                   double r_cy, int *cycle_start, int *period, int *n_cycle,
                   double *st, double *cd_mean, double *cl_rms)
{ /* S:28-131 and find_peaks S:183-214; indices are 1-based rows of the force series */
    const double D = 2.0 * r_cy;
    const double f_norm = ((0.50 * rho_infty) * (uo * uo)) * D;
    int peak1 = -1, peak2 = -1;
    for (int i = 2; i <= nstep - 1; i++) {
        const double a = fl[i - 1], prev = fl[i - 2], next = fl[i];
        if (a > prev && a > next) {
            if (peak1 == -1) peak1 = i;
            else if (peak2 == -1) { peak2 = i; break; }
        }
    }
    if (peak1 == -1 || peak2 == -1) return -1;
    const int cs = peak1, per = peak2 - peak1;
    const int nc = (nstep - cs) / per;
    const int ce = cs + nc * per;
    double cd_avg = 0.0, cl_avg = 0.0;
    for (int i = cs; i <= ce; i++) {
        const int mc = i - cs;
        const double cd = fd[i - 1] / f_norm, cl = fl[i - 1] / f_norm;
        cd_avg = ((double)mc * cd_avg + cd) / (double)(mc + 1);
        cl_avg = ((double)mc * cl_avg + (cl * cl)) / (double)(mc + 1);
    }
    *cycle_start = cs;
    *period = per;
    *n_cycle = nc;
    *st = D / ((double)per * uo);
    *cd_mean = cd_avg;
    *cl_rms = sqrt(cl_avg);
    return 0;
}
