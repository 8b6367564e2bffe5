/*
 * lbm_oracle.h -- FP64 CPU restatement of LUMA v1.7.12's lattice-Boltzmann time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker: only tests/, the smoke check in
 * __graft_entry__.py and the cpu_baseline leg of bench.py may load it.  The product
 * (luma_amd/, libluma_amd.so) never links, loads or calls it.
 *
 * It keeps LUMA's CPU data layout on purpose, so that its arithmetic order can be read
 * line-by-line against the reference:
 *   site index   id = k + j*K + i*K*M           (LUMA/inc/IVector.h:94-101)
 *   populations  f[v + Q*id], fNew[v + Q*id]    (LUMA/src/GridObj_ops_lbm_optimised.cpp:79)
 *   velocity     u[d + D*id], force_xyz[d + D*id], force_i[v + Q*id], rho[id]
 * Labels use LUMA's eType numbering (LUMA/inc/Enumerations.h:84-96).
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { O_SOLID = 0, O_FLUID, O_REFINED, O_TL2C, O_TL2F, O_BFL, O_VELOCITY, O_PRESSURE, O_SLIP,
       O_EXTRAP_RIGHT };

typedef struct ora_desc {
    int dims, Q;                 /* 2/9 or 3/19 */
    int N, M, K;                 /* N_lim, M_lim, K_lim (K = 1 in 2D) */
    int regularised;             /* L_REGULARISED_BOUNDARIES */
    int forced;                  /* L_IBM_ON || L_GRAVITY_ON: force_xyz, force_i and +F/2 active */
    int fz_compat;               /* reproduce the 3D macro quirk (:833) */
    int level;                   /* refinement level (0 = coarsest) */
    double omega;
    double rho_in;               /* L_RHOIN */
    double xmax, ymax, zmax;     /* global_edges[e?Max][0] of level 0 (for wall classification) */
    double wall_thk[6];          /* L_WALL_THICKNESS_{LEFT,RIGHT,BOTTOM,TOP,FRONT,BACK} */
    int coarse_lims[6];          /* CoarseLimsX/Y/Z[min,max] on the parent (children only) */
} ora_desc;

typedef struct ora_grid ora_grid;

ora_grid *ora_create(const ora_desc *d);
void ora_destroy(ora_grid *g);                     /* destroys children too */
#define ORA_MAX_KIDS 16                           /* refined regions per grid */
int ora_add_child(ora_grid *parent, ora_grid *child);   /* -1 past ORA_MAX_KIDS */

/* Arrays in LUMA layout; all are copied in. */
void ora_set_labels(ora_grid *g, const int *lattyp);
void ora_set_positions(ora_grid *g, const double *xpos, const double *ypos, const double *zpos);
void ora_set_inlet(ora_grid *g, const double *ux_in, const double *uy_in, const double *uz_in);
void ora_set_macro(ora_grid *g, const double *rho, const double *u);
void ora_set_force(ora_grid *g, const double *force_xyz);
void ora_set_f(ora_grid *g, const double *f);
void ora_init_feq(ora_grid *g);                    /* f = fNew = feq(rho,u) (GridObj_init_grids.cpp:315-333) */

void ora_get_f(const ora_grid *g, double *f);
void ora_get_macro(const ora_grid *g, double *rho, double *u);
void ora_get_force(const ora_grid *g, double *force_xyz);
int ora_get_t(const ora_grid *g);

/* One LBM_multi_opt(subcycle) call on g and (recursively) its children. */
void ora_multi_opt(ora_grid *g, int subcycle);
/* The two halves, for callers that put IBM work in between (no recursion into children). */
void ora_sweep1(ora_grid *g, int subcycle);        /* stream + BC + macro */
void ora_sweep2(ora_grid *g);                      /* force + collide + swap + ++t */
void ora_reset_forces(ora_grid *g);                /* _LBM_resetForces */
void ora_set_gravity(ora_grid *g, double gravity, int dir);   /* L_GRAVITY_ON in the reset */
void ora_set_ramp(ora_grid *g, double ramp_time, double dt);  /* L_VELOCITY_RAMP (0: off), GridObj::dt */
/* L_REYNOLDS_RAMP T (0: off) over L_RE with GridObj::dt, dh of the grid; omega is then set
 * at the start of every step */
void ora_set_reynolds_ramp(ora_grid *g, double re, double ramp_time, double dt, double dh);
double ora_get_omega(const ora_grid *g);
/* L_PRESSURE_DELTA (dimensionless) with GridObj::dh, dt, dm of the grid */
void ora_set_pressure_delta(ora_grid *g, double p, double dh, double dt, double dm);
void ora_macro_site(ora_grid *g, int i, int j, int k);  /* _LBM_macro_opt on one site */
/* Run n steps of LBM_multi_opt on g with an optional OpenMP sweep (nthreads > 1). */
void ora_run(ora_grid *g, int nsteps, int nthreads);

/* Per-site helpers exposed for unit tests. */
double ora_equilibrium(const ora_grid *g, int id, int v);
void ora_lattice(int Q, int *c_out /* Q*3 */, double *w_out /* Q */, int *opp_out /* Q */);
double ora_cs(void);

/* IBM hook: called between sweeps when set (used by the IBM oracle). */
typedef void (*ora_ibm_hook)(ora_grid *g, void *ctx);
void ora_set_ibm_hook(ora_grid *g, ora_ibm_hook fn, void *ctx);

/* Raw access for the IBM oracle. */
double *ora_u(ora_grid *g);
double *ora_rho(ora_grid *g);
double *ora_force(ora_grid *g);
double *ora_fnew(ora_grid *g);
const int *ora_labels(const ora_grid *g);
const ora_desc *ora_get_desc(const ora_grid *g);
const double *ora_pos(const ora_grid *g, int axis);
void ora_set_u_n_copy(ora_grid *g, int on);       /* hasFlexibleBodies: u_n = u after sweep 1 */
double *ora_u_n(ora_grid *g);

#ifdef __cplusplus
}
#endif
#endif
