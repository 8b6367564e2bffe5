/*
 * lbm_oracle.c -- FP64 CPU restatement of LUMA v1.7.12's GridObj::LBM_multi_opt.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  See lbm_oracle.h.  Built with
 * -ffp-contract=off and no fast-math so that every floating-point operation happens in the
 * order written, as in LUMA's own x86-64 -O3 build (no FMA without -march).
 *
 * Each function names the reference lines whose arithmetic it restates; the expression
 * trees (operand order, int*double promotions, the explicit 0*f terms) are kept as LUMA
 * writes them so that results can be compared bit for bit.
 */
#include "lbm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- lattice tables: LUMA/src/stdafx.cpp:81-153, GridUtils.cpp:54-67 ---------------- */
static const int C19[19][3] = {
    {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},  {0, 0, -1}, {1, 1, 0},
    {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {0, 1, 1},  {0, -1, -1}, {0, 1, -1}, {0, -1, 1},
    {1, 0, 1},  {-1, 0, -1}, {-1, 0, 1}, {1, 0, -1}, {0, 0, 0}};
static const double W19[19] = {
    1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0,
    1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
    1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 3.0};
static const int OPP19[19] = {1, 0, 3, 2, 5, 4, 7, 6, 9, 8, 11, 10, 13, 12, 15, 14, 17, 16, 18};

static const int C9[9][3] = {{1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {1, 1, 0},
                             {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {0, 0, 0}};
static const double W9[9] = {1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0, 1.0 / 36.0,
                             1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 4.0 / 9.0};
static const int OPP9[9] = {1, 0, 3, 2, 5, 4, 7, 6, 8};

#define SQ(x) ((x) * (x))

struct ora_grid {
    ora_desc d;
    int D, Q;
    const int (*c)[3];
    const double *w;
    const int *opp;
    double cs;
    size_t ncell;
    double *f, *fNew, *u, *u_n, *rho, *force_xyz, *force_i;
    int *LatTyp;
    double *ux_in, *uy_in, *uz_in;
    double *XPos, *YPos, *ZPos;
    int t;
    ora_grid *parent;
    ora_grid *kids[ORA_MAX_KIDS];   /* GridObj::subGrid, in the order they step */
    int nkids;
    ora_ibm_hook hook;
    void *hook_ctx;
    int copy_u_n;
    int gravity_on, gravity_dir;   /* L_GRAVITY_ON, L_GRAVITY_DIRECTION */
    double gravity;                /* GridObj::gravity (lattice units) */
    double ramp_time, dt;          /* L_VELOCITY_RAMP (0: undefined), GridObj::dt */
    double pressure_drho;          /* GridUnits::pd2dlbm(L_PRESSURE_DELTA) */
    double re, re_ramp, re_dt, re_dh;   /* L_RE, L_REYNOLDS_RAMP (0: undefined), GridObj::dt, dh */
};

/* GridUtils::getVelocityRampCoefficient (GridUtils.cpp:1808-1816) */
static double ramp_coefficient(const ora_grid *g, double t) {
    if (g->ramp_time > 0.0 && t <= g->ramp_time) return (1.0 - cos(3.14159265358979323846 * t / g->ramp_time)) / 2.0;
    return 1.0;
}
void ora_set_ramp(ora_grid *g, double ramp_time, double dt) {
    g->ramp_time = ramp_time;
    g->dt = dt;
}
/* L_REYNOLDS_RAMP: GridUtils::getReynoldsRampCoefficient (GridUtils.cpp:1825-1833), which
 * is 1 - cos(pi t / T) up to t = T (rising to 2) and 1 after */
void ora_set_reynolds_ramp(ora_grid *g, double re, double ramp_time, double dt, double dh) {
    g->re = re;
    g->re_ramp = ramp_time;
    g->re_dt = dt;
    g->re_dh = dh;
}
/* _LBM_updateReynolds (GridObj_ops_lbm_optimised.cpp:1313-1321) with Re = L_RE * coefficient
 * at (t + 1) dt, at the start of LBM_multi_opt (:39-42): nu = nud2nulbm(1 / Re)
 * (GridUnits.h:128-131), omega = 1 / (nu / cs^2 + 0.5) */
static void update_reynolds(ora_grid *g) {
    const double t = (g->t + 1) * g->re_dt;
    double coef = 1.0;
    if (t <= g->re_ramp) coef = 1.0 - cos(3.14159265358979323846 * t / g->re_ramp);
    const double re = g->re * coef;
    const double nu = ((1.0 / re) * g->re_dt) / (g->re_dh * g->re_dh);
    const double cs = 1.0 / sqrt(3.0);
    g->d.omega = 1.0 / ((nu / (cs * cs)) + 0.5);
}

double ora_get_omega(const ora_grid *g) { return g->d.omega; }

/* L_PRESSURE_DELTA added to the pressure sites' density (GridObj_ops_lbm_optimised.cpp:343-345)
 * through GridUnits::pd2dlbm (GridUnits.h:177-180): p dh dt^2 / dm / cs^2 */
void ora_set_pressure_delta(ora_grid *g, double p, double dh, double dt, double dm) {
    const double cs = 1.0 / sqrt(3.0);
    g->pressure_drho = (p * dh * (dt * dt) / dm) / (cs * cs);
}

double ora_cs(void) { return 1.0 / sqrt(3.0); }   /* stdafx.cpp:152 */

void ora_lattice(int Q, int *c_out, double *w_out, int *opp_out) {
    for (int v = 0; v < Q; ++v) {
        for (int a = 0; a < 3; ++a) c_out[3 * v + a] = (Q == 19) ? C19[v][a] : C9[v][a];
        w_out[v] = (Q == 19) ? W19[v] : W9[v];
        opp_out[v] = (Q == 19) ? OPP19[v] : OPP9[v];
    }
}

ora_grid *ora_create(const ora_desc *d) {
    ora_grid *g = (ora_grid *)calloc(1, sizeof(ora_grid));
    g->d = *d;
    g->D = d->dims;
    g->Q = d->Q;
    g->c = (d->Q == 19) ? C19 : C9;
    g->w = (d->Q == 19) ? W19 : W9;
    g->opp = (d->Q == 19) ? OPP19 : OPP9;
    g->cs = ora_cs();
    g->ncell = (size_t)d->N * d->M * d->K;
    g->f = (double *)calloc(g->ncell * g->Q, sizeof(double));
    g->fNew = (double *)calloc(g->ncell * g->Q, sizeof(double));
    g->force_i = (double *)calloc(g->ncell * g->Q, sizeof(double));
    g->u = (double *)calloc(g->ncell * g->D, sizeof(double));
    g->u_n = (double *)calloc(g->ncell * g->D, sizeof(double));
    g->force_xyz = (double *)calloc(g->ncell * g->D, sizeof(double));
    g->rho = (double *)calloc(g->ncell, sizeof(double));
    g->LatTyp = (int *)calloc(g->ncell, sizeof(int));
    for (size_t i = 0; i < g->ncell; ++i) g->LatTyp[i] = O_FLUID;
    g->ux_in = (double *)calloc(d->M, sizeof(double));
    g->uy_in = (double *)calloc(d->M, sizeof(double));
    g->uz_in = (double *)calloc(d->M, sizeof(double));
    g->XPos = (double *)calloc(d->N, sizeof(double));
    g->YPos = (double *)calloc(d->M, sizeof(double));
    g->ZPos = (double *)calloc(d->K, sizeof(double));
    return g;
}

void ora_destroy(ora_grid *g) {
    if (!g) return;
    for (int k = 0; k < g->nkids; ++k) ora_destroy(g->kids[k]);
    free(g->f); free(g->fNew); free(g->force_i); free(g->u); free(g->u_n);
    free(g->force_xyz); free(g->rho); free(g->LatTyp);
    free(g->ux_in); free(g->uy_in); free(g->uz_in);
    free(g->XPos); free(g->YPos); free(g->ZPos);
    free(g);
}

int ora_add_child(ora_grid *parent, ora_grid *child) {
    if (parent->nkids >= ORA_MAX_KIDS) return -1;
    parent->kids[parent->nkids++] = child;
    child->parent = parent;
    return 0;
}

/* ---- GridUtils::getSubGrid: GridUtils.cpp:1761-1799 ----------------------------------
 * The region whose edges hold the site's position; for the sites that ask (TL2F sites, which
 * lie inside their region) that is the region whose coarse limits hold the site's indices. */
static const ora_grid *sub_grid(const ora_grid *g, int i, int j, int k) {
    for (int n = 0; n < g->nkids; ++n) {
        const int *L = g->kids[n]->d.coarse_lims;
        if (i >= L[0] && i <= L[1] && j >= L[2] && j <= L[3] && (g->D == 2 || (k >= L[4] && k <= L[5])))
            return g->kids[n];
    }
    return NULL;
}

void ora_set_labels(ora_grid *g, const int *lt) { memcpy(g->LatTyp, lt, g->ncell * sizeof(int)); }
void ora_set_positions(ora_grid *g, const double *x, const double *y, const double *z) {
    memcpy(g->XPos, x, g->d.N * sizeof(double));
    memcpy(g->YPos, y, g->d.M * sizeof(double));
    memcpy(g->ZPos, z, g->d.K * sizeof(double));
}
void ora_set_inlet(ora_grid *g, const double *ux, const double *uy, const double *uz) {
    memcpy(g->ux_in, ux, g->d.M * sizeof(double));
    memcpy(g->uy_in, uy, g->d.M * sizeof(double));
    memcpy(g->uz_in, uz, g->d.M * sizeof(double));
}
void ora_set_macro(ora_grid *g, const double *rho, const double *u) {
    memcpy(g->rho, rho, g->ncell * sizeof(double));
    memcpy(g->u, u, g->ncell * g->D * sizeof(double));
    memcpy(g->u_n, u, g->ncell * g->D * sizeof(double));
}
void ora_set_force(ora_grid *g, const double *fx) { memcpy(g->force_xyz, fx, g->ncell * g->D * sizeof(double)); }
void ora_set_f(ora_grid *g, const double *f) {
    memcpy(g->f, f, g->ncell * g->Q * sizeof(double));
    memcpy(g->fNew, f, g->ncell * g->Q * sizeof(double));
}
void ora_get_f(const ora_grid *g, double *f) { memcpy(f, g->f, g->ncell * g->Q * sizeof(double)); }
void ora_get_macro(const ora_grid *g, double *rho, double *u) {
    if (rho) memcpy(rho, g->rho, g->ncell * sizeof(double));
    if (u) memcpy(u, g->u, g->ncell * g->D * sizeof(double));
}
void ora_get_force(const ora_grid *g, double *fx) { memcpy(fx, g->force_xyz, g->ncell * g->D * sizeof(double)); }
int ora_get_t(const ora_grid *g) { return g->t; }
void ora_set_ibm_hook(ora_grid *g, ora_ibm_hook fn, void *ctx) { g->hook = fn; g->hook_ctx = ctx; }
double *ora_u(ora_grid *g) { return g->u; }
double *ora_u_n(ora_grid *g) { return g->u_n; }
double *ora_rho(ora_grid *g) { return g->rho; }
double *ora_force(ora_grid *g) { return g->force_xyz; }
double *ora_fnew(ora_grid *g) { return g->fNew; }
const int *ora_labels(const ora_grid *g) { return g->LatTyp; }
const ora_desc *ora_get_desc(const ora_grid *g) { return &g->d; }
const double *ora_pos(const ora_grid *g, int axis) { return axis == 0 ? g->XPos : axis == 1 ? g->YPos : g->ZPos; }
void ora_set_u_n_copy(ora_grid *g, int on) { g->copy_u_n = on; }

/* ---- _LBM_equilibrium_opt: GridObj_ops_lbm_optimised.cpp:674-705 ------------------- */
double ora_equilibrium(const ora_grid *g, int id, int v) {
    const int *c = g->c[v];
    const double *u = &g->u[(size_t)id * g->D];
    const double cs = g->cs;
    double A, B;
    if (g->D == 3) {
        A = (c[0] * u[0]) + (c[1] * u[1]) + (c[2] * u[2]);
        B = (SQ(c[0]) - SQ(cs)) * SQ(u[0]) + (SQ(c[1]) - SQ(cs)) * SQ(u[1]) +
            (SQ(c[2]) - SQ(cs)) * SQ(u[2]) + 2 * c[0] * c[1] * u[0] * u[1] +
            2 * c[0] * c[2] * u[0] * u[2] + 2 * c[1] * c[2] * u[1] * u[2];
    } else {
        A = (c[0] * u[0]) + (c[1] * u[1]);
        B = (SQ(c[0]) - SQ(cs)) * SQ(u[0]) + (SQ(c[1]) - SQ(cs)) * SQ(u[1]) +
            2 * c[0] * c[1] * u[0] * u[1];
    }
    return g->rho[id] * g->w[v] * (1.0 + (A / SQ(cs)) + (B / (2.0 * SQ(cs) * SQ(cs))));
}

void ora_init_feq(ora_grid *g) {
    /* GridObj_init_grids.cpp:315-333: f = feq; feq = f; fNew = f */
    for (size_t id = 0; id < g->ncell; ++id)
        for (int v = 0; v < g->Q; ++v) g->f[v + g->Q * id] = ora_equilibrium(g, (int)id, v);
    memcpy(g->fNew, g->f, g->ncell * g->Q * sizeof(double));
}

/* ---- GridUtils::getFineIndices / getCoarseIndices: GridUtils.cpp:410-466 ------------ */
static void fine_indices(const ora_grid *g, int ci, int x0, int cj, int y0, int ck, int z0, int out[3]) {
    out[0] = 2 * (ci - x0 + 1) - 2;
    out[1] = 2 * (cj - y0 + 1) - 2;
    out[2] = (g->D == 3) ? 2 * (ck - z0 + 1) - 2 : 0;
}
static void coarse_indices(const ora_grid *g, int fi, int x0, int fj, int y0, int fk, int z0, int out[3]) {
    if ((fi % 2) != 0) fi = fi - 1;
    if ((fj % 2) != 0) fj = fj - 1;
    if ((fk % 2) != 0) fk = fk - 1;
    out[0] = (fi / 2) + x0;
    out[1] = (fj / 2) + y0;
    out[2] = (g->D == 3) ? (fk / 2) + z0 : 0;
}

/* ---- _LBM_explode_opt: GridObj_ops_lbm_optimised.cpp:646-663 ------------------------- */
static void explode(ora_grid *g, int id, int v, int sx, int sy, int sz) {
    int p[3];
    coarse_indices(g, sx, g->d.coarse_lims[0], sy, g->d.coarse_lims[2], sz, g->d.coarse_lims[4], p);
    const ora_grid *P = g->parent;
    g->fNew[v + id * g->Q] =
        P->f[v + p[2] * P->Q + p[1] * P->Q * P->d.K + (size_t)p[0] * P->Q * P->d.K * P->d.M];
}

/* ---- _LBM_coalesce_opt: GridObj_ops_lbm_optimised.cpp:591-636 ------------------------ */
static void coalesce(ora_grid *g, const ora_grid *C, int i, int j, int k, int id, int v) {
    int cM = C->d.M, cK = C->d.K, Q = g->Q;
    int ci[3];
    fine_indices(g, i, C->d.coarse_lims[0], j, C->d.coarse_lims[2], k, C->d.coarse_lims[4], ci);
    double acc = 0.0;
    for (int ii = 0; ii < 2; ++ii)
        for (int jj = 0; jj < 2; ++jj)
            for (int kk = 0; kk < (g->D == 3 ? 2 : 1); ++kk)
                acc += C->f[v + (ci[2] + kk) * Q + (ci[1] + jj) * Q * cK + (size_t)(ci[0] + ii) * Q * cK * cM];
    if (g->D == 3) acc /= 8.0;
    else acc /= 4.0;
    g->fNew[v + id * Q] = acc;
}

/* ---- _LBM_stream_opt: GridObj_ops_lbm_optimised.cpp:206-297 (BFL and slip are out) ---- */
static void stream_site(ora_grid *g, int i, int j, int k, int id, int type_local, int subcycle) {
    const int N = g->d.N, M = g->d.M, K = g->d.K, Q = g->Q;
    for (int v = 0; v < Q; ++v) {
        int sx = (i - g->c[v][0] + N) % N;
        int sy = (j - g->c[v][1] + M) % M;
        int sz = (k - g->c[v][2] + K) % K;
        int sid = sz + sy * K + sx * K * M;
        int st = g->LatTyp[sid];
        if (st == O_SOLID) {
            g->fNew[v + id * Q] = g->f[g->opp[v] + id * Q];
        } else if (st == O_EXTRAP_RIGHT) {
            g->fNew[v + id * Q] = g->f[v + (sid - 2 * (K * M)) * Q];
        } else if (!g->d.regularised && st == O_VELOCITY) {
            if (g->ramp_time > 0.0) {
                /* :258-266: the source's u set from the ramped profile of this site's row j */
                const double r = ramp_coefficient(g, g->t * g->dt);
                g->u[0 + (size_t)sid * g->D] = g->ux_in[j] * r;
                g->u[1 + (size_t)sid * g->D] = g->uy_in[j] * r;
                if (g->D == 3) g->u[2 + (size_t)sid * g->D] = g->uz_in[j] * r;
            }
            g->fNew[v + id * Q] = ora_equilibrium(g, sid, v);
        } else if (g->parent && st == O_TL2C && subcycle == 0) {
            explode(g, id, v, sx, sy, sz);
        } else if (g->nkids && st == O_REFINED && type_local == O_TL2F && sub_grid(g, i, j, k)) {
            coalesce(g, sub_grid(g, i, j, k), i, j, k, id, v);
        } else {
            g->fNew[v + id * Q] = g->f[v + sid * Q];
        }
    }
}

/* ---- _LBM_macro_opt: GridObj_ops_lbm_optimised.cpp:800-893 ---------------------------- */
static void macro_site(ora_grid *g, int i, int j, int k, int id, int type_local) {
    const int Q = g->Q, D = g->D;
    if (type_local == O_FLUID || type_local == O_BFL || type_local == O_TL2F || type_local == O_SLIP) {
        double rho_t = 0.0, rx = 0.0, ry = 0.0, rz = 0.0;
        const double *fn = &g->fNew[(size_t)id * Q];
        for (int v = 0; v < Q; ++v) {
            rho_t += fn[v];
            rx += g->c[v][0] * fn[v];
            ry += g->c[v][1] * fn[v];
            if (D == 3) rz += g->c[v][2] * fn[v];
        }
        if (g->d.forced) {
            const double *F = &g->force_xyz[(size_t)id * D];
            rx += 0.5 * F[0];
            ry += 0.5 * F[1];
            if (D == 3) {
                if (g->d.fz_compat) rx += 0.5 * F[2];   /* :833 adds F_z/2 to x-momentum */
                else rz += 0.5 * F[2];
            }
        }
        double *u = &g->u[(size_t)id * D];
        u[0] = rx / rho_t;
        u[1] = ry / rho_t;
        if (D == 3) u[2] = rz / rho_t;
        g->rho[id] = rho_t;
    }
    if (type_local == O_TL2F && g->nkids && sub_grid(g, i, j, k)) {
        /* :850-891 copy u, rho into the child TL cluster */
        ora_grid *C = (ora_grid *)sub_grid(g, i, j, k);
        int ci[3];
        fine_indices(g, i, C->d.coarse_lims[0], j, C->d.coarse_lims[2], k, C->d.coarse_lims[4], ci);
        for (int ii = 0; ii < 2; ++ii)
            for (int jj = 0; jj < 2; ++jj)
                for (int kk = 0; kk < (D == 3 ? 2 : 1); ++kk) {
                    size_t cid = (size_t)(ci[2] + kk) + (size_t)(ci[1] + jj) * C->d.K +
                                 (size_t)(ci[0] + ii) * C->d.K * C->d.M;
                    for (int d = 0; d < D; ++d) C->u[d + cid * D] = g->u[d + (size_t)id * D];
                    C->rho[cid] = g->rho[id];
                }
    }
}

void ora_macro_site(ora_grid *g, int i, int j, int k) {
    int id = k + j * g->d.K + i * g->d.K * g->d.M;
    macro_site(g, i, j, k, id, g->LatTyp[id]);
}

/* ---- GridUtils::isWithinDomainWall: GridUtils.cpp:1369-1430 --------------------------- */
static int within_wall(const ora_grid *g, double px, double py, double pz, int nrm[3], int *ndir) {
    const double *th = g->d.wall_thk;
    int cnt = 0;
    nrm[0] = nrm[1] = nrm[2] = 0;
    *ndir = -1;
    if (px > 0.0 && px < th[0]) { *ndir = 0; nrm[0] = 1; cnt++; }
    if (px < g->d.xmax && px > g->d.xmax - th[1]) { *ndir = 0; nrm[0] = -1; cnt++; }
    if (py > 0.0 && py < th[2]) { *ndir = 1; nrm[1] = 1; cnt++; }
    if (py < g->d.ymax && py > g->d.ymax - th[3]) { *ndir = 1; nrm[1] = -1; cnt++; }
    if (g->D == 3) {
        if (pz > 0.0 && pz < th[4]) { *ndir = 2; nrm[2] = 1; cnt++; }
        if (pz < g->d.zmax && pz > g->d.zmax - th[5]) { *ndir = 2; nrm[2] = -1; cnt++; }
    }
    return cnt;
}

/* ---- _LBM_updateInteriorLatticeSite (:1422-1434) / _LBM_updateAndExtrapolate (:1353-1412)
 *      with GridUtils::extrapolate (GridUtils.h:231-280), order 1 on rho ------------------ */
static void update_interior_site(ora_grid *g, int i, int j, int k, int subcycle) {
    int id = k + j * g->d.K + i * g->d.K * g->d.M;
    int type_local = g->LatTyp[id];
    stream_site(g, i, j, k, id, type_local, subcycle);
    macro_site(g, i, j, k, id, type_local);
}
static int off_grid(const ora_grid *g, int i, int j, int k) {
    return i < 0 || i >= g->d.N || j < 0 || j >= g->d.M || k < 0 || k >= g->d.K;
}
static double update_and_extrapolate_rho(ora_grid *g, int subcycle, const int dir[3], int i, int j, int k,
                                         int *err) {
    int id = k + j * g->d.K + i * g->d.K * g->d.M;
    int i1 = i + dir[0], i2 = i1 + dir[0];
    int j1 = j + dir[1], j2 = j1 + dir[1];
    int k1 = k + dir[2], k2 = k1 + dir[2];
    if (off_grid(g, i1, j1, k1) || off_grid(g, i2, j2, k2)) { *err = 1; return 0.0; }
    int id1 = k1 + j1 * g->d.K + i1 * g->d.K * g->d.M;
    int id2 = k2 + j2 * g->d.K + i2 * g->d.K * g->d.M;
    if (id1 > id) update_interior_site(g, i1, j1, k1, subcycle);
    if (id2 > id) update_interior_site(g, i2, j2, k2, subcycle);
    return 2.0 * g->rho[id1] - g->rho[id2];
}
/* the same for velocity component d (the pressure face's tangential velocities, :393-399) */
static double update_and_extrapolate_u(ora_grid *g, int subcycle, const int dir[3], int i, int j, int k, int d,
                                       int *err) {
    int id = k + j * g->d.K + i * g->d.K * g->d.M;
    int i1 = i + dir[0], i2 = i1 + dir[0];
    int j1 = j + dir[1], j2 = j1 + dir[1];
    int k1 = k + dir[2], k2 = k1 + dir[2];
    if (off_grid(g, i1, j1, k1) || off_grid(g, i2, j2, k2)) { *err = 1; return 0.0; }
    int id1 = k1 + j1 * g->d.K + i1 * g->d.K * g->d.M;
    int id2 = k2 + j2 * g->d.K + i2 * g->d.K * g->d.M;
    if (id1 > id) update_interior_site(g, i1, j1, k1, subcycle);
    if (id2 > id) update_interior_site(g, i2, j2, k2, subcycle);
    return 2.0 * g->u[d + (size_t)id1 * g->D] - g->u[d + (size_t)id2 * g->D];
}

/* ---- _LBM_regularised_opt: GridObj_ops_lbm_optimised.cpp:313-510 (velocity sites; pressure
 *      sites on faces, :389-402, with rho = L_RHOIN + pd2dlbm(L_PRESSURE_DELTA), :317, :343) --- */
static void regularised_site(ora_grid *g, int i, int j, int k, int id, int type, int subcycle) {
    const int Q = g->Q, D = g->D;
    const double cs = g->cs;
    double vel[3] = {0, 0, 0};
    double dens = g->d.rho_in;
    dens += g->pressure_drho;
    int nrm[3], ndir;
    double f_plus = 0.0, f_zero = 0.0;
    double Sxx = 0, Syy = 0, Sxy = 0, Szz = 0, Sxz = 0, Syz = 0;
    double *fn = &g->fNew[(size_t)id * Q];
    int cnt = within_wall(g, g->XPos[i], g->YPos[j], g->ZPos[k], nrm, &ndir);
    if (cnt == 0) abort();   /* L_ERROR: velocity site not on a domain wall */
    const double ramp = ramp_coefficient(g, (g->t + 1) * g->dt);   /* :326 */
    vel[0] = g->ux_in[j] * ramp;
    vel[1] = g->uy_in[j] * ramp;
    vel[2] = g->uz_in[j] * ramp;
    if (cnt > 1) {
        if (type != O_VELOCITY) abort();   /* L_ERROR: pressure BC not supported at corner/edge */
        int err = 0;
        dens = update_and_extrapolate_rho(g, subcycle, nrm, i, j, k, &err);
        if (err) abort();
    } else {
        for (int v = 0; v < Q; ++v) {
            if (g->c[v][ndir] == -nrm[ndir]) f_plus += fn[v];
            else if (g->c[v][ndir] == 0) f_zero += fn[v];
        }
        if (type == O_PRESSURE) {
            for (int d = 0; d < D; d++)
                if (d != ndir) {
                    int err = 0;
                    vel[d] = update_and_extrapolate_u(g, subcycle, nrm, i, j, k, d, &err);
                    if (err) abort();
                }
            vel[ndir] = 1.0 - ((1.0 / dens) * (2.0 * f_plus + f_zero));
            if (nrm[ndir] == -1) vel[ndir] *= -1.0;
        } else {
            double nv = vel[ndir];
            if (nrm[ndir] == -1) nv *= -1.0;
            dens = (1.0 / (1.0 - nv)) * (2.0 * f_plus + f_zero);
        }
    }
    g->rho[id] = dens;
    g->u[0 + (size_t)id * D] = vel[0];
    g->u[1 + (size_t)id * D] = vel[1];
    if (D == 3) g->u[2 + (size_t)id * D] = vel[2];

    for (int v = 0; v < Q; ++v) {
        const int *c = g->c[v];
        int o = g->opp[v];
        if (cnt == 1 && c[ndir] == nrm[ndir]) {
            fn[v] = ora_equilibrium(g, id, v) + (fn[o] - ora_equilibrium(g, id, o));
        } else if (cnt > 1 && (c[0] == nrm[0] || c[1] == nrm[1] || (D == 3 && c[2] == nrm[2]))) {
            int dp = 0;
            double mag = 0.0;
            for (int d = 0; d < D; d++) {
                dp += c[d] * nrm[d];
                mag += ((double)c[d] * (double)c[d]);
            }
            mag = sqrt(mag);
            if (dp == 0 && mag > 1.0) fn[v] = ora_equilibrium(g, id, v);
            else fn[v] = ora_equilibrium(g, id, v) + (fn[o] - ora_equilibrium(g, id, o));
        }
        double fneq = fn[v] - ora_equilibrium(g, id, v);
        Sxx += c[0] * c[0] * fneq;
        Syy += c[1] * c[1] * fneq;
        Sxy += c[0] * c[1] * fneq;
        if (D == 3) {
            Szz += c[2] * c[2] * fneq;
            Sxz += c[0] * c[2] * fneq;
            Syz += c[1] * c[2] * fneq;
        }
    }
    for (int v = 0; v < Q; v++) {
        const int *c = g->c[v];
        fn[v] = ora_equilibrium(g, id, v) +
                (g->w[v] / (2.0 * SQ(cs) * SQ(cs))) *
                    (((c[0] * c[0] - SQ(cs)) * Sxx) + ((c[1] * c[1] - SQ(cs)) * Syy) +
                     ((c[2] * c[2] - SQ(cs)) * Szz) + (2.0 * c[0] * c[1] * Sxy) +
                     (2.0 * c[0] * c[2] * Sxz) + (2.0 * c[1] * c[2] * Syz));
    }
}

/* ---- _LBM_forceGrid_opt: GridObj_ops_lbm_optimised.cpp:928-990 ----------------------- */
static void force_site(ora_grid *g, int id) {
    const int Q = g->Q, D = g->D;
    const double cs = g->cs;
    double *fi = &g->force_i[(size_t)id * Q];
    const double *u = &g->u[(size_t)id * D];
    const double *F = &g->force_xyz[(size_t)id * D];
    memset(fi, 0, sizeof(double) * Q);
    for (int v = 0; v < Q; v++) {
        double beta_v = 0.0;
        double lambda_v = (1 - 0.5 * g->d.omega) * (g->w[v] / (cs * cs));
        for (int d = 0; d < D; d++) beta_v += (g->c[v][d] * u[d]);
        beta_v = beta_v * (1 / (cs * cs));
        for (int d = 0; d < D; d++) fi[v] += F[d] * (g->c[v][d] * (1 + beta_v) - u[d]);
        fi[v] *= lambda_v;
    }
}

/* ---- _LBM_collide_opt (BGK): GridObj_ops_lbm_optimised.cpp:765-790 ------------------- */
static void collide_site(ora_grid *g, int id) {
    const int Q = g->Q;
    double *fn = &g->fNew[(size_t)id * Q];
    const double om = g->d.omega;
    if (g->d.forced) {
        const double *fi = &g->force_i[(size_t)id * Q];
        for (int v = 0; v < Q; ++v) fn[v] += om * (ora_equilibrium(g, id, v) - fn[v]) + fi[v];
    } else {
        for (int v = 0; v < Q; ++v) fn[v] += om * (ora_equilibrium(g, id, v) - fn[v]);
    }
}

/* ---- _LBM_resetForces: :1327-1337 -------------------------------------------------------
 * Gravity off: every component zeroed.  Gravity on: only the gravity component is written,
 * rho[id] * gravity * refinement_ratio with refinement_ratio = 1 / 2^level (GridObj.cpp:54);
 * the other components keep their values. */
void ora_reset_forces(ora_grid *g) {
    if (!g->gravity_on) {
        memset(g->force_xyz, 0, g->ncell * g->D * sizeof(double));
        return;
    }
    const double rr = 1.0 / pow(2.0, (double)g->d.level);
    for (size_t id = 0; id < g->ncell; ++id)
        g->force_xyz[g->gravity_dir + id * g->D] = g->rho[id] * g->gravity * rr;
}
void ora_set_gravity(ora_grid *g, double gravity, int dir) {
    g->gravity_on = 1;
    g->gravity = gravity;
    g->gravity_dir = dir;
}

/* ---- sweep 1 of LBM_multi_opt: :68-107 ------------------------------------------------ */
static void sweep1_plane(ora_grid *g, int i, int subcycle) {
    const int M = g->d.M, K = g->d.K;
    for (int j = 0; j < M; ++j)
        for (int k = 0; k < K; ++k) {
            int id = k + j * K + i * K * M;
            int tl = g->LatTyp[id];
            if (tl == O_REFINED || tl == O_SOLID || (!g->d.regularised && tl == O_VELOCITY)) continue;
            stream_site(g, i, j, k, id, tl, subcycle);
            if (g->d.regularised && (tl == O_VELOCITY || tl == O_PRESSURE))
                regularised_site(g, i, j, k, id, tl, subcycle);
            macro_site(g, i, j, k, id, tl);
        }
}
void ora_sweep1(ora_grid *g, int subcycle) {
    for (int i = 0; i < g->d.N; ++i) sweep1_plane(g, i, subcycle);
}

/* ---- sweep 2 of LBM_multi_opt: :125-156, swap :159, ++t :172 ------------------------- */
static void sweep2_plane(ora_grid *g, int i) {
    const int M = g->d.M, K = g->d.K;
    for (int j = 0; j < M; ++j)
        for (int k = 0; k < K; ++k) {
            int id = k + j * K + i * K * M;
            int tl = g->LatTyp[id];
            if (g->d.forced && tl != O_SOLID) force_site(g, id);
            if (tl != O_TL2C) collide_site(g, id);
        }
}
static void swap_and_tick(ora_grid *g) {
    double *tmp = g->f;
    g->f = g->fNew;
    g->fNew = tmp;
    ++g->t;
}
void ora_sweep2(ora_grid *g) {
    for (int i = 0; i < g->d.N; ++i) sweep2_plane(g, i);
    swap_and_tick(g);
}

/* ---- LBM_multi_opt: GridObj_ops_lbm_optimised.cpp:36-193 ---------------------------- */
void ora_multi_opt(ora_grid *g, int subcycle) {
    if (g->re_ramp > 0.0) update_reynolds(g);
    for (int n = 0; n < g->nkids; ++n) {
        ora_multi_opt(g->kids[n], 0);
        ora_multi_opt(g->kids[n], 1);
    }
    if (g->hook) ora_reset_forces(g);
    ora_sweep1(g, subcycle);
    if (g->copy_u_n) memcpy(g->u_n, g->u, g->ncell * g->D * sizeof(double));
    if (g->hook) g->hook(g, g->hook_ctx);
    ora_sweep2(g);
}

/* OpenMP timing mode for the CPU baseline: the same per-site arithmetic, sweeps split over
 * the x index like LUMA's own experimental pragma (:69-71).  Only an all-fluid grid runs in
 * parallel: a regularised edge or corner site updates its neighbours on the fly (:1376) and a
 * forced-equilibrium source under a ramp writes the source's u (:258-266), neither of which a
 * parallel sweep orders as the serial one does, so every other grid steps serially. */
static int all_fluid(const ora_grid *g) {
    for (size_t id = 0; id < g->ncell; ++id)
        if (g->LatTyp[id] != O_FLUID) return 0;
    return 1;
}
void ora_run(ora_grid *g, int nsteps, int nthreads) {
    if (nthreads <= 1 || g->nkids || g->hook || g->re_ramp > 0.0 || !all_fluid(g)) {
        for (int s = 0; s < nsteps; ++s) ora_multi_opt(g, 0);
        return;
    }
#ifdef _OPENMP
    for (int s = 0; s < nsteps; ++s) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int i = 0; i < g->d.N; ++i) sweep1_plane(g, i, 0);
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int i = 0; i < g->d.N; ++i) sweep2_plane(g, i);
        swap_and_tick(g);
    }
#else
    for (int s = 0; s < nsteps; ++s) ora_multi_opt(g, 0);
#endif
}
