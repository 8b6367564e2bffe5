/*
 * ibm_oracle.c -- FP64 CPU restatement of LUMA v1.7.12's immersed-boundary step
 * (ObjectManager::ibm_apply and the set-up it depends on).
 *
 * TEST INFRASTRUCTURE ONLY (the checker); see lbm_oracle.h.  Serial semantics: every marker
 * and support site is on "this rank".  Each function names the reference lines it restates.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "fem_oracle.h"
#include "lbm_oracle.h"

#define SMAX 27

typedef struct {
    double pos[3], pos0[3], vel[3], vel_km1[3], force[3], mom[3];
    double rho, eps, ds, dilation, area;
    int ns;
    int si[SMAX], sj[SMAX], sk[SMAX];
    double sx[SMAX], sy[SMAX], sz[SMAX], delta[SMAX];
} omarker;

typedef struct {
    int nm;
    omarker *m;
    int flexible;
    ofem *fem;
} obody;

typedef struct ora_ibm {
    ora_grid *g;
    int nb;
    obody b[16];
    double dh, origin[3];
    double dt, ux0;   /* grid dt and L_UX0 for the sub-iteration residual */
    int err;
    int sub_it;       /* sub-iterations of the last step */
    double sub_res;
    int universal;    /* L_UNIVERSAL_EPSILON_CALC */
} ora_ibm;

static int D_of(const ora_ibm *s) { return ora_get_desc(s->g)->dims; }

/* ObjectManager::ibm_deltaKernel, ObjectManager_ops_ibm.cpp:255-272 (Roma et al. 3-point) */
double ora_ibm_delta(double radius, double dilation) {
    double mag_r = fabs(radius) / dilation, value;
    if (mag_r > 1.5) value = 0.0;
    else if (mag_r > 0.5) value = (5.0 - (3.0 * mag_r) - sqrt(-3.0 * pow(1.0 - mag_r, 2) + 1.0)) / 6.0;
    else value = (1.0 + sqrt(1.0 - 3.0 * pow(mag_r, 2))) / 3.0;
    return value;
}

ora_ibm *ora_ibm_create(ora_grid *g, double dh, const double *origin) {
    ora_ibm *s = (ora_ibm *)calloc(1, sizeof(ora_ibm));
    s->g = g;
    s->dh = dh;
    for (int a = 0; a < 3; ++a) s->origin[a] = origin[a];
    return s;
}

void ora_ibm_destroy(ora_ibm *s) {
    if (!s) return;
    for (int i = 0; i < s->nb; ++i) {
        free(s->b[i].m);
        ora_fem_destroy(s->b[i].fem);
    }
    free(s);
}

/* Body<IBMarker>::addMarker + IBMarker ctor (IBMarker.cpp:56-86): eps 0, area 1, dilation 1,
 * ds 1, markerVel 0 */
int ora_ibm_add_body(ora_ibm *s, int n, const double *pos, int flexible) {
    if (s->nb >= 16) return -1;
    obody *b = &s->b[s->nb];
    b->nm = n;
    b->flexible = flexible;
    b->m = (omarker *)calloc(n, sizeof(omarker));
    for (int i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) b->m[i].pos[a] = b->m[i].pos0[a] = pos[3 * i + a];
        b->m[i].area = 1.0;
        b->m[i].dilation = 1.0;
        b->m[i].ds = 1.0;
    }
    return s->nb++;
}

/* IBBody filament ctor with an FEM body (IBBody.cpp:240-258, Body.h:144-187, 2D): markers at
 * floor(L/dh) + 1 points along the filament, FEMBody of nEl corotational beam elements */
int ora_ibm_add_filament(ora_ibm *s, const double *start, double length, double height, double depth,
                         double angle_deg, int nEl, int clamped, double density, double E, double dt, double dm) {
    const double dh = s->dh;
    const int n = (int)floor(length / dh) + 1;
    const double spacing = length / (n - 1);
    const double spacing_h = spacing * cos(angle_deg * 3.14159265358979323846 / 180);
    double *pos = (double *)malloc(sizeof(double) * 3 * n);
    for (int i = 0; i < n; ++i) {
        pos[3 * i] = start[0] + i * spacing_h * cos(0.0 * 3.14159265358979323846 / 180.0);
        pos[3 * i + 1] = start[1] + i * spacing * sin(angle_deg * 3.14159265358979323846 / 180.0);
        pos[3 * i + 2] = start[2] + i * spacing_h * sin(0.0 * 3.14159265358979323846 / 180.0);
    }
    int ib = ora_ibm_add_body(s, n, pos, 1);
    free(pos);
    if (ib < 0) return ib;
    s->b[ib].fem = ora_fem_create(start, length, height, depth, angle_deg, nEl, clamped, density, E, dh, dt, dm);
    s->dt = dt;
    s->ux0 = 1.0;
    return ib;
}

/* ibm_initialiseSupport, :462-491 */
static void init_support(ora_ibm *s, omarker *m, const double est[3]) {
    double dx = (est[0] - m->pos[0]) / s->dh;
    double dy = (est[1] - m->pos[1]) / s->dh;
    double val = ora_ibm_delta(dx, m->dilation) * ora_ibm_delta(dy, m->dilation);
    if (D_of(s) == 3) {
        double dz = (est[2] - m->pos[2]) / s->dh;
        val = ora_ibm_delta(dx, m->dilation) * ora_ibm_delta(dy, m->dilation) * ora_ibm_delta(dz, m->dilation);
    }
    m->delta[m->ns] = val;
}

static int within_domain(const ora_ibm *s, const double p[3]) {
    /* GridUtils::isWithinDomain (GridUtils.cpp:1311-1330), level-0 edges from 0 */
    const ora_desc *d = ora_get_desc(s->g);
    if (p[0] < 0.0 || p[0] > d->xmax || p[1] < 0.0 || p[1] > d->ymax) return 0;
    if (d->dims == 3 && (p[2] < 0.0 || p[2] > d->zmax)) return 0;
    return 1;
}

static int enclosing(const ora_ibm *s, double x, int axis) {
    /* GridUtils::getEnclosingVoxel, serial (GridUtils.cpp:1630-1719) */
    return (int)floor((x - s->origin[axis]) / s->dh);
}

/* ibm_findSupport, ObjectManager_ops_ibm.cpp:279-453 */
static int find_support(ora_ibm *s, obody *b) {
    const ora_desc *d = ora_get_desc(s->g);
    const int D = d->dims;
    const double *X = ora_pos(s->g, 0), *Y = ora_pos(s->g, 1), *Z = ora_pos(s->g, 2);
    for (int mi = 0; mi < b->nm; ++mi) {
        omarker *m = &b->m[mi];
        m->ns = 0;
        int inear = enclosing(s, m->pos[0], 0), jnear = enclosing(s, m->pos[1], 1);
        int knear = (D == 3) ? enclosing(s, m->pos[2], 2) : 0;
        if (inear < 0 || jnear < 0 || knear < 0 || inear >= d->N || jnear >= d->M || knear >= d->K) return -1;
        double nearpos[3] = {X[inear], Y[jnear], (D == 3) ? Z[knear] : 0.0};
        m->si[0] = inear;
        m->sj[0] = jnear;
        m->sk[0] = knear;
        m->sx[0] = nearpos[0];
        m->sy[0] = nearpos[1];
        m->sz[0] = nearpos[2];
        init_support(s, m, nearpos);
        m->ns = 1;
        for (int i = inear - 5; i <= inear + 5; i++)
            for (int j = jnear - 5; j <= jnear + 5; j++)
                for (int k = (D == 3 ? knear - 5 : 0); k <= (D == 3 ? knear + 5 : 0); k++) {
                    double est[3] = {0.0, 0.0, 0.0};
                    est[0] = nearpos[0] + (i - inear) * s->dh;
                    est[1] = nearpos[1] + (j - jnear) * s->dh;
                    if (D == 3) est[2] = nearpos[2] + (k - knear) * s->dh;
                    int in = (fabs(m->pos[0] - est[0]) / s->dh < 1.5 * m->dilation) &&
                             (fabs(m->pos[1] - est[1]) / s->dh < 1.5 * m->dilation);
                    if (D == 3) in = in && (fabs(m->pos[2] - est[2]) / s->dh < 1.5 * m->dilation);
                    if (in && within_domain(s, est)) {
                        if (i != inear || j != jnear || (D == 3 && k != knear)) {
                            if (m->ns >= SMAX) return -2;
                            if (i < 0 || j < 0 || k < 0 || i >= d->N || j >= d->M || k >= d->K) return -3;
                            m->si[m->ns] = i;
                            m->sj[m->ns] = j;
                            m->sk[m->ns] = k;
                            m->sx[m->ns] = est[0];
                            m->sy[m->ns] = est[1];
                            m->sz[m->ns] = est[2];
                            init_support(s, m, est);
                            m->ns++;
                        }
                    }
                }
    }
    return 0;
}

/* ibm_computeDs, :927-981 (GridUtils::vecnorm = sqrt of the sum of pow(d, 2)) */
static void compute_ds(ora_ibm *s, obody *b) {
    for (int mi = 0; mi < b->nm; ++mi) {
        double ds = 10.0;
        for (int n = 0; n < b->nm; ++n) {
            if (n == mi) continue;
            double acc = 0.0;
            for (int a = 0; a < 3; ++a) acc += pow(b->m[mi].pos[a] - b->m[n].pos[a], 2);
            double dist = sqrt(acc) / s->dh;
            if (dist < ds) ds = dist;
        }
        b->m[mi].ds = ds;
    }
}

/* Dense solve A x = b by LU with partial pivoting.  GridUtils::solveLinearSystem
 * (GridUtils.cpp:356-395) hands LAPACK dgetrf_/dgetrs_('T') the row-major A, i.e. factors
 * A^T = P L U with row pivoting and solves through the transpose.  Restated here as LAPACK's
 * unblocked dgetf2 + dgetrs('T') sequence. */
int ora_lu_solve_t(int n, double *a /* row-major A, destroyed */, double *b) {
    /* column-major view: M(i,j) = a[j*n + i] is A^T */
#define MM(i, j) a[(size_t)(j) * n + (i)]
    int *piv = (int *)malloc(sizeof(int) * n);
    for (int j = 0; j < n; ++j) {
        int p = j;
        double best = fabs(MM(j, j));
        for (int i = j + 1; i < n; ++i)
            if (fabs(MM(i, j)) > best) { best = fabs(MM(i, j)); p = i; }
        piv[j] = p;
        if (MM(p, j) == 0.0) { free(piv); return -1; }
        if (p != j)
            for (int k = 0; k < n; ++k) { double t = MM(j, k); MM(j, k) = MM(p, k); MM(p, k) = t; }
        double r = 1.0 / MM(j, j);
        for (int i = j + 1; i < n; ++i) MM(i, j) *= r;
        for (int k = j + 1; k < n; ++k) {
            double t = MM(j, k);
            if (t != 0.0)
                for (int i = j + 1; i < n; ++i) MM(i, k) -= MM(i, j) * t;
        }
    }
    /* solve M^T x = b with M = P L U: U^T y = b, L^T z = y, x = P z */
    for (int j = 0; j < n; ++j) {
        double t = b[j];
        for (int i = 0; i < j; ++i) t -= MM(i, j) * b[i];
        b[j] = t / MM(j, j);
    }
    for (int j = n - 1; j >= 0; --j) {
        double t = b[j];
        for (int i = j + 1; i < n; ++i) t -= MM(i, j) * b[i];
        b[j] = t;
    }
    for (int j = n - 1; j >= 0; --j)
        if (piv[j] != j) { double t = b[j]; b[j] = b[piv[j]]; b[piv[j]] = t; }
    free(piv);
    return 0;
#undef MM
}

/* ibm_findEpsilon, :773-920 */
static int find_epsilon(ora_ibm *s, obody *b) {
    const int n = b->nm, D = D_of(s);
    double *A = (double *)calloc((size_t)n * n, sizeof(double));
    double *rhs = (double *)malloc(sizeof(double) * n);
    for (int I = 0; I < n; ++I) {
        omarker *mI = &b->m[I];
        for (int J = 0; J < n; ++J) {
            omarker *mJ = &b->m[J];
            double acc = 0.0;
            for (int k = 0; k < mI->ns; ++k) {
                double dI = mI->delta[k], dJ;
                if (D == 3)
                    dJ = ora_ibm_delta((mJ->pos[0] - mI->sx[k]) / s->dh, mJ->dilation) *
                         ora_ibm_delta((mJ->pos[1] - mI->sy[k]) / s->dh, mJ->dilation) *
                         ora_ibm_delta((mJ->pos[2] - mI->sz[k]) / s->dh, mJ->dilation);
                else
                    dJ = ora_ibm_delta((mJ->pos[0] - mI->sx[k]) / s->dh, mJ->dilation) *
                         ora_ibm_delta((mJ->pos[1] - mI->sy[k]) / s->dh, mJ->dilation);
                acc += dI * dJ * mI->area;
            }
            A[(size_t)I * n + J] = acc * mJ->ds;
        }
    }
    for (int i = 0; i < n; ++i) rhs[i] = 1.0;
    int r = ora_lu_solve_t(n, A, rhs);
    for (int i = 0; i < n && r == 0; ++i) b->m[i].eps = rhs[i];
    free(A);
    free(rhs);
    return r;
}

/* ibm_findEpsilon over the level (:773-920): per body, or with L_UNIVERSAL_EPSILON_CALC one
 * system over every marker of the level in body order (ibm_universalEpsilonGather, :1067-1107),
 * its epsilon scattered back in the same order (ibm_universalEpsilonScatter, :1113-1141) */
static int find_epsilon_all(ora_ibm *s) {
    if (!s->universal) {
        for (int i = 0; i < s->nb; ++i)
            if (find_epsilon(s, &s->b[i])) return -1;
        return 0;
    }
    int n = 0;
    for (int i = 0; i < s->nb; ++i) n += s->b[i].nm;
    if (n == 0) return 0;
    obody all = {0};
    all.nm = n;
    all.m = (omarker *)malloc(sizeof(omarker) * (size_t)n);
    for (int i = 0, k = 0; i < s->nb; ++i)
        for (int m = 0; m < s->b[i].nm; ++m) all.m[k++] = s->b[i].m[m];
    int r = find_epsilon(s, &all);
    for (int i = 0, k = 0; i < s->nb && r == 0; ++i)
        for (int m = 0; m < s->b[i].nm; ++m) s->b[i].m[m].eps = all.m[k++].eps;
    free(all.m);
    return r;
}

void ora_ibm_set_universal_epsilon(ora_ibm *s, int on) { s->universal = on ? 1 : 0; }

/* ObjectManager::ibm_initialise, :190-246 */
int ora_ibm_initialise(ora_ibm *s) {
    for (int i = 0; i < s->nb; ++i)
        if (find_support(s, &s->b[i])) return -1;
    for (int i = 0; i < s->nb; ++i) compute_ds(s, &s->b[i]);
    if (find_epsilon_all(s)) return -2;
    return 0;
}

static size_t site_id(const ora_desc *d, int i, int j, int k) { return (size_t)k + (size_t)j * d->K + (size_t)i * d->K * d->M; }

/* ibm_interpolate (:498-591), ibm_computeForce (:598-614), ibm_spread (:621-689),
 * ibm_updateMacroscopic (:696-766) */
static void ibm_apply_core(ora_grid *g, ora_ibm *s);
static int has_flexible(const ora_ibm *s) {
    for (int i = 0; i < s->nb; ++i)
        if (s->b[i].flexible) return 1;
    return 0;
}

/* ibm_moveBodies, ObjectManager_ops_ibm.cpp:61-104 (serial) */
static void move_bodies(ora_ibm *s) {
    for (int ib = 0; ib < s->nb; ++ib) {
        obody *b = &s->b[ib];
        if (!b->flexible || !b->fem) continue;
        const int n = b->nm;
        double *force = (double *)malloc(sizeof(double) * 3 * n), *eps = (double *)malloc(sizeof(double) * n);
        double *pos = (double *)malloc(sizeof(double) * 3 * n), *pos0 = (double *)malloc(sizeof(double) * 3 * n);
        double *vel = (double *)malloc(sizeof(double) * 3 * n), *vkm1 = (double *)malloc(sizeof(double) * 3 * n);
        for (int i = 0; i < n; ++i) {
            eps[i] = b->m[i].eps;
            for (int a = 0; a < 3; ++a) {
                force[3 * i + a] = b->m[i].force[a];
                pos[3 * i + a] = b->m[i].pos[a];
                pos0[3 * i + a] = b->m[i].pos0[a];
                vel[3 * i + a] = b->m[i].vel[a];
                vkm1[3 * i + a] = b->m[i].vel_km1[a];
            }
        }
        if (ora_fem_dynamic(b->fem, force, eps, pos, pos0, vel, vkm1)) s->err = 1;
        for (int i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) {
                b->m[i].pos[a] = pos[3 * i + a];
                b->m[i].vel[a] = vel[3 * i + a];
                b->m[i].vel_km1[a] = vkm1[3 * i + a];
            }
        free(force); free(eps); free(pos); free(pos0); free(vel); free(vkm1);
    }
    for (int ib = 0; ib < s->nb; ++ib)
        if (s->b[ib].flexible && find_support(s, &s->b[ib])) s->err = 2;
    for (int ib = 0; ib < s->nb; ++ib) compute_ds(s, &s->b[ib]);
    if (find_epsilon_all(s)) s->err = 3;
}

/* ibm_checkVelDiff, :988-1032 */
static double vel_diff(const ora_ibm *s) {
    double res = 0.0;
    for (int ib = 0; ib < s->nb; ++ib) {
        const obody *b = &s->b[ib];
        if (!b->flexible) continue;
        for (int i = 0; i < b->nm; ++i) {
            double acc = 0.0;
            for (int a = 0; a < 3; ++a) acc += pow(b->m[i].vel[a] - b->m[i].vel_km1[a], 2);
            double v = sqrt(acc);
            v = v / (s->ux0 * s->dt / s->dh);
            if (fabs(v) > res) res = fabs(v);
        }
    }
    return res;
}

/* ObjectManager::ibm_apply (:33-54) and ibm_subIterate (:111-185) */
static void ibm_apply_flag(ora_grid *g, ora_ibm *s, int doSub) {
    ibm_apply_core(g, s);
    if (!has_flexible(s)) return;
    move_bodies(s);
    if (!doSub) return;
    const ora_desc *d = ora_get_desc(g);
    const size_t nu = (size_t)d->N * d->M * d->K * d->dims;
    double res;
    int it = 0;
    do {
        memcpy(ora_u(g), ora_u_n(g), nu * sizeof(double));
        ora_reset_forces(g);
        ibm_apply_flag(g, s, 0);
        res = vel_diff(s);
        it++;
    } while (res > 1e-4 && it < 10);
    s->sub_it = it;
    s->sub_res = res;
    for (int ib = 0; ib < s->nb; ++ib)
        if (s->b[ib].flexible && s->b[ib].fem) ora_fem_commit(s->b[ib].fem);
}

static void ibm_apply(ora_grid *g, void *ctx) { ibm_apply_flag(g, (ora_ibm *)ctx, 1); }

static void ibm_apply_core(ora_grid *g, ora_ibm *s) {
    const ora_desc *d = ora_get_desc(g);
    const int D = d->dims;
    double *rho = ora_rho(g), *u = ora_u(g), *F = ora_force(g);
    for (int ib = 0; ib < s->nb; ++ib) {
        obody *b = &s->b[ib];
        for (int mi = 0; mi < b->nm; ++mi) {
            omarker *m = &b->m[mi];
            m->mom[0] = m->mom[1] = m->mom[2] = 0.0;
            m->rho = 0.0;
            for (int k = 0; k < m->ns; ++k) {
                size_t id = site_id(d, m->si[k], m->sj[k], m->sk[k]);
                m->rho += rho[id] * m->delta[k] * m->area;
                for (int dir = 0; dir < D; ++dir) m->mom[dir] += rho[id] * u[dir + D * id] * m->delta[k] * m->area;
            }
        }
    }
    for (int ib = 0; ib < s->nb; ++ib)
        for (int mi = 0; mi < s->b[ib].nm; ++mi) {
            omarker *m = &s->b[ib].m[mi];
            for (int dir = 0; dir < D; ++dir) m->force[dir] = 2.0 * (m->mom[dir] - m->rho * m->vel[dir]) / 1.0;
        }
    for (int ib = 0; ib < s->nb; ++ib)
        for (int mi = 0; mi < s->b[ib].nm; ++mi) {
            omarker *m = &s->b[ib].m[mi];
            for (int k = 0; k < m->ns; ++k) {
                double volWidth = m->eps, volDepth = (D == 3) ? m->ds : 1.0;
                size_t id = site_id(d, m->si[k], m->sj[k], m->sk[k]);
                for (int dir = 0; dir < D; ++dir)
                    F[dir + D * id] -= m->delta[k] * m->force[dir] * volWidth * volDepth * m->ds;
            }
        }
    for (int ib = 0; ib < s->nb; ++ib)
        for (int mi = 0; mi < s->b[ib].nm; ++mi) {
            omarker *m = &s->b[ib].m[mi];
            for (int k = 0; k < m->ns; ++k) ora_macro_site(g, m->si[k], m->sj[k], m->sk[k]);
        }
}

void ora_ibm_attach(ora_ibm *s) {
    ora_set_ibm_hook(s->g, ibm_apply, s);
    if (has_flexible(s)) ora_set_u_n_copy(s->g, 1);
}
int ora_ibm_error(const ora_ibm *s) { return s->err; }
void ora_ibm_subiter(const ora_ibm *s, int *it, double *res) {
    *it = s->sub_it;
    *res = s->sub_res;
}
int ora_ibm_fem_state(const ora_ibm *s, int ib, double *U, double *Udot, int *it, double *res) {
    if (!s->b[ib].fem) return 0;
    ora_fem_state(s->b[ib].fem, U, Udot, it, res);
    return ora_fem_ndof(s->b[ib].fem);
}

void ora_ibm_set_velocity(ora_ibm *s, int ib, const double *vel) {
    for (int i = 0; i < s->b[ib].nm; ++i)
        for (int a = 0; a < 3; ++a) s->b[ib].m[i].vel[a] = vel[3 * i + a];
}
void ora_ibm_set_epsilon(ora_ibm *s, int ib, const double *eps) {
    for (int i = 0; i < s->b[ib].nm; ++i) s->b[ib].m[i].eps = eps[i];
}

/* per-marker record: pos[3] vel[3] force[3] mom[3] rho eps ds ns = 16 doubles */
void ora_ibm_get_markers(const ora_ibm *s, int ib, double *out) {
    for (int i = 0; i < s->b[ib].nm; ++i) {
        const omarker *m = &s->b[ib].m[i];
        double *o = out + 19 * (size_t)i;
        for (int a = 0; a < 3; ++a) {
            o[a] = m->pos[a];
            o[3 + a] = m->vel[a];
            o[6 + a] = m->force[a];
            o[9 + a] = m->mom[a];
            o[16 + a] = m->vel_km1[a];
        }
        o[12] = m->rho;
        o[13] = m->eps;
        o[14] = m->ds;
        o[15] = m->ns;
    }
}

/* support list of a marker: (i, j, k) triples and delta values; returns the count */
int ora_ibm_get_support(const ora_ibm *s, int ib, int mi, int *ijk, double *delta) {
    const omarker *m = &s->b[ib].m[mi];
    for (int k = 0; k < m->ns; ++k) {
        ijk[3 * k] = m->si[k];
        ijk[3 * k + 1] = m->sj[k];
        ijk[3 * k + 2] = m->sk[k];
        delta[k] = m->delta[k];
    }
    return m->ns;
}
