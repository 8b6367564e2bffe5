/*
 * fem_oracle.c -- FP64 CPU restatement of LUMA v1.7.12's corotational beam FEM
 * (FEMBody.cpp, FEMElement.cpp, FEMNode.cpp) for a 2D filament coupled to IB markers.
 *
 * TEST INFRASTRUCTURE ONLY (the checker); see lbm_oracle.h.  Matrices are dense row-major;
 * the small products follow GridUtils::matrix_multiply (GridUtils.cpp:257-308): rows, then
 * columns, then the inner index, each accumulating from 0.0.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "fem_oracle.h"

#define SQ(x) ((x) * (x))
#define TH(x) ((x) * (x) * (x))
#define L_PI 3.14159265358979323846
#define NB_ALPHA 0.25   /* L_NB_ALPHA, definitions.h:223 */
#define NB_DELTA 0.5    /* L_NB_DELTA, definitions.h:224 */
#define RELAX 0.5       /* L_RELAX, definitions.h:225 */

static void mat6_vec(const double A[6][6], const double *x, double *y) {
    for (int r = 0; r < 6; ++r) {
        y[r] = 0.0;
        for (int c = 0; c < 6; ++c) y[r] += A[r][c] * x[c];
    }
}
static void mat6_T_vec(const double A[6][6], const double *x, double *y) {   /* A^T x */
    for (int r = 0; r < 6; ++r) {
        y[r] = 0.0;
        for (int c = 0; c < 6; ++c) y[r] += A[c][r] * x[c];
    }
}
/* T^T * L * T, as matrix_multiply(matrix_multiply(transpose(T), L), T) */
static void congruence(const double T[6][6], const double L[6][6], double G[6][6]) {
    double TL[6][6];
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) {
            TL[r][c] = 0.0;
            for (int k = 0; k < 6; ++k) TL[r][c] += T[k][r] * L[k][c];
        }
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) {
            G[r][c] = 0.0;
            for (int k = 0; k < 6; ++k) G[r][c] += TL[r][k] * T[k][c];
        }
}

static void set_T(ofem_el *e) {
    memset(e->T, 0, sizeof(e->T));
    e->T[0][0] = e->T[1][1] = e->T[3][3] = e->T[4][4] = cos(e->angles);
    e->T[0][1] = e->T[3][4] = sin(e->angles);
    e->T[1][0] = e->T[4][3] = -sin(e->angles);
    e->T[2][2] = e->T[5][5] = 1.0;
}

/* FEMBody::computeNodeMapping, FEMBody.cpp:378-424 */
static void node_mapping(ofem *f, int nIB, int nFEM) {
    const double r = (nFEM - 1.0) / (nIB - 1.0);
    f->nib = nIB;
    f->par_el = (int *)calloc(nIB, sizeof(int));
    f->par_zeta = (double *)calloc(nIB, sizeof(double));
    f->par_el[0] = 0;
    f->par_zeta[0] = -1.0;
    for (int i = 1; i < nIB; ++i) {
        if (fmod(i * r, 1.0) == 0.0) {
            f->par_el[i] = (int)floor(i * r) - 1;
            f->par_zeta[i] = 1.0;
        } else {
            f->par_el[i] = (int)floor(i * r);
            f->par_zeta[i] = fmod(i * r, 1.0) * 2.0 - 1.0;
        }
    }
    for (int el = 0; el < f->nel; ++el) {
        ofem_el *e = &f->el[el];
        e->nchild = 0;
        for (int node = 0; node < nIB; ++node) {
            double n1 = -1 + 2.0 * ((node - 0.5) * r - el) / 1.0;
            double n2 = -1 + 2.0 * ((node + 0.5) * r - el) / 1.0;
            if ((n1 > -1.0 && n1 < 1.0) || (n2 > -1.0 && n2 < 1.0)) {
                if (n1 < -1.0) n1 = -1.0;
                if (n2 > 1.0) n2 = 1.0;
                e->cnode[e->nchild] = node;
                e->cz1[e->nchild] = n1;
                e->cz2[e->nchild] = n2;
                e->nchild++;
            }
        }
    }
}

/* FEMBody ctor (FEMBody.cpp:62-131) with FEMElement (FEMElement.cpp:61-102) and FEMNode
 * (FEMNode.cpp:49-68), 2D: depth = dh, horizontal angle 0 */
ofem *ora_fem_create(const double *start, double length, double height, double depth, double angle_deg,
                     int nEl, int clamped, double density, double E, double dh, double dt, double dm) {
    (void)depth;
    ofem *f = (ofem *)calloc(1, sizeof(ofem));
    f->nel = nEl;
    f->nnode = nEl + 1;
    f->n = (nEl + 1) * 3;
    f->bc = clamped ? 3 : 2;
    f->dh = dh;
    f->dt = dt;
    f->dm = dm;
    const double spacing = length / nEl;
    const double spacing_h = spacing * cos(angle_deg * L_PI / 180);
    f->nd = (ofem_node *)calloc(f->nnode, sizeof(ofem_node));
    for (int i = 0; i < f->nnode; ++i) {
        f->nd[i].pos0[0] = start[0] + i * spacing_h * cos(0.0 * L_PI / 180.0);
        f->nd[i].pos0[1] = start[1] + i * spacing * sin(angle_deg * L_PI / 180.0);
        f->nd[i].pos0[2] = start[2] + i * spacing_h * sin(0.0 * L_PI / 180.0);
        memcpy(f->nd[i].pos, f->nd[i].pos0, sizeof(f->nd[i].pos));
        f->nd[i].angles0 = angle_deg * L_PI / 180.0;
        f->nd[i].angles = f->nd[i].angles0;
    }
    const double dep = dh;   /* 2D: depth is one lattice spacing (FEMBody.cpp:103-105) */
    f->el = (ofem_el *)calloc(nEl, sizeof(ofem_el));
    for (int i = 0; i < nEl; ++i) {
        ofem_el *e = &f->el[i];
        e->length0 = spacing;
        e->length = spacing;
        e->area = height * dep;
        e->angles = angle_deg * L_PI / 180.0;
        e->I = dep * TH(height) / 12.0;
        e->E = E;
        e->density = density;
        for (int k = 0; k < 6; ++k) e->dofs[k] = i * 3 + k;
        set_T(e);
    }
    const int nIB = (int)floor(length / dh) + 1;
    node_mapping(f, nIB, f->nnode);
    size_t nn = (size_t)f->n * f->n;
    f->M = (double *)calloc(nn, sizeof(double));
    f->K = (double *)calloc(nn, sizeof(double));
    double **vecs[] = {&f->R, &f->F, &f->U, &f->U_n, &f->delU, &f->Udot, &f->Udot_n, &f->Udotdot, &f->Udotdot_n};
    for (int k = 0; k < 9; ++k) *vecs[k] = (double *)calloc(f->n, sizeof(double));
    return f;
}

void ora_fem_destroy(ofem *f) {
    if (!f) return;
    free(f->nd); free(f->el); free(f->par_el); free(f->par_zeta); free(f->M); free(f->K);
    free(f->R); free(f->F); free(f->U); free(f->U_n); free(f->delU); free(f->Udot); free(f->Udot_n);
    free(f->Udotdot); free(f->Udotdot_n);
    free(f);
}

/* FEMBody::updateFEMValues, FEMBody.cpp:297-325 */
static void update_values(ofem *f) {
    for (int n = 0; n < f->nnode; ++n) {
        for (int d = 0; d < 2; ++d) f->nd[n].pos[d] = f->nd[n].pos0[d] + f->U[n * 3 + d];
        f->nd[n].angles = f->nd[n].angles0 + f->U[n * 3 + 2];
    }
    for (int el = 0; el < f->nel; ++el) {
        double v[3];
        for (int a = 0; a < 3; ++a) v[a] = f->nd[el + 1].pos[a] - f->nd[el].pos[a];
        ofem_el *e = &f->el[el];
        e->angles = atan2(v[1], v[0]);
        double acc = 0.0;
        for (int a = 0; a < 3; ++a) acc += pow(v[a], 2);
        e->length = sqrt(acc);
        e->T[0][0] = e->T[1][1] = e->T[3][3] = e->T[4][4] = cos(e->angles);
        e->T[0][1] = e->T[3][4] = sin(e->angles);
        e->T[1][0] = e->T[4][3] = -sin(e->angles);
        e->T[2][2] = e->T[5][5] = 1.0;
    }
}

/* FEMElement::loadVector, FEMElement.cpp:133-173 */
static void load_vector(ofem *f, ofem_el *e, const double *mforce /* 3 per marker */, const double *meps) {
    const double forceScale = f->dm / SQ(f->dt);
    for (int c = 0; c < e->nchild; ++c) {
        const int ib = e->cnode[c];
        const double a = e->cz1[c], b = e->cz2[c], L = e->length;
        double s = meps[ib] * 1.0 * forceScale;
        double fs[2] = {mforce[3 * ib] * s, mforce[3 * ib + 1] * s};
        double F[2];
        F[0] = 0.0; F[0] += e->T[0][0] * fs[0]; F[0] += e->T[0][1] * fs[1];
        F[1] = 0.0; F[1] += e->T[1][0] * fs[0]; F[1] += e->T[1][1] * fs[1];
        double Rl[6];
        Rl[0] = F[0] * 0.5 * L * (0.5 * b - 0.5 * a + 0.25 * SQ(a) - 0.25 * SQ(b));
        Rl[1] = F[1] * 0.5 * L * (0.5 * b - 0.5 * a - SQ(a) * SQ(a) / 16.0 + SQ(b) * SQ(b) / 16.0 + 3.0 * SQ(a) / 8.0 - 3.0 * SQ(b) / 8.0);
        Rl[2] = F[1] * 0.5 * L * (L * (-SQ(a) * SQ(a) + SQ(b) * SQ(b)) / 32.0 - L * (-TH(a) + TH(b)) / 24.0 - L * (-SQ(a) + SQ(b)) / 16.0 + L * (b - a) / 8.0);
        Rl[3] = F[0] * 0.5 * L * (-0.25 * SQ(a) + 0.25 * SQ(b) + 0.5 * b - 0.5 * a);
        Rl[4] = F[1] * 0.5 * L * (0.5 * b - 0.5 * a + SQ(a) * SQ(a) / 16.0 - SQ(b) * SQ(b) / 16.0 - 3.0 * SQ(a) / 8.0 + 3.0 * SQ(b) / 8.0);
        Rl[5] = F[1] * 0.5 * L * (L * (-SQ(a) * SQ(a) + SQ(b) * SQ(b)) / 32.0 + L * (-TH(a) + TH(b)) / 24.0 - L * (-SQ(a) + SQ(b)) / 16.0 - L * (b - a) / 8.0);
        double Rg[6];
        mat6_T_vec(e->T, Rl, Rg);
        for (int k = 0; k < 6; ++k) f->R[e->dofs[k]] += Rg[k];
    }
}

static void assemble_mat(ofem *f, const ofem_el *e, const double G[6][6], double *Mg) {
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) Mg[(size_t)e->dofs[i] * f->n + e->dofs[j]] += G[i][j];
}

/* FEMElement::massMatrix, FEMElement.cpp:178-213 */
static void mass_matrix(ofem *f, ofem_el *e) {
    double Ml[6][6] = {{0}};
    const double C1 = e->density * e->area * e->length0 / 420.0, l0 = e->length0;
    Ml[0][0] = C1 * 140.0;
    Ml[0][3] = C1 * 70.0;
    Ml[1][1] = C1 * 156.0;
    Ml[1][2] = C1 * 22.0 * l0;
    Ml[1][4] = C1 * 54;
    Ml[1][5] = C1 * (-13.0 * l0);
    Ml[2][2] = C1 * 4.0 * SQ(l0);
    Ml[2][4] = C1 * 13.0 * l0;
    Ml[2][5] = C1 * (-3.0 * SQ(l0));
    Ml[3][3] = C1 * 140.0;
    Ml[4][4] = C1 * 156.0;
    Ml[4][5] = C1 * (-22.0 * l0);
    Ml[5][5] = C1 * 4.0 * SQ(l0);
    for (int i = 1; i < 6; i++)
        for (int j = 0; j < i; j++) Ml[i][j] = Ml[j][i];
    double G[6][6];
    congruence(e->T, Ml, G);
    assemble_mat(f, e, G, f->M);
}

/* FEMElement::stiffMatrix, FEMElement.cpp:218-268 */
static void stiff_matrix(ofem *f, ofem_el *e) {
    double Kl[6][6] = {{0}};
    const double E = e->E, A = e->area, I = e->I, l0 = e->length0;
    Kl[0][0] = E * A / l0;
    Kl[0][3] = -E * A / l0;
    Kl[1][1] = 12.0 * E * I / TH(l0);
    Kl[1][2] = 6.0 * E * I / SQ(l0);
    Kl[1][4] = -12.0 * E * I / TH(l0);
    Kl[1][5] = 6.0 * E * I / SQ(l0);
    Kl[2][2] = 4.0 * E * I / l0;
    Kl[2][4] = -6.0 * E * I / SQ(l0);
    Kl[2][5] = 2.0 * E * I / l0;
    Kl[3][3] = E * A / l0;
    Kl[4][4] = 12.0 * E * I / TH(l0);
    Kl[4][5] = -6.0 * E * I / SQ(l0);
    Kl[5][5] = 4.0 * E * I / l0;
    for (int i = 1; i < 6; i++)
        for (int j = 0; j < i; j++) Kl[i][j] = Kl[j][i];
    const double F0 = -e->F[0], V0 = e->F[4];
    Kl[0][1] += -V0 / l0;
    Kl[0][4] += V0 / l0;
    Kl[1][0] += -V0 / l0;
    Kl[1][1] += F0 / l0;
    Kl[1][3] += V0 / l0;
    Kl[1][4] += -F0 / l0;
    Kl[3][1] += V0 / l0;
    Kl[3][4] += -V0 / l0;
    Kl[4][0] += V0 / l0;
    Kl[4][1] += -F0 / l0;
    Kl[4][3] += -V0 / l0;
    Kl[4][4] += F0 / l0;
    double G[6][6];
    congruence(e->T, Kl, G);
    assemble_mat(f, e, G, f->K);
}

/* FEMElement::forceVector, FEMElement.cpp:273-302 */
static void force_vector(ofem *f, ofem_el *e, int id) {
    const double angle1 = f->nd[id].angles, angle2 = f->nd[id + 1].angles;
    const double u = (SQ(e->length) - SQ(e->length0)) / (e->length + e->length0);
    const double th1 = atan2(cos(e->angles) * sin(angle1) - sin(e->angles) * cos(angle1),
                             cos(e->angles) * cos(angle1) + sin(e->angles) * sin(angle1));
    const double th2 = atan2(cos(e->angles) * sin(angle2) - sin(e->angles) * cos(angle2),
                             cos(e->angles) * cos(angle2) + sin(e->angles) * sin(angle2));
    const double F0 = (e->E * e->area / e->length0) * u;
    const double M1 = (2 * e->E * e->I / e->length0) * (2.0 * th1 + th2);
    const double M2 = (2 * e->E * e->I / e->length0) * (th1 + 2.0 * th2);
    e->F[0] = -F0;
    e->F[1] = (1.0 / e->length0) * (M1 + M2);
    e->F[2] = M1;
    e->F[3] = F0;
    e->F[4] = -(1.0 / e->length0) * (M1 + M2);
    e->F[5] = M2;
    double Fg[6];
    mat6_T_vec(e->T, e->F, Fg);
    for (int k = 0; k < 6; ++k) f->F[e->dofs[k]] += Fg[k];
}

/* GridUtils::solveLinearSystem(A, b, BC) (GridUtils.cpp:356-395): the (n-BC) block at
 * offset BC solved through the transposed view; the first BC results are zero */
extern int ora_lu_solve_t(int n, double *a, double *b);
static int solve_bc(ofem *f) {
    const int n = f->n, bc = f->bc, m = n - bc;
    double *a = (double *)malloc(sizeof(double) * m * m);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) a[i * m + j] = f->K[(size_t)(bc + i) * n + (bc + j)];
    for (int i = 0; i < m; ++i) f->delU[bc + i] = f->F[bc + i];
    int r = ora_lu_solve_t(m, a, f->delU + bc);
    for (int i = 0; i < bc; ++i) f->delU[i] = 0.0;
    free(a);
    return r;
}

/* FEMBody::setNewmark, FEMBody.cpp:264-293 */
static void set_newmark(ofem *f) {
    const double Dt = f->dt;
    const double a0 = 1.0 / (NB_ALPHA * SQ(Dt)), a2 = 1.0 / (NB_ALPHA * Dt), a3 = 1.0 / (2.0 * NB_ALPHA) - 1.0;
    const int n = f->n;
    double *Meff = (double *)malloc(sizeof(double) * n), *MF = (double *)malloc(sizeof(double) * n);
    for (int i = 0; i < n; i++) Meff[i] = a0 * (f->U_n[i] - f->U[i]) + a2 * f->Udot[i] + a3 * f->Udotdot[i];
    for (int r = 0; r < n; ++r) {
        MF[r] = 0.0;
        for (int c = 0; c < n; ++c) MF[r] += f->M[(size_t)r * n + c] * Meff[c];
    }
    for (int i = 0; i < n; i++) {
        f->F[i] = f->R[i] - f->F[i] + MF[i];
        for (int j = 0; j < n; j++) f->K[(size_t)i * n + j] += a0 * f->M[(size_t)i * n + j];
    }
    free(Meff);
    free(MF);
}

/* FEMBody::newtonRaphsonIterator, :183-218 */
static int nr_iteration(ofem *f) {
    memset(f->F, 0, sizeof(double) * f->n);
    memset(f->M, 0, sizeof(double) * f->n * f->n);
    memset(f->K, 0, sizeof(double) * f->n * f->n);
    for (int el = 0; el < f->nel; ++el) {
        force_vector(f, &f->el[el], el);
        mass_matrix(f, &f->el[el]);
        stiff_matrix(f, &f->el[el]);
    }
    set_newmark(f);
    if (solve_bc(f)) return -1;
    for (int i = 0; i < f->n; i++) f->U[i] = f->U[i] + f->delU[i];
    update_values(f);
    return 0;
}

/* FEMBody::checkNRConvergence, :224-231 */
static double nr_residual(const ofem *f) {
    const double L = f->el[0].length0 * f->nel;
    double dp = 0.0;
    for (int i = 0; i < f->n; ++i) dp += f->delU[i] * f->delU[i];
    return sqrt(dp / (double)f->n) / L;
}

/* FEMBody::finishNewmark, :236-255 */
static void finish_newmark(ofem *f) {
    const double Dt = f->dt;
    const double a6 = 1.0 / (NB_ALPHA * SQ(Dt)), a7 = -1.0 / (NB_ALPHA * Dt), a8 = -(1.0 / (2.0 * NB_ALPHA) - 1.0);
    const double a9 = Dt * (1.0 - NB_DELTA), a10 = NB_DELTA * Dt;
    for (int i = 0; i < f->n; i++) {
        double st = a6 * (f->U[i] - f->U_n[i]) + a7 * f->Udot[i] + a8 * f->Udotdot[i];
        f->Udot[i] = f->Udot[i] + a9 * f->Udotdot[i] + a10 * st;
        f->Udotdot[i] = st;
    }
}

/* FEMElement::shapeFuns, FEMElement.cpp:109-128 */
static void shape(const ofem_el *e, const double *v, double zeta, double out[2]) {
    const double h = (zeta + 1.0) / 2.0;
    const double N0 = 1.0 - h, N1 = 1.0 - 3.0 * SQ(h) + 2.0 * TH(h), N2 = (h - 2.0 * SQ(h) + TH(h)) * e->length;
    const double N3 = h, N4 = 3.0 * SQ(h) - 2.0 * TH(h), N5 = (-SQ(h) + TH(h)) * e->length;
    out[0] = v[0] * N0 + v[3] * N3;
    out[1] = v[1] * N1 + v[2] * N2 + v[4] * N4 + v[5] * N5;
}

/* FEMBody::updateIBMarkers, :329-370 */
static void update_markers(ofem *f, double *mpos, const double *mpos0, double *mvel, double *mvel_km1) {
    for (int node = 0; node < f->nib; ++node) {
        const ofem_el *e = &f->el[f->par_el[node]];
        const double zeta = f->par_zeta[node];
        double dU[6], dUd[6], lU[6], lUd[6], sU[2], sUd[2];
        for (int k = 0; k < 6; ++k) {
            dU[k] = f->U[e->dofs[k]];
            dUd[k] = f->Udot[e->dofs[k]];
        }
        mat6_vec(e->T, dU, lU);
        mat6_vec(e->T, dUd, lUd);
        shape(e, lU, zeta, sU);
        shape(e, lUd, zeta, sUd);
        /* transpose(Tsub) * v with Tsub = T[0..1][0..1] */
        double gU[2], gUd[2];
        for (int r = 0; r < 2; ++r) {
            gU[r] = 0.0;
            gUd[r] = 0.0;
            for (int c = 0; c < 2; ++c) {
                gU[r] += e->T[c][r] * sU[c];
                gUd[r] += e->T[c][r] * sUd[c];
            }
        }
        const double sc = f->dt / f->dh;
        for (int d = 0; d < 2; ++d) {
            const double ud = gUd[d] * sc;
            mpos[3 * node + d] = mpos0[3 * node + d] + gU[d];
            mvel_km1[3 * node + d] = mvel[3 * node + d];
            mvel[3 * node + d] = RELAX * ud + (1.0 - RELAX) * mvel_km1[3 * node + d];
        }
    }
}

/* FEMBody::dynamicFEM, FEMBody.cpp:136-179 */
int ora_fem_dynamic(ofem *f, const double *mforce, const double *meps, double *mpos, const double *mpos0,
                    double *mvel, double *mvel_km1) {
    memcpy(f->U, f->U_n, sizeof(double) * f->n);
    memcpy(f->Udot, f->Udot_n, sizeof(double) * f->n);
    memcpy(f->Udotdot, f->Udotdot_n, sizeof(double) * f->n);
    update_values(f);
    memset(f->R, 0, sizeof(double) * f->n);
    for (int el = 0; el < f->nel; ++el) load_vector(f, &f->el[el], mforce, meps);
    f->it = 0;
    do {
        if (nr_iteration(f)) return -1;
        f->res = nr_residual(f);
        f->it++;
    } while (f->res > 1e-10 && f->it < 20);
    finish_newmark(f);
    update_markers(f, mpos, mpos0, mvel, mvel_km1);
    return 0;
}

/* end of a step's sub-iterations: U_n = U etc. (ObjectManager_ops_ibm.cpp:150-158) */
void ora_fem_commit(ofem *f) {
    memcpy(f->U_n, f->U, sizeof(double) * f->n);
    memcpy(f->Udot_n, f->Udot, sizeof(double) * f->n);
    memcpy(f->Udotdot_n, f->Udotdot, sizeof(double) * f->n);
}

void ora_fem_state(const ofem *f, double *U, double *Udot, int *it, double *res) {
    if (U) memcpy(U, f->U, sizeof(double) * f->n);
    if (Udot) memcpy(Udot, f->Udot, sizeof(double) * f->n);
    if (it) *it = f->it;
    if (res) *res = f->res;
}
int ora_fem_ndof(const ofem *f) { return f->n; }

/* Test access: assemble the tangent stiffness and mass matrices at the current state (the
 * loops of newtonRaphsonIterator before Newmark), full n x n row-major. */
void ora_fem_assemble(ofem *f, double *K, double *M) {
    memset(f->F, 0, sizeof(double) * f->n);
    memset(f->M, 0, sizeof(double) * f->n * f->n);
    memset(f->K, 0, sizeof(double) * f->n * f->n);
    for (int el = 0; el < f->nel; ++el) {
        force_vector(f, &f->el[el], el);
        mass_matrix(f, &f->el[el]);
        stiff_matrix(f, &f->el[el]);
    }
    memcpy(K, f->K, sizeof(double) * f->n * f->n);
    memcpy(M, f->M, sizeof(double) * f->n * f->n);
}

void ora_fem_set_state(ofem *f, const double *U, const double *Udot, const double *Udotdot) {
    memcpy(f->U_n, U, sizeof(double) * f->n);
    memcpy(f->Udot_n, Udot, sizeof(double) * f->n);
    memcpy(f->Udotdot_n, Udotdot, sizeof(double) * f->n);
    memcpy(f->U, U, sizeof(double) * f->n);
    update_values(f);
}

/* One time step of dynamicFEM with a given external load vector R (no IB markers). */
int ora_fem_step_load(ofem *f, const double *R) {
    memcpy(f->U, f->U_n, sizeof(double) * f->n);
    memcpy(f->Udot, f->Udot_n, sizeof(double) * f->n);
    memcpy(f->Udotdot, f->Udotdot_n, sizeof(double) * f->n);
    update_values(f);
    memcpy(f->R, R, sizeof(double) * f->n);
    f->it = 0;
    do {
        if (nr_iteration(f)) return -1;
        f->res = nr_residual(f);
        f->it++;
    } while (f->res > 1e-10 && f->it < 20);
    finish_newmark(f);
    ora_fem_commit(f);
    return f->it;
}

/* Test access: the markers of the current state through updateIBMarkers (FEMBody.cpp:329-370):
 * positions pos0 + N(zeta) U of each marker's parent element, velocities relaxed from Udot. */
void ora_fem_markers(ofem *f, const double *mpos0, double *mpos, double *mvel, double *mvel_km1) {
    update_markers(f, mpos, mpos0, mvel, mvel_km1);
}
int ora_fem_nib(const ofem *f) { return f->nib; }
