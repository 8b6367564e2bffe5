/* fem_oracle.h -- TEST INFRASTRUCTURE ONLY: FP64 restatement of LUMA's beam FEM. */
#ifndef FEM_ORACLE_H
#define FEM_ORACLE_H

typedef struct {
    double length0, length, area, E, density, angles, I;
    double T[6][6];
    double F[6];
    int dofs[6];
    int nchild;
    int cnode[64];
    double cz1[64], cz2[64];
} ofem_el;

typedef struct {
    double pos0[3], pos[3], angles0, angles;
} ofem_node;

typedef struct ofem {
    int nel, nnode, n, bc;
    ofem_el *el;
    ofem_node *nd;
    double *M, *K, *R, *F, *U, *U_n, *delU, *Udot, *Udot_n, *Udotdot, *Udotdot_n;
    int nib;
    int *par_el;
    double *par_zeta;
    double dt, dh, dm;
    int it;
    double res;
} ofem;

ofem *ora_fem_create(const double *start, double length, double height, double depth, double angle_deg,
                     int nEl, int clamped, double density, double E, double dh, double dt, double dm);
void ora_fem_destroy(ofem *f);
int ora_fem_dynamic(ofem *f, const double *mforce, const double *meps, double *mpos, const double *mpos0,
                    double *mvel, double *mvel_km1);
void ora_fem_commit(ofem *f);
void ora_fem_state(const ofem *f, double *U, double *Udot, int *it, double *res);
int ora_fem_ndof(const ofem *f);
/* test access (pinning tests) */
void ora_fem_assemble(ofem *f, double *K, double *M);
void ora_fem_set_state(ofem *f, const double *U, const double *Udot, const double *Udotdot);
int ora_fem_step_load(ofem *f, const double *R);
void ora_fem_markers(ofem *f, const double *mpos0, double *mpos, double *mvel, double *mvel_km1);
int ora_fem_nib(const ofem *f);

#endif
