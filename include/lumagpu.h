/*
 * lumagpu.h -- C ABI of the MI355X (gfx950) lattice-Boltzmann time step.
 *
 * The reference has no FFI layer: its boundary is the C++ class API of GridObj and
 * ObjectManager (LUMA/inc/GridObj.h:44-188, LUMA/inc/ObjectManager.h:141-194).  Each entry
 * point below names the reference member it replaces.  All arrays crossing this boundary
 * use LUMA's own host layouts:
 *   site id       id = k + j*K + i*K*M            (LUMA/inc/IVector.h:94-101)
 *   populations   f[v + Q*id], LUMA's c_opt order (LUMA/src/stdafx.cpp:81-153)
 *   vectors       u[d + D*id], force_xyz[d + D*id]; scalars rho[id]
 *   labels        int eType values                (LUMA/inc/Enumerations.h:84-96)
 * Pointers are borrowed for the duration of the call only.  Every function returns 0 on
 * success or a negative LUMAGPU_E* code; lumagpu_last_error() then holds a message
 * (thread-local).  Calls on one grid are ordered on that grid's HIP stream; functions that
 * return host data synchronise that stream.  A grid handle is used by one host thread.
 */
#ifndef LUMAGPU_H
#define LUMAGPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LUMAGPU_ABI_VERSION 1

#define LUMAGPU_OK 0
#define LUMAGPU_EINVAL (-1)       /* bad argument or descriptor */
#define LUMAGPU_EHIP (-2)         /* HIP runtime error */
#define LUMAGPU_ENOMEM (-3)       /* device allocation failed */
#define LUMAGPU_EUNSUPPORTED (-4) /* configuration not supported by this build */
#define LUMAGPU_ESTATE (-5)       /* call out of order (e.g. step before state upload) */
#define LUMAGPU_EPHYS (-6)        /* physically invalid input (omega >= 2, ...) */

/* storage/arithmetic form of the populations */
#define LUMAGPU_FP64 0   /* plain FP64 storage and arithmetic, LUMA's operation order */
#define LUMAGPU_FP32S 1  /* FP32 storage of f - w*rho0, FP32 collision */
#define LUMAGPU_FP16S 2  /* FP16 storage of f - w*rho0, FP32 collision */
#define LUMAGPU_FP64T 3  /* tangent form: FP64 values with their derivative along one parameter
                            (dual numbers, 16 B per population); values equal LUMAGPU_FP64's */
#define LUMAGPU_FP32T 4  /* shifted tangent form: LUMAGPU_FP32S on float dual numbers (8 B per
                            population); macro arrays, force field and IB in FP64 duals */

/* moments returned by lumagpu_moments, all FP64 */
#define LUMAGPU_MOMENT_MASS 0
#define LUMAGPU_MOMENT_MOMX 1
#define LUMAGPU_MOMENT_MOMY 2
#define LUMAGPU_MOMENT_MOMZ 3
#define LUMAGPU_MOMENT_KE 4
#define LUMAGPU_NUM_MOMENTS 5

typedef struct lumagpu_grid lumagpu_grid;

/* Replaces the compile-time configuration of LUMA/inc/definitions.h for one grid
 * (GridObj(int level), GridObj_init_grids.cpp:155-384 / :390-536). */
typedef struct lumagpu_grid_desc {
    int dims;             /* L_DIMS: 2 or 3 */
    int Q;                /* L_NUM_VELS: 9 (D2Q9) or 19 (D3Q19) */
    int N, M, K;          /* N_lim, M_lim, K_lim (K = 1 in 2D) */
    int precision;        /* LUMAGPU_FP64 / FP32S / FP16S / FP64T / FP32T */
    int regularised;      /* L_REGULARISED_BOUNDARIES */
    int forced;           /* L_IBM_ON || L_GRAVITY_ON: force_xyz enters macro and collision */
    int fz_compat;        /* 3D: add F_z/2 to x-momentum as GridObj_ops_lbm_optimised.cpp:833 */
    int level;            /* refinement level */
    double omega;         /* GridObj::omega */
    double rho_in;        /* L_RHOIN */
    double rho0;          /* reference density of the shifted forms: must be 1 for them */
    double xmax, ymax, zmax;   /* GridManager::global_edges[e?Max][0] of level 0 */
    double wall_thk[6];        /* L_WALL_THICKNESS_{LEFT,RIGHT,BOTTOM,TOP,FRONT,BACK} */
    int coarse_lims[6];        /* CoarseLimsX/Y/Z[min,max] on the parent (level > 0) */
    double dh;                 /* GridObj::dh (IB support search and distances) */
    double origin[3];          /* global_edges[e?Min] of this grid's level and region */
    double dt;                 /* GridObj::dt (FEM time step, sub-iteration residual scale) */
    double dm;                 /* GridObj::dm (FEM load scaling dm/dt^2) */
} lumagpu_grid_desc;

/* ---- lifetime ------------------------------------------------------------------------ */
/* GridObj::GridObj(int level) + LBM_initGrid allocation (GridObj.cpp:51-67) */
int lumagpu_grid_create(const lumagpu_grid_desc *desc, lumagpu_grid **out);
/* GridObj::~GridObj (GridObj.cpp:39-43); children must be destroyed by the caller first */
int lumagpu_grid_destroy(lumagpu_grid *g);

/* GridObj::LBM_addSubGrid / LBM_initSubGrid (GridObj.cpp:92-118, GridObj_init_grids.cpp:390-536):
 * attach `child` (level + 1, extents 2*(coarse_lims max - min + 1)) as a refined region of
 * `parent` (GridObj::subGrid; up to 16 per grid, disjoint, else LUMAGPU_EINVAL /
 * LUMAGPU_EUNSUPPORTED).  Stepping the parent then steps each child twice first, in the order
 * they were added, as LBM_multi_opt does. */
int lumagpu_grid_add_child(lumagpu_grid *parent, lumagpu_grid *child);

/* ---- state upload / download --------------------------------------------------------- */
/* LatTyp + XPos/YPos/ZPos (GridObj_init_grids.cpp:234-273; labels of LBM_initBoundLab) */
int lumagpu_set_labels(lumagpu_grid *g, const int *lattyp, const double *xpos, const double *ypos,
                       const double *zpos);
/* ux_in/uy_in/uz_in, one value per j (GridObj_init_grids.cpp:1322-1360) */
int lumagpu_set_inlet(lumagpu_grid *g, const double *ux_in, const double *uy_in, const double *uz_in);
/* GridObj::_LBM_initGetInletProfileFromFile (GridObj_init_grids.cpp:1233-1312, L_USE_INLET_PROFILE):
 * ux_in/uy_in/uz_in from a file of x y z ux uy uz rows, interpolated along y at the grid's YPos
 * with the reference's end extrapolation; path replaces ./input/inlet_profile.in.  Values are
 * taken as the file gives them, as the reference does. */
int lumagpu_read_inlet_profile(lumagpu_grid *g, const char *path);
/* GridObj::LBM_initVelocity with L_INIT_VELOCITY_FROM_FILE (GridObj_init_grids.cpp:43-78): fills
 * u[d + D*id] (LUMA layout, for lumagpu_init_equilibrium) from the first N*M*K rows of an
 * x y z ux uy uz file in dimensionless units (ud2ulbm: u dt/dh), each row at the voxel that
 * encloses its position; path replaces ./input/initial_velocity.in.  Fewer rows than sites is
 * an error, as in the reference. */
int lumagpu_read_initial_velocity(const lumagpu_grid *g, const char *path, double *u);
/* rho and u, then f = feq(rho, u) (GridObj_init_grids.cpp:315-333) */
int lumagpu_init_equilibrium(lumagpu_grid *g, const double *rho, const double *u);
/* GridObj::f after a step (post-collision populations) and the macro arrays rho, u */
int lumagpu_upload_f(lumagpu_grid *g, const double *f);
int lumagpu_upload_macro(lumagpu_grid *g, const double *rho, const double *u);
int lumagpu_download_f(lumagpu_grid *g, double *f);
int lumagpu_download_macro(lumagpu_grid *g, double *rho, double *u);
/* Tangent forms only: derivatives of f (post-collision, LUMA layout), rho and u along the
 * seeded direction; any output may be NULL.  No counterpart in LUMA (the tangent form is this
 * build's forward-mode derivative of LBM_multi_opt). */
int lumagpu_download_tangent(lumagpu_grid *g, double *df, double *drho, double *du);
/* FP64 tangent form: seeds d(omega) = d_omega for this grid's collision (GridObj::omega,
 * GridObj_init_grids.cpp:339-356).  Grids of a hierarchy are seeded one by one; a refined
 * level's omega follows its parent's as omega_f = 1/(2(1/omega_c - 1/2) + 1/2) (:527), so
 * its seed is d_omega_c * domega_f/domega_c. */
int lumagpu_tangent_seed_omega(lumagpu_grid *g, double d_omega);
/* GridObj::force_xyz as left by the caller (body force); requires desc.forced */
int lumagpu_set_force(lumagpu_grid *g, const double *force_xyz);

/* ---- time step ----------------------------------------------------------------------- */
/* GridObj::LBM_multi_opt(int subcycle) (GridObj.h:188, GridObj_ops_lbm_optimised.cpp:36-193):
 * the children's two sub-steps first, then this grid's step with its IB bodies' step
 * (ObjectManager::ibm_apply) between the stream/macro and force/collide halves. */
int lumagpu_step(lumagpu_grid *g, int subcycle);
/* The step of one grid in the two halves LBM_multi_opt's sweeps form, so that a caller can run
 * the IB stages (below), or code of its own, in between; the children must have stepped
 * already (lumagpu_step(child, 0), lumagpu_step(child, 1)).  lumagpu_stream_bc_macro is
 * sweep 1 (GridObj_ops_lbm_optimised.cpp:53-107: _LBM_resetForces on a level holding IB bodies,
 * then stream, regularised BCs and _LBM_macro_opt): afterwards rho and u hold sweep 1's values
 * at every fluid, eTransitionToFiner and regularised velocity-face site (edge/corner velocity
 * and pressure sites receive theirs in the second half).  lumagpu_force_collide_swap is sweep 2
 * (:125-172: _LBM_forceGrid_opt, _LBM_collide_opt, f.swap(fNew), ++t); it streams again in
 * the same pass (the populations stay in place until then).  stream_bc_macro, the IB stages
 * and force_collide_swap in that order equal lumagpu_step bit for bit; lumagpu_step and
 * lumagpu_run are refused while a split step is open. */
int lumagpu_stream_bc_macro(lumagpu_grid *g, int subcycle);
int lumagpu_force_collide_swap(lumagpu_grid *g);
/* GridObj::_LBM_resetForces (GridObj_ops_lbm_optimised.cpp:1327-1337): force_xyz = 0 at every
 * site, or with L_GRAVITY_ON the gravity component rho * gravity * refinement_ratio (the other
 * components left as they are); requires desc.forced.  lumagpu_stream_bc_macro already does
 * this on a level holding IB bodies; this entry is for callers that sub-iterate themselves. */
int lumagpu_reset_forces(lumagpu_grid *g);
/* nsteps consecutive lumagpu_step(g, 0) calls */
int lumagpu_run(lumagpu_grid *g, int nsteps);
/* as lumagpu_run, timing with HIP events on the grid's stream: total wall time of the
 * nsteps and the mean duration of one step kernel (event pair around each launch) */
int lumagpu_run_timed(lumagpu_grid *g, int nsteps, double *ms_total, double *ms_kernel_mean);
/* write rho and u from every step (LUMA keeps them in GridObj::rho/u); off by default */
int lumagpu_set_store_macro(lumagpu_grid *g, int on);
/* L_GRAVITY_ON with L_GRAVITY_DIRECTION (definitions.h:114-117): gravity = GridObj::gravity in
 * lattice units (fd2flbm(L_GRAVITY_FORCE), GridObj_init_grids.cpp:185).  Each step of a grid
 * that holds IB bodies then starts with _LBM_resetForces (GridObj_ops_lbm_optimised.cpp:53-57,
 * :1327-1337): the gravity component of force_xyz at every site becomes rho * gravity *
 * refinement_ratio (refinement_ratio = 1/2^level) from the previous step's rho, and, as in the
 * reference, the other components are not reset, so IB forces in them add up over the steps.
 * LUMA resets nothing on levels without IB bodies: there the force stays at its initial
 * rho * gravity * refinement_ratio (:291-297), which lumagpu_set_force gives.  Turns the macro
 * store on.  Flexible bodies are refused at the step. */
int lumagpu_set_gravity(lumagpu_grid *g, double gravity, int direction);
/* L_VELOCITY_RAMP (definitions.h:244): the inlet profile is scaled by
 * GridUtils::getVelocityRampCoefficient (GridUtils.cpp:1808-1816), (1 - cos(pi t / T)) / 2 for
 * t <= T and 1 after, with t = (GridObj::t + 1) dt at regularised velocity sites (:326) and
 * t = GridObj::t dt at forced-equilibrium sources (:258-266; these take the profile of the
 * pulling site's row, as the reference writes it).  ramp_time = T in dimensionless time; 0 turns
 * it off.  Needs desc.dt.  lumagpu_run steps eagerly while a ramp is set. */
int lumagpu_set_velocity_ramp(lumagpu_grid *g, double ramp_time);
/* Temporal blocking of LBM_multi_opt (GridObj_ops_lbm_optimised.cpp:36-193) for a periodic D3Q19
 * box without labels, force, refinement or bodies: lumagpu_run and lumagpu_run_timed then
 * advance such a grid two steps per kernel launch (even and odd AA step fused through LDS, a
 * TX x TY tile of columns per block marching `zchunk` z planes; shape 0 = 64 x 8, 1 = 32 x 16,
 * 2 = 128 x 4, plus 8 for two alternating LDS copies (FP32- and FP16-shifted only), plus 16
 * (with 8) for loads one plane ahead, plus 32 (with 8) for LDS copies holding only the tile's
 * read windows; 24 with zchunk 128 is the fastest at 256^3; N and M whole numbers of tiles),
 * bit for bit the same populations and macro arrays as
 * two single steps in FP64 and FP32-shifted storage, and within one unit in the last place of
 * the stored half per value in FP16-shifted.  Allocates a second population array (the kernel
 * reads one and writes the other); shape < 0 frees it and turns the two-step kernel off.
 * Grids that stop qualifying (labels or a force set later) step one step at a time.  Long
 * lumagpu_run calls replay a captured graph of two two-step launches. */
int lumagpu_set_temporal_blocking(lumagpu_grid *g, int shape, int zchunk);
/* L_REYNOLDS_RAMP (definitions.h:204): at the start of every step of this grid and of each grid
 * below it (LBM_multi_opt, GridObj_ops_lbm_optimised.cpp:39-42, before the children's
 * sub-steps), omega is reset by _LBM_updateReynolds (:1313-1321) to
 * 1 / (nu / cs^2 + 0.5) with nu = (dt / Re) / dh^2 (GridUnits::nud2nulbm, GridUnits.h:128-131)
 * and Re = re * getReynoldsRampCoefficient((t + 1) dt) (GridUtils.cpp:1825-1833), the
 * coefficient being 1 - cos(pi t / T) for t <= T, which rises to 2 at T, and 1 after, as the
 * reference computes it.  re = L_RE, ramp_time = T in dimensionless time (0 turns it off),
 * with each level's own desc.dt and desc.dh.  The omega seed of the tangent form stays
 * constant; the adjoint refuses a grid with a Reynolds ramp. */
int lumagpu_set_reynolds_ramp(lumagpu_grid *g, double re, double ramp_time);
/* GridObj::omega as the last step used it (the descriptor's until a Reynolds ramp changes it) */
int lumagpu_get_omega(const lumagpu_grid *g, double *omega);
/* L_PRESSURE_DELTA (definitions.h:245): regularised pressure sites hold
 * rho = L_RHOIN + GridUnits::pd2dlbm(p) = L_RHOIN + (p dh dt^2 / dm) / cs^2
 * (GridObj_ops_lbm_optimised.cpp:317, :343-345; GridUnits.h:177-180), p dimensionless.  The
 * initial density and the forced-equilibrium sources keep L_RHOIN.  Needs desc.dh, dt, dm. */
int lumagpu_set_pressure_delta(lumagpu_grid *g, double p);
/* completed steps on this grid (GridObj::t) */
int lumagpu_get_time(const lumagpu_grid *g, int *t);
int lumagpu_sync(lumagpu_grid *g);

/* ---- reductions (FP64, deterministic) --------------------------------------------------
 * mass, momentum and kinetic energy of the post-collision state; LUMA has no counterpart
 * (its logs hold per-marker forces only, ObjectManager_ops_io.cpp:65-128). */
int lumagpu_moments(lumagpu_grid *g, double out[LUMAGPU_NUM_MOMENTS]);
/* lumagpu_moments without any allocation: no counterpart in the reference; for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Writes the same LUMAGPU_NUM_MOMENTS doubles, bit for bit, into
 * out_dev (device memory of the caller) asynchronously on the grid's stream (lumagpu_sync
 * waits).  Its per-block partial sums live in a static device array of the library, one per
 * GPU, sized for at most 2^18 row blocks (ceil(N/256) * M * K), ~10 MB; larger grids are
 * refused with LUMAGPU_EINVAL.  Calls that overlap on one GPU share that array and race. */
int lumagpu_moments_noalloc(lumagpu_grid *g, double *out_dev);

/* ---- immersed boundary (ObjectManager, LUMA/src/ObjectManager_ops_ibm.cpp) ------------
 * Bodies live on one grid; their step (interpolate, force, spread, macroscopic update,
 * :33-54) runs inside lumagpu_step between the stream/macro and force/collide halves.  The
 * grid must be created with desc.forced = 1 (L_IBM_ON).  Positions use the units of the
 * grid's XPos/YPos/ZPos (lumagpu_set_labels / lumagpu_set_positions) and desc.dh. */
/* L_UNIVERSAL_EPSILON_CALC (definitions.h:220): ibm_findEpsilon solves one RKPM system over
 * every marker of the grid's bodies, in body order (ibm_universalEpsilonGather /
 * ibm_universalEpsilonScatter, ObjectManager_ops_ibm.cpp:775-786, :905-909, :1067-1141), for
 * bodies whose supports overlap.  Set before lumagpu_ibm_initialise; 0 = per body (default).
 * The adjoint chains the positions through the same joint system. */
int lumagpu_ibm_set_universal_epsilon(lumagpu_grid *g, int on);
/* Body<IBMarker>::addMarker (Body.h:144-255) + IBMarker ctor (IBMarker.cpp:56-86):
 * a body of n markers at pos[3*m + a]; epsilon 0, ds 1, markerVel 0 until initialised. */
int lumagpu_ibm_add_body(lumagpu_grid *g, int nmarkers, const double *pos, int flexible, int *body);
/* IBBody filament with an FEMBody (IBBody.cpp:240-258, FEMBody.cpp:62-131), 2D: floor(L/dh) + 1
 * markers from `start` at angle_deg, n_elements corotational beams, clamped or supported base,
 * density and Young's modulus as LUMA's geometry config takes them (ObjectManager_ops_io.cpp:455-567).
 * Such a body makes every step run LUMA's FSI sub-iterations (ObjectManager_ops_ibm.cpp:111-185). */
int lumagpu_ibm_add_filament(lumagpu_grid *g, const double *start, double length, double height, double depth,
                             double angle_deg, int n_elements, int clamped, double density, double youngs,
                             int *body);
/* The IB stages of one grid's bodies, between lumagpu_stream_bc_macro and
 * lumagpu_force_collide_swap (ESTATE elsewhere; no-ops on a grid without bodies):
 *   lumagpu_ibm_apply            ObjectManager::ibm_apply(g, true) (ObjectManager_ops_ibm.cpp:33-54):
 *                                all of the below, then for flexible bodies ibm_moveBodies and the
 *                                FSI sub-iterations (ibm_subIterate, :111-185) -- what lumagpu_step runs
 *   lumagpu_ibm_interpolate      ibm_interpolate (:498-591): rho and rho u at the markers from sweep 1's
 *                                values at their supports (recomputed from the populations)
 *   lumagpu_ibm_compute_force    ibm_computeForce (:598-614): F = 2 (rho u - rho U) per marker
 *   lumagpu_ibm_spread           ibm_spread (:621-689) into force_xyz, in LUMA's accumulation order
 *   lumagpu_ibm_update_macro     ibm_updateMacroscopic (:696-766): rho, u of the supports with F/2
 * interpolate, compute_force, spread, update_macro in that order equal lumagpu_ibm_apply for rigid
 * bodies (flexible bodies move only in lumagpu_ibm_apply). */
int lumagpu_ibm_apply(lumagpu_grid *g);
int lumagpu_ibm_interpolate(lumagpu_grid *g);
int lumagpu_ibm_compute_force(lumagpu_grid *g);
int lumagpu_ibm_spread(lumagpu_grid *g);
int lumagpu_ibm_update_macro(lumagpu_grid *g);
/* sub-iterations and residual of the last step's FSI loop (ObjectManager::timeav_subResidual) */
int lumagpu_ibm_subiterations(const lumagpu_grid *g, int *iterations, double *residual);
/* FSI sub-iteration counts of every step since lumagpu_ibm_initialise (the per-step values of
 * ObjectManager::timeav_subIterations): *n receives the number of steps, up to max are copied.
 * And a prescribed schedule: the next n steps run exactly counts[k] sub-iterations (1..10)
 * instead of stopping on the residual (n = 0 clears).  No counterpart in LUMA: pinning lets
 * finite differences in a parameter follow the schedule a derivative run took. */
int lumagpu_ibm_subiteration_log(const lumagpu_grid *g, int *counts, int max, int *n);
int lumagpu_ibm_pin_subiterations(lumagpu_grid *g, const int *counts, int n);
/* ObjectManager::ibm_initialise (:190-246): support stencils, ds and epsilon of every body */
int lumagpu_ibm_initialise(lumagpu_grid *g);
/* ObjectManager::io_readInGeomConfig (ObjectManager_ops_io.cpp:322-800): builds the IBM bodies
 * of a geometry.config file on the grids (matched by level, region 0): FROM_FILE point clouds
 * (io_readInCloud :812-1049, one marker per voxel; the cloud file is read from the geometry
 * file's directory, as LUMA reads both from ./input/; rigid only), CIRCLE_SPHERE, SQUARE_CUBE,
 * PLATE (3D) and FILAMENT_ARRAY (FLEXIBLE filaments get their FEM beam, 2D).  shift[3] is the
 * wall-thickness shift LUMA adds to the prefabs when the left/bottom/front walls are solid.
 * Bounce-back and BFL bodies return LUMAGPU_EUNSUPPORTED.  *nbodies receives the count built. */
int lumagpu_ibm_read_geometry(lumagpu_grid *const *grids, int ngrids, const char *path, const double *shift,
                              int *nbodies);
/* number of IB bodies on the grid and, for body >= 0, the markers of that body (iBody.size(),
 * iBody[b].markers.size()); either output may be NULL */
int lumagpu_ibm_count(const lumagpu_grid *g, int *nbodies, int body, int *nmarkers);
/* IBMarker::markerVel for every marker of a body (vel[3*m + a]) */
int lumagpu_ibm_set_velocity(lumagpu_grid *g, int body, const double *vel);
/* IBMarker::epsilon for every marker of a body (replaces the ibm_findEpsilon result) */
int lumagpu_ibm_set_epsilon(lumagpu_grid *g, int body, const double *eps);
/* Marker state after the last IB step, 16 doubles per marker: position[3], markerVel[3],
 * force_xyz[3], interpMom[3], interpRho, epsilon, ds, number of support sites. */
int lumagpu_ibm_get_markers(lumagpu_grid *g, int body, double *rec);
/* FP64 sum over the body's markers of force_xyz * epsilon * volDepth * ds (the per-marker
 * terms of ObjectManager::io_writeLiftDrag, ObjectManager_ops_io.cpp:98-121) */
int lumagpu_ibm_body_force(lumagpu_grid *g, int body, double out[3]);
/* tangent form: the same sum and its derivative */
int lumagpu_ibm_body_force_tangent(lumagpu_grid *g, int body, double out[3], double dout[3]);
/* tangent form: derivatives of the 16 lumagpu_ibm_get_markers values of each marker (the
 * support count's is 0) */
int lumagpu_ibm_get_marker_tangents(lumagpu_grid *g, int body, double *rec);
/* tangent form: seeds the derivative direction -- d(Young's modulus)/ds = dE for every element
 * of a flexible filament (FEMElement::E, FEMElement.cpp:61-104), so every derivative the grid
 * reports is d/ds along E -> E + s dE.  Nothing else is seeded. */
int lumagpu_ibm_seed_youngs(lumagpu_grid *g, int body, double dE);
/* XPos/YPos/ZPos of a grid that has no labels to set */
int lumagpu_set_positions(lumagpu_grid *g, const double *xpos, const double *ypos, const double *zpos);

/* ---- one GPU's share of a sharded box ---------------------------------------------------
 * MpiManager's Cartesian decomposition (MpiManager.cpp:103-226) with its exchange after every
 * grid step (GridObj_ops_lbm_optimised.cpp:186-191), for a z-sharded D3Q19 periodic box: with
 * on = 1 the grid's first and last z planes become halo planes, the step updates planes
 * 1..K-2 only, and around every odd (neighbour-access) step the five populations that cross
 * each face are exchanged -- here with the grid itself as both neighbours, the copy a rank
 * would replace with RCCL send/receive of the same contiguous planes. */
int lumagpu_set_zslab(lumagpu_grid *g, int on);

/* ---- LUMA's on-disk formats --------------------------------------------------------------
 * Restart lines of this grid (GridObj::io_restart, GridObj_ops_io.cpp:406-590): level, region,
 * X, Y, Z, u in dimensionless units (GridUnits::ulbm2ud), rho, and the time-scaled
 * non-equilibrium populations ((f - feq) omega)/(feq dt); written at full precision (LUMA
 * writes 6 significant digits).  append = 0 starts the file (level 0), 1 appends (sub-grids).
 * Reading takes the lines of this grid's level and region, places them by position and sets
 * f = feq (1 + dt fneq / omega) from the file's rho and u (LUMA evaluates that equilibrium on
 * the level-0 grid's arrays for every level; here each grid uses its own).  Needs desc.dh/dt. */
int lumagpu_write_restart(lumagpu_grid *g, const char *path, int append, int region);
int lumagpu_read_restart(lumagpu_grid *g, const char *path, int region);
/* GridObj::io_lite (GridObj_ops_io.cpp:706-810): one row per site, k-j-i order: rank, eType,
 * X, Y, Z, rho, u (uz = 0.000000 in 2D), f[Q], fNew[Q], fixed format with the given number of
 * decimals (L_OUTPUT_PRECISION).  This build keeps one population array, so the fNew block
 * repeats the post-collision f. */
int lumagpu_io_lite(lumagpu_grid *g, const char *path, int rank, int precision);
/* ObjectManager::io_writeLiftDrag (ObjectManager_ops_io.cpp:64-122) for one body: appends
 * "t<TAB>time" and, per marker, force * epsilon * volDepth * ds * dm dh / dt^2 (/dh in 2D) per
 * dimension; the header line when t == 0. */
int lumagpu_ibm_write_ld(lumagpu_grid *g, const char *path, int body, int t, double time, int precision);
/* ObjectManager::io_writeTipPositions (ObjectManager_ops_io.cpp:1220-1245, L_WRITE_TIP_POSITIONS)
 * for one flexible body: appends "t<TAB>time<TAB>TipX<TAB>TipY<TAB>TipZ" of its last marker
 * (precision = L_OUTPUT_PRECISION, the stream's default format), the header line when t == 0;
 * time = t * the level-0 grid's dt as the reference writes it; the reference names the file
 * Body_<id>_TipPositions.out. */
int lumagpu_ibm_write_tip(lumagpu_grid *g, const char *path, int body, int t, double time, int precision);
/* ObjectManager::io_restart (ObjectManager_ops_io.cpp:136-290) for the bodies of a hierarchy's
 * grids, in grid order: the body count, then per body "\t/\t" n "\t/\t" and its marker
 * positions (plus position0 for flexible bodies), at full precision (LUMA writes 6 digits).
 * Reading checks the body and marker counts against the bodies already built and applies
 * the positions only when the whole file is accepted; the bodies then need
 * lumagpu_ibm_initialise (main_lbm.cpp:307-318).  LUMA's reader takes the marker values from
 * the wrong '/' segment and sets nothing; this one reads each segment whole.  Flexible bodies
 * return LUMAGPU_EUNSUPPORTED on read (their FEM state is not in the file). */
int lumagpu_ibm_write_restart(lumagpu_grid *const *grids, int ngrids, const char *path);
int lumagpu_ibm_read_restart(lumagpu_grid *const *grids, int ngrids, const char *path);

/* ---- device checkpoints --------------------------------------------------------------------
 * A copy in HBM of a grid's populations, storage layout, macro arrays and step counter
 * (the in-memory counterpart of io_restart, GridObj_ops_io.cpp:406-590, with no rounding):
 * restoring it resumes the run bit for bit.  IB marker state is not included (rigid bodies
 * do not change; flexible bodies are refused).  Free with lumagpu_checkpoint_free. */
typedef struct lumagpu_checkpoint lumagpu_checkpoint;
int lumagpu_checkpoint_save(lumagpu_grid *g, lumagpu_checkpoint **out);
int lumagpu_checkpoint_restore(lumagpu_grid *g, const lumagpu_checkpoint *ck);
int lumagpu_checkpoint_free(lumagpu_checkpoint *ck);

/* ---- discrete adjoint (the differentiation form) ------------------------------------------
 * No LUMA counterpart.  Objective: J = (1/T) sum over the next T steps of dir . (the IB body
 * force, summed over all bodies of the grid: the per-marker terms LUMA writes to
 * Body_<id>_LD, ObjectManager_ops_io.cpp:98-121).  lumagpu_adj_run advances the grid T steps
 * (as T lumagpu_step calls; it is left at the end of them), keeping a copy of the populations
 * every checkpoint_every steps; with backward = 1 it then runs the discrete adjoint of those
 * steps, recomputing each segment from its checkpoint, and returns
 *   res->J, res->dJ_domega (relaxation frequency; under a Reynolds ramp the omega after the
 *   ramp, which L_RE sets for every step), res->dJ_dinlet (derivative with respect to
 *   a common scale s of the inlet profile, u_in -> s u_in, at s = 1),
 *   dJ_dpos[3*m + a] (marker positions of all bodies in marker order, through the delta
 *   weights, epsilon and ds; NULL to skip) and lambda0 (dJ/df of the initial post-collision
 *   populations in LUMA's layout f[v + Q*id]; NULL to skip).
 * Covered: FP64 and FP32-shifted forms, one grid level, fluid / solid / forced-equilibrium
 * velocity / regularised velocity (faces, edges, corners) / regularised pressure face /
 * ExtrapolateRight sites, body force, rigid IB bodies (supports on fluid sites), a velocity
 * ramp, a Reynolds ramp.  Refused (LUMAGPU_EUNSUPPORTED): refinement, z-slabs, gravity,
 * flexible bodies.  Memory: ceil(T/K) + K + 1
 * population arrays plus three adjoint arrays (and, while lambda0 is written, one array of
 * Q FP64 values per cell for LUMA's layout). */
typedef struct lumagpu_adj_result {
    double J;
    double dJ_domega;
    double dJ_dinlet;
} lumagpu_adj_result;
int lumagpu_adj_run(lumagpu_grid *g, int nsteps, int checkpoint_every, const double *dir, int backward,
                    lumagpu_adj_result *res, double *dJ_dpos, double *lambda0);

/* ---- diagnostics --------------------------------------------------------------------- */
const char *lumagpu_last_error(void);   /* replaces L_ERROR's log line (stdafx.h:119-149) */
int lumagpu_abi_version(void);
/* bytes of population storage per cell and the population array size in bytes */
int lumagpu_footprint(const lumagpu_grid *g, size_t *bytes_per_cell, size_t *f_bytes);

#ifdef __cplusplus
}
#endif
#endif /* LUMAGPU_H */
