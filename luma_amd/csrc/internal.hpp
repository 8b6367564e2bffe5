// internal.hpp -- the grid handle behind the C ABI and helpers shared by the host units.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lumagpu.h"
#include "launch.hpp"

namespace lumagpu {
struct IbmSystem;
}

struct lumagpu_grid {
    lumagpu_grid_desc d;
    int Q = 0, D = 0;
    size_t ncell = 0, ps = 0, elem = 0;
    size_t melem = 8;                 // bytes per element of the macro, force and side arrays
    void *f = nullptr;
    void *f2 = nullptr;               // second population array of the two-step (temporal-blocking) kernel
    int fuse_shape = -1, fuse_zc = 0; // its tile shape (FusedShape, -1: off) and z planes per block
    uint8_t *flags = nullptr;
    uint32_t *plain = nullptr;        // plain-site bitmap, rebuilt with every flag upload
    int host_forced = 0;              // some host flag carries the force bit (at the last upload)
    // refined sites that sites other than TL2F read, and their initial populations
    void *stale = nullptr;            // [Q][nstale] storage elements
    int *stale_idx = nullptr;         // [ncell] index into stale, -1 elsewhere
    int *stale_cells = nullptr;       // [nstale] their cells
    std::vector<int> stale_list;      // host copy of stale_cells
    // large labelled grids: the sites the plain-site kernel leaves to the labelled one
    int *mlist = nullptr;
    int nmlist = -1;                  // < 0: one labelled launch over the grid
    std::vector<uint8_t> mlist_types; // the site types mlist was built from
    // the captured two-step graph of this hierarchy (root grids), reused while its key holds
    int ib_fold = 0;                  // the step kernel does this step's IB spread
    hipGraphExec_t gexec = nullptr;
    std::vector<uint8_t> gkey;
    int nstale = 0;
    double *force = nullptr;
    double *uin = nullptr;
    double *rho = nullptr, *u = nullptr;
    double *partials = nullptr;
    double *stage = nullptr;       // LUMA-layout staging (f AoS sized)
    size_t stage_bytes = 0;
    double *moments_out = nullptr;
    hipStream_t stream = nullptr;     // the stream this grid's work goes to: the root's, in a hierarchy
    hipStream_t own_stream = nullptr; // created with the grid
    int device = 0;                   // the GPU the grid was created on
    hipEvent_t ev_done = nullptr;     // recorded on this grid's stream after its sub-steps
    hipEvent_t ev_ready = nullptr;    // recorded before a child reads this grid's populations
    lumagpu_grid *parent = nullptr;   // GridObj::parentGrid (non-owning)
    std::vector<lumagpu_grid *> kids; // GridObj::subGrid (non-owning), stepped in this order
    lumagpu::ChildRef *d_kids = nullptr;   // device table of the kids for the step kernels
    double *xr = nullptr;             // ExtrapolateRight side buffers, 2 x Q*M*K (this step's, the next's)
    int has_outlet = 0;
    int xr_sel = 0;                   // which of the two is this step's
    int xr_ready = 0;                 // this step's holds the sources (written by the previous step)
    int xr_fold = 0;                  // the step kernels write the next step's (every outlet row's x = N-3
                                      // site is one they step)
    int nedge = 0;                    // regularised velocity sites on edges and corners
    int *edge = nullptr;              // [4*nedge]: cells, packed normals, neighbour x+n, x+2n
    double *erho = nullptr;           // [8*nedge] neighbour rho, u of the current step (edge/pressure)
    double d_omega = 0.0;             // tangent form: seeded derivative of omega
    int gravity_on = 0, gravity_dir = 0;   // L_GRAVITY_ON, L_GRAVITY_DIRECTION
    double gravity = 0.0;             // GridObj::gravity (lattice units)
    int all_forced = 0;               // this step reset gravity: every site carries force_xyz
    int sweep1_zero = 0;              // this step's IB reset zeroes force_xyz (sweep 1 sees no force)
    double ramp_time = 0.0;           // L_VELOCITY_RAMP in dimensionless time (0: off)
    double pressure_drho = 0.0;       // GridUnits::pd2dlbm(L_PRESSURE_DELTA)
    double re = 0.0, re_ramp = 0.0;   // L_RE and L_REYNOLDS_RAMP in dimensionless time (0: off)
    int masked = 0;
    int zslab = 0;                    // planes 0 and K-1 are halo planes of a z-sharded box
    int has_state = 0;
    int half = 0;                     // between lumagpu_stream_bc_macro and lumagpu_force_collide_swap
    int half_sub = 0;                 // the subcycle of that split step
    int store_macro = 0;
    int parity = 0;
    int t = 0;
    std::vector<uint8_t> hflags;
    std::vector<double> xpos, ypos, zpos;   // GridObj::XPos/YPos/ZPos (IB support search)
    lumagpu::IbmSystem *ibm = nullptr;      // IB bodies owned by this grid
};

namespace lumagpu {

// thread-local message behind lumagpu_last_error()
int fail(int code, const std::string &msg);

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(LUMAGPU_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
    } while (0)

inline StepArgs make_args(const lumagpu_grid *g) {
    StepArgs a;
    std::memset(&a, 0, sizeof a);   // padding included: the bytes key the cached step graph
    a.f = g->f;
    a.ps = g->ps;
    a.nx = g->d.N;
    a.ny = g->d.M;
    a.nz = g->d.K;
    a.flags = g->flags;
    a.plain = g->plain;
    a.force = g->force;
    a.uin = g->uin;
    a.rho_out = g->rho;
    a.u_out = g->u;
    a.ncell = g->ncell;
    a.omega = g->d.omega;
    a.d_omega = g->d_omega;
    a.rho0 = g->d.rho0;
    a.rho_in = g->d.rho_in;
    // tmpDensity = L_RHOIN; tmpDensity += pd2dlbm(L_PRESSURE_DELTA) (GridObj_ops_lbm_optimised.cpp:317, :343-345)
    a.rho_p = g->d.rho_in;
    a.rho_p += g->pressure_drho;
    const double cs = 1.0 / std::sqrt(3.0);    // LUMA/src/stdafx.cpp:152
    a.cs2 = cs * cs;
    a.den2 = 2.0 * (cs * cs) * (cs * cs);
    a.om_hi = float(g->d.omega);
    a.om_lo = float(g->d.omega - double(a.om_hi));
    a.zoff = g->zslab ? 1 : 0;
    a.regularised = g->d.regularised;
    a.fz_compat = g->d.fz_compat;
    a.all_forced = g->all_forced;
    // GridUtils::getVelocityRampCoefficient (GridUtils.cpp:1808-1816) at (t+1) dt (regularised
    // sites, GridObj_ops_lbm_optimised.cpp:326) and t dt (forced-equilibrium sources, :258-266)
    auto ramp = [g](double t) {
        if (g->ramp_time > 0.0 && t <= g->ramp_time)
            return (1.0 - std::cos(3.14159265358979323846 * t / g->ramp_time)) / 2.0;
        return 1.0;
    };
    a.ramp_on = g->ramp_time > 0.0;
    a.ramp_reg = ramp((g->t + 1) * g->d.dt);
    a.ramp_feq = ramp(g->t * g->d.dt);
    a.sweep1_zero = g->sweep1_zero;
    if (g->parent) {
        const lumagpu_grid *p = g->parent;
        a.pf = p->f;
        a.pps = p->ps;
        a.pnx = p->d.N;
        a.pny = p->d.M;
        a.pnz = p->d.K;
        a.pparity = p->parity;
        a.pflags = p->flags;
        a.clx = g->d.coarse_lims[0];
        a.cly = g->d.coarse_lims[2];
        a.clz = g->d.coarse_lims[4];
    }
    a.kids = g->d_kids;
    a.nkids = (int)g->kids.size();
    a.xr = g->xr ? g->xr + (size_t)g->xr_sel * g->Q * g->d.M * g->d.K * (g->melem / sizeof(double)) : nullptr;
    a.xr_next = (g->xr && g->xr_fold) ? g->xr + (size_t)(g->xr_sel ^ 1) * g->Q * g->d.M * g->d.K * (g->melem / sizeof(double))
                                       : nullptr;
    a.stale = g->stale;
    a.stale_idx = g->stale_idx;
    a.nstale = g->nstale;
    return a;
}

// g and every grid below it, parents before their children
template <class G, class Fn>
void for_each_level(G *g, Fn &&fn) {
    fn(g);
    for (lumagpu_grid *k : g->kids) for_each_level(static_cast<G *>(k), fn);
}

// IB step between the stream/macro and the force/collide halves (ObjectManager::ibm_apply);
// defined in ibm.hip
// fold_ok: the spread may be left to the step kernel that follows (see StepArgs::ib_site_of)
int ibm_apply(lumagpu_grid *g, bool fold_ok = false);
// the step kernel's spread arguments when the last IB application left the spread to it
void ibm_fold_args(const lumagpu_grid *g, StepArgs &a);
void ibm_destroy(lumagpu_grid *g);
int ensure_flags(lumagpu_grid *g);
// append what the IB step of g passes to its kernels (ibm.hip), for the step graph's key
void ibm_graph_key(const lumagpu_grid *g, std::vector<uint8_t> &key);
// copy the host flag bytes to the device and rebuild the plain-site bitmap (api.hip)
int upload_flags(lumagpu_grid *g);
// copy the initial populations of the refined sites other sites read (api.hip); the grid is
// in the post-local layout (parity 0) right after its state was set
int capture_stale(lumagpu_grid *g);
// refined regions per grid (GridObj::subGrid)
constexpr size_t kMaxKids = 16;
// rewrite the device table of g's refined regions (api.hip)
int upload_kids(lumagpu_grid *g);
// after a whole-field force reset: no site holds an IB force, the next spread flags its sites
void ibm_forget_flags(lumagpu_grid *g);
// the grid carries a flexible (FEM) body, whose step runs host work
bool ibm_has_flexible(const lumagpu_grid *g);
// markers of all IB bodies on the grid
int ibm_marker_count(const lumagpu_grid *g);
// one LBM_multi_opt of the grid (children first), as lumagpu_step; defined in api.hip
int step_grid(lumagpu_grid *g, int subcycle);
// ExtrapolateRight side buffer from the grid's current populations; defined in api.hip
hipError_t outlet_prepass(const lumagpu_grid *g);
// edge/corner and pressure sites: phase 0 the neighbours' sweep-1 rho, u (erho), phase 1 the
// sites' step; defined in api.hip
hipError_t launch_edges(const lumagpu_grid *g, const StepArgs &a, int phase);
// omega of the coming step under a Reynolds ramp, and its derivative with respect to the
// omega after the ramp; defined in api.hip
void update_reynolds(lumagpu_grid *g);
double reynolds_omega_weight(const lumagpu_grid *g);
// k_site_macro at a list of cells (every cell when cells is null); defined in ibm.hip
hipError_t launch_sites(const lumagpu_grid *g, const StepArgs &a, bool odd, int n, const int *cells, int ib_force);

}  // namespace lumagpu
