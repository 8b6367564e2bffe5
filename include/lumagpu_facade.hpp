// lumagpu_facade.hpp -- C++ host side over the C ABI in lumagpu.h, mirroring LUMA's own class
// interface for the time step: GridObj (LUMA/inc/GridObj.h:44-188) and ObjectManager
// (LUMA/inc/ObjectManager.h:141-194), with the same member names and argument meaning, so code
// written against the reference's API drives the GPU step.  Header-only; link libluma_amd.so.
//
// Differences from the reference, by construction: the configuration of definitions.h is a
// runtime descriptor (lumagpu_grid_desc) instead of compile-time macros; the host arrays
// (f, u, rho, ...) are mirrors, filled from the device by sync() and pushed by LBM_initGrid();
// errors go through L_ERROR's behaviour (message, then exit(12345), stdafx.h:119-149) unless a
// handler is installed with set_error_handler (tests throw instead).
#ifndef LUMAGPU_FACADE_HPP
#define LUMAGPU_FACADE_HPP

#include <array>
#include <cstdlib>
#include <functional>
#include <iostream>
#include <string>
#include <utility>
#include <vector>

#include "lumagpu.h"

namespace luma {

// eIOFlag (LUMA/inc/Enumerations.h)
enum eIOFlag { eWrite, eRead };

using ErrorHandler = std::function<void(const std::string &)>;
inline ErrorHandler &error_handler() {
    static ErrorHandler h = [](const std::string &msg) {
        std::cerr << "Error: " << msg << std::endl;
        std::exit(12345);
    };
    return h;
}
inline void set_error_handler(ErrorHandler h) { error_handler() = std::move(h); }
// L_ERROR on a failed ABI call
inline void check(int rc) {
    if (rc != LUMAGPU_OK) error_handler()(lumagpu_last_error());
}

class GridObj {
  public:
    // GridObj(int level) + the allocation of LBM_initGrid / LBM_initSubGrid: one grid of the
    // hierarchy from its descriptor; the host arrays are sized in LUMA's layouts
    explicit GridObj(const lumagpu_grid_desc &desc) : d_(desc) {
        check(lumagpu_grid_create(&d_, &h_));
        level = d_.level;
        N_lim = d_.N;
        M_lim = d_.M;
        K_lim = d_.K;
        dh = d_.dh;
        dt = d_.dt;
        dm = d_.dm;
        omega = d_.omega;
        const size_t n = (size_t)N_lim * M_lim * K_lim;
        f.assign(n * d_.Q, 0.0);
        u.assign(n * d_.dims, 0.0);
        rho.assign(n, d_.rho_in);
        LatTyp.assign(n, 1);   // eFluid
        XPos.assign(N_lim, 0.0);
        YPos.assign(M_lim, 0.0);
        ZPos.assign(K_lim, 0.0);
        ux_in.assign(M_lim, 0.0);
        uy_in.assign(M_lim, 0.0);
        uz_in.assign(M_lim, 0.0);
    }
    GridObj(const GridObj &) = delete;
    GridObj &operator=(const GridObj &) = delete;
    // ~GridObj (GridObj.cpp:39-43): sub-grids are owned and destroyed first
    ~GridObj() {
        for (GridObj *c : subGrid) delete c;
        if (h_) lumagpu_grid_destroy(h_);
    }

    // public state of GridObj.h:54-128 (host mirrors)
    int level = 0, region_number = 0, t = 0;
    int N_lim = 0, M_lim = 0, K_lim = 0;
    double dh = 1.0, dt = 1.0, dm = 1.0, omega = 1.0;
    std::vector<double> f, u, rho, XPos, YPos, ZPos, ux_in, uy_in, uz_in;
    std::vector<int> LatTyp;
    std::vector<GridObj *> subGrid;
    GridObj *parentGrid = nullptr;

    // LBM_initGrid (GridObj_init_grids.cpp:155-384) after the host arrays are filled: labels
    // and positions, inlet profile, and f = feq(rho, u)
    void LBM_initGrid() {
        check(lumagpu_set_labels(h_, LatTyp.data(), XPos.data(), YPos.data(), ZPos.data()));
        check(lumagpu_set_inlet(h_, ux_in.data(), uy_in.data(), uz_in.data()));
        check(lumagpu_init_equilibrium(h_, rho.data(), u.data()));
        check(lumagpu_set_store_macro(h_, 1));
    }
    // _LBM_initGetInletProfileFromFile (GridObj_init_grids.cpp:1233-1312): ux_in/uy_in/uz_in from
    // a profile file (L_USE_INLET_PROFILE); path replaces ./input/inlet_profile.in
    void LBM_initGetInletProfileFromFile(const std::string &path) {
        check(lumagpu_read_inlet_profile(h_, path.c_str()));
    }
    // LBM_initVelocity with L_INIT_VELOCITY_FROM_FILE (GridObj_init_grids.cpp:43-78): u from a
    // velocity file; path replaces ./input/initial_velocity.in
    void LBM_initVelocityFromFile(const std::string &path) { check(lumagpu_read_initial_velocity(h_, path.c_str(), u.data())); }
    // start from LUMA-layout post-collision populations already in f (a restart)
    void LBM_initFromF() {
        check(lumagpu_upload_f(h_, f.data()));
        check(lumagpu_upload_macro(h_, rho.data(), u.data()));
    }
    // LBM_addSubGrid (GridObj.cpp:92-118); the child is owned by this grid from now on
    void LBM_addSubGrid(GridObj *child) {
        check(lumagpu_grid_add_child(h_, child->h_));
        child->parentGrid = this;
        subGrid.push_back(child);
    }
    // LBM_multi_opt (GridObj_ops_lbm_optimised.cpp:36-193): children twice, then this grid
    void LBM_multi_opt(int subcycle = 0) {
        check(lumagpu_step(h_, subcycle));
        advance(1);
    }
    // LBM_multi_opt in the halves its two sweeps form, for a caller that runs the IB stages
    // (ObjectManager::ibm_interpolate ... ibm_updateMacroscopic) or code of its own in between;
    // the sub-grids must have stepped first (their own LBM_multi_opt(0) and (1)).  Sweep 1:
    // _LBM_resetForces on a level holding IB bodies, stream, regularised BCs, _LBM_macro_opt
    // (GridObj_ops_lbm_optimised.cpp:53-107); sweep 2: _LBM_forceGrid_opt, _LBM_collide_opt, the
    // swap and ++t (:125-172).  The host mirrors are not touched (sync() fills them).
    void LBM_stream_bc_macro(int subcycle = 0) { check(lumagpu_stream_bc_macro(h_, subcycle)); }
    void LBM_force_collide_swap() {
        check(lumagpu_force_collide_swap(h_));
        t += 1;
    }
    // _LBM_resetForces (:1327-1337): force_xyz = 0, or the gravity component with L_GRAVITY_ON
    void _LBM_resetForces() { check(lumagpu_reset_forces(h_)); }
    // nsteps LBM_multi_opt calls of the top-level grid (replayed as a graph when possible)
    void LBM_multi_run(int nsteps) {
        check(lumagpu_run(h_, nsteps));
        advance(nsteps);
    }
    // host mirrors <- device: f (post-collision), rho, u, for this grid and its sub-grids
    void sync() {
        check(lumagpu_download_f(h_, f.data()));
        check(lumagpu_download_macro(h_, rho.data(), u.data()));
        for (GridObj *c : subGrid) c->sync();
    }
    // io_restart (GridObj_ops_io.cpp:406-590); path replaces LUMA's fixed file names
    void io_restart(eIOFlag flag, const std::string &path) {
        if (flag == eWrite) {
            check(lumagpu_write_restart(h_, path.c_str(), level != 0, region_number));
            for (GridObj *c : subGrid) c->io_restart(eWrite, path);
        } else {
            check(lumagpu_read_restart(h_, path.c_str(), region_number));
            for (GridObj *c : subGrid) c->io_restart(eRead, path);
        }
    }
    // io_lite (GridObj_ops_io.cpp:706-810) into a directory, LUMA's file name
    void io_lite(double tval, const std::string &dir, int rank = 0, int precision = 8) {
        const std::string name = dir + "/io_lite.Lev" + std::to_string(level) + ".Reg" +
                                 std::to_string(region_number) + ".Rnk" + std::to_string(rank) + "." +
                                 std::to_string((int)tval) + ".dat";
        check(lumagpu_io_lite(h_, name.c_str(), rank, precision));
        for (GridObj *c : subGrid) c->io_lite(tval, dir, rank, precision);
    }
    lumagpu_grid *handle() const { return h_; }
    const lumagpu_grid_desc &desc() const { return d_; }

  private:
    void advance(int n) {
        t += n;
        for (GridObj *c : subGrid) c->advance(2 * n);
    }
    lumagpu_grid_desc d_;
    lumagpu_grid *h_ = nullptr;
};

// ObjectManager (ObjectManager.h:141-194): the immersed bodies of a grid hierarchy.  Their
// step (ibm_apply, ObjectManager_ops_ibm.cpp:33-54) runs inside LBM_multi_opt on the device.
class ObjectManager {
  public:
    static ObjectManager *getInstance(GridObj *g = nullptr) {
        ObjectManager *&m = instance();
        if (!m) m = new ObjectManager(g);
        return m;
    }
    static void destroyInstance() {
        delete instance();
        instance() = nullptr;
    }

    // io_readInGeomConfig (ObjectManager_ops_io.cpp:322-800) over the hierarchy
    void io_readInGeomConfig(const std::string &path, const std::array<double, 3> &shift = {0.0, 0.0, 0.0}) {
        std::vector<lumagpu_grid *> hs;
        for (GridObj *x = _Grids; x; x = x->subGrid.empty() ? nullptr : x->subGrid[0]) hs.push_back(x->handle());
        int n = 0;
        check(lumagpu_ibm_read_geometry(hs.data(), (int)hs.size(), path.c_str(), shift.data(), &n));
    }
    // io_restart (ObjectManager_ops_io.cpp:136-290) for the bodies of every grid of the
    // hierarchy; path replaces LUMA's restart_IBBody_Rnk<r>.out.  After a read, call
    // ibm_initialise, as main_lbm.cpp:307-318 does.
    void io_restart(eIOFlag flag, const std::string &path) {
        std::vector<lumagpu_grid *> hs;
        for (GridObj *x = _Grids; x; x = x->subGrid.empty() ? nullptr : x->subGrid[0]) hs.push_back(x->handle());
        if (flag == eWrite) check(lumagpu_ibm_write_restart(hs.data(), (int)hs.size(), path.c_str()));
        else check(lumagpu_ibm_read_restart(hs.data(), (int)hs.size(), path.c_str()));
    }
    // a rigid (or flexible) Body<IBMarker> from marker positions (3 per marker) on grid g
    int addBody(GridObj &g, const std::vector<double> &pos, bool flexible = false) {
        int b = -1;
        check(lumagpu_ibm_add_body(g.handle(), (int)(pos.size() / 3), pos.data(), flexible ? 1 : 0, &b));
        return b;
    }
    // ibm_initialise (ObjectManager_ops_ibm.cpp:190-246) on every grid that holds bodies
    void ibm_initialise() {
        for (GridObj *x = _Grids; x; x = x->subGrid.empty() ? nullptr : x->subGrid[0]) {
            int nb = 0;
            check(lumagpu_ibm_count(x->handle(), &nb, -1, nullptr));
            if (nb > 0) check(lumagpu_ibm_initialise(x->handle()));
        }
    }
    // The IB step of the bodies on grid level `level` (one region per level), between a grid's
    // LBM_stream_bc_macro and LBM_force_collide_swap.  ibm_apply (ObjectManager_ops_ibm.cpp:33-54,
    // with doSubIterate = true as LBM_multi_opt calls it): all stages, the FEM move and the FSI
    // sub-iterations of flexible bodies.  ibm_interpolate (:498-591), ibm_computeForce (:598-614),
    // ibm_spread (:621-689), ibm_updateMacroscopic (:696-766): one stage each, for rigid bodies.
    void ibm_apply(GridObj *g) { check(lumagpu_ibm_apply(g->handle())); }
    void ibm_interpolate(int level) { check(lumagpu_ibm_interpolate(grid_at(level)->handle())); }
    void ibm_computeForce(int level) { check(lumagpu_ibm_compute_force(grid_at(level)->handle())); }
    void ibm_spread(int level) { check(lumagpu_ibm_spread(grid_at(level)->handle())); }
    void ibm_updateMacroscopic(int level) { check(lumagpu_ibm_update_macro(grid_at(level)->handle())); }
    // FP64 sum of a body's marker forces (the terms of io_writeLiftDrag)
    std::array<double, 3> bodyForce(GridObj &g, int body) {
        std::array<double, 3> out{};
        check(lumagpu_ibm_body_force(g.handle(), body, out.data()));
        return out;
    }
    // io_writeLiftDrag (ObjectManager_ops_io.cpp:64-122) for one body of grid g
    void io_writeLiftDrag(GridObj &g, int body, const std::string &path, int precision = 8) {
        check(lumagpu_ibm_write_ld(g.handle(), path.c_str(), body, _Grids->t, _Grids->t * _Grids->dt, precision));
    }
    // io_writeTipPositions (ObjectManager_ops_io.cpp:1220-1245) for one flexible body of grid g
    void io_writeTipPositions(GridObj &g, int body, const std::string &path, int precision = 10) {
        check(lumagpu_ibm_write_tip(g.handle(), path.c_str(), body, _Grids->t, _Grids->t * _Grids->dt, precision));
    }

  private:
    explicit ObjectManager(GridObj *g) : _Grids(g) {}
    GridObj *grid_at(int level) {
        for (GridObj *x = _Grids; x; x = x->subGrid.empty() ? nullptr : x->subGrid[0])
            if (x->level == level) return x;
        error_handler()("no grid on level " + std::to_string(level));
        return _Grids;
    }
    static ObjectManager *&instance() {
        static ObjectManager *m = nullptr;
        return m;
    }
    GridObj *_Grids;
};

}  // namespace luma

#endif
