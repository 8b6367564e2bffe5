// api.hip -- the C ABI (include/lumagpu.h): grid handles, state transfer and the step driver.
//
// A grid owns one population array in the AA in-place layout, a per-cell flag byte (only
// for labelled grids), optional FP64 macro arrays and force field, and one HIP stream at the
// default priority.  `parity` records which layout the population array is in:
//   0  post-local: A[opp v][x] = f_v(x)   (after upload, init, or an even step)
//   1  post-neighbour: A[v][x + c_v] = f_v(x), or A[opp v][x] where x + c_v is solid
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "internal.hpp"

using namespace lumagpu;

namespace lumagpu {
namespace {
thread_local std::string g_last_error;
}
int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}
}  // namespace lumagpu

namespace {
hipError_t launch_step(const lumagpu_grid *g, const LaunchSpec &ls, const StepArgs &a) {
    switch (g->d.precision) {
        case LUMAGPU_FP64: return launch_step_f64(ls, a, g->stream);
        case LUMAGPU_FP32S: return launch_step_f32s(ls, a, g->stream);
        case LUMAGPU_FP64T: return launch_step_dual(ls, a, g->stream);
        case LUMAGPU_FP32T: return launch_step_dual32(ls, a, g->stream);
        default: return launch_step_f16s(ls, a, g->stream);
    }
}

hipError_t launch_aux(const lumagpu_grid *g, int op, const double *p0, const double *p1, double *out) {
    const StepArgs a = make_args(g);
    switch (g->d.precision) {
        case LUMAGPU_FP64: return launch_aux_f64(op, g->Q, a, p0, p1, out, g->parity, g->masked, g->stream);
        case LUMAGPU_FP32S: return launch_aux_f32s(op, g->Q, a, p0, p1, out, g->parity, g->masked, g->stream);
        case LUMAGPU_FP64T: return launch_aux_dual(op, g->Q, a, p0, p1, out, g->parity, g->masked, g->stream);
        case LUMAGPU_FP32T: return launch_aux_dual32(op, g->Q, a, p0, p1, out, g->parity, g->masked, g->stream);
        default: return launch_aux_f16s(op, g->Q, a, p0, p1, out, g->parity, g->masked, g->stream);
    }
}

int ensure_stage(lumagpu_grid *g, size_t bytes) {
    if (g->stage_bytes >= bytes) return 0;
    if (g->stage) {
        HIP_TRY(hipStreamSynchronize(g->stream));
        HIP_TRY(hipFree(g->stage));
        g->stage = nullptr;
        g->stage_bytes = 0;
    }
    if (hipMalloc(&g->stage, bytes) != hipSuccess) return fail(LUMAGPU_ENOMEM, "staging allocation failed");
    g->stage_bytes = bytes;
    return 0;
}

int check_grid(const lumagpu_grid *g) {
    if (!g) return fail(LUMAGPU_EINVAL, "null grid handle");
    return 0;
}

// GridUtils::isWithinDomainWall, LUMA/src/GridUtils.cpp:1369-1430: returns edge count and
// the code of the last inward normal found (1..6 = +x,-x,+y,-y,+z,-z).
// GridUtils::isWithinDomainWall (GridUtils.cpp:1369-1430): the number of walls the position
// lies in, the code of the last one (1..6 = +x,-x,+y,-y,+z,-z inward normal) and the combined
// inward vector n.
int wall_normal(const lumagpu_grid_desc &d, double px, double py, double pz, int *code, int n[3]) {
    const double *th = d.wall_thk;
    int cnt = 0;
    *code = 0;
    n[0] = n[1] = n[2] = 0;
    if (px > 0.0 && px < th[0]) { *code = 1; n[0] = 1; cnt++; }
    if (px < d.xmax && px > d.xmax - th[1]) { *code = 2; n[0] = -1; cnt++; }
    if (py > 0.0 && py < th[2]) { *code = 3; n[1] = 1; cnt++; }
    if (py < d.ymax && py > d.ymax - th[3]) { *code = 4; n[1] = -1; cnt++; }
    if (d.dims == 3) {
        if (pz > 0.0 && pz < th[4]) { *code = 5; n[2] = 1; cnt++; }
        if (pz < d.zmax && pz > d.zmax - th[5]) { *code = 6; n[2] = -1; cnt++; }
    }
    return cnt;
}

hipError_t launch_zslab(const lumagpu_grid *g, const StepArgs &a, int phase) {
    switch (g->d.precision) {
        case LUMAGPU_FP64: return launch_zslab_f64(a, phase, g->stream);
        case LUMAGPU_FP32S: return launch_zslab_f32s(a, phase, g->stream);
        case LUMAGPU_FP16S: return launch_zslab_f16s(a, phase, g->stream);
        default: return hipErrorInvalidValue;   // refused by lumagpu_set_zslab
    }
}

}  // namespace

namespace lumagpu {
hipError_t launch_edges(const lumagpu_grid *g, const StepArgs &a, int phase) {
    EdgeArgs e{};
    e.n = g->nedge;
    e.cell = g->edge;
    e.nrm = g->edge + g->nedge;
    e.n1 = g->edge + 2 * g->nedge;
    e.n2 = g->edge + 3 * g->nedge;
    e.erho = g->erho;
    const int store = (g->store_macro && a.rho_out) ? 1 : 0;
    const bool odd = g->parity == 0;
    switch (g->d.precision) {
        case LUMAGPU_FP64: return launch_edges_f64(g->Q, a, e, odd, phase, store, g->stream);
        case LUMAGPU_FP32S: return launch_edges_f32s(g->Q, a, e, odd, phase, store, g->stream);
        case LUMAGPU_FP64T: return launch_edges_dual(g->Q, a, e, odd, phase, store, g->stream);
        case LUMAGPU_FP32T: return launch_edges_dual32(g->Q, a, e, odd, phase, store, g->stream);
        default: return launch_edges_f16s(g->Q, a, e, odd, phase, store, g->stream);
    }
}

// _LBM_updateReynolds (GridObj_ops_lbm_optimised.cpp:1313-1321) at the start of
// LBM_multi_opt (:39-42): Re = L_RE * getReynoldsRampCoefficient((t + 1) dt)
// (GridUtils.cpp:1825-1833: 1 - cos(pi t / T) up to T, 1 after), nu = nud2nulbm(1 / Re)
// (GridUnits.h:128-131), omega = 1 / (nu / cs^2 + 0.5)
void update_reynolds(lumagpu_grid *g) {
    const double t = (g->t + 1) * g->d.dt;
    double coef = 1.0;
    if (t <= g->re_ramp) coef = 1.0 - std::cos(3.14159265358979323846 * t / g->re_ramp);
    const double re = g->re * coef;
    const double nu = ((1.0 / re) * g->d.dt) / (g->d.dh * g->d.dh);
    const double cs = 1.0 / std::sqrt(3.0);
    g->d.omega = 1.0 / ((nu / (cs * cs)) + 0.5);
}

// d omega_t / d omega_end for the step update_reynolds sets up: omega_t = 1 / (nu_t / cs^2 + 0.5)
// with nu_t = c / (L_RE coef_t), and omega_end, the omega after the ramp (coef = 1), follows
// L_RE the same way, so d omega_t / d omega_end = omega_t^2 nu_t / (omega_end^2 nu_end)
double reynolds_omega_weight(const lumagpu_grid *g) {
    const double t = (g->t + 1) * g->d.dt;
    double coef = 1.0;
    if (t <= g->re_ramp) coef = 1.0 - std::cos(3.14159265358979323846 * t / g->re_ramp);
    const double cs2 = 1.0 / 3.0, c = g->d.dt / (g->d.dh * g->d.dh);
    const double nt = c / (g->re * coef), ne = c / g->re;
    const double ot = 1.0 / (nt / cs2 + 0.5), oe = 1.0 / (ne / cs2 + 0.5);
    return (ot * ot * nt) / (oe * oe * ne);
}
}  // namespace lumagpu

namespace {

// LUMA-layout rho, u <-> the grid's macro arrays (Duals with zero derivative in the tangent
// form); macro_to_luma takes the values, or with tan the derivatives
hipError_t macro_from_luma(lumagpu_grid *g, const double *rho, const double *u) {
    const dim3 b = cell_block(g->d.N), gr = cell_grid(g->d.N, g->d.M, g->d.K);
    const int N = g->d.N, M = g->d.M, K = g->d.K;
    const bool t = g->melem != sizeof(double);
    if (g->D == 2) {
        if (t) k_macro_from_luma<2, Dual><<<gr, b, 0, g->stream>>>(g->rho, g->u, g->ncell, N, M, K, rho, u);
        else k_macro_from_luma<2><<<gr, b, 0, g->stream>>>(g->rho, g->u, g->ncell, N, M, K, rho, u);
    } else {
        if (t) k_macro_from_luma<3, Dual><<<gr, b, 0, g->stream>>>(g->rho, g->u, g->ncell, N, M, K, rho, u);
        else k_macro_from_luma<3><<<gr, b, 0, g->stream>>>(g->rho, g->u, g->ncell, N, M, K, rho, u);
    }
    return hipGetLastError();
}
hipError_t macro_to_luma(lumagpu_grid *g, double *rho, double *u, bool tan) {
    const dim3 b = cell_block(g->d.N), gr = cell_grid(g->d.N, g->d.M, g->d.K);
    const int N = g->d.N, M = g->d.M, K = g->d.K;
    const bool t = g->melem != sizeof(double);
    if (g->D == 2) {
        if (t && tan) k_macro_to_luma<2, Dual, true><<<gr, b, 0, g->stream>>>(g->rho, g->u, g->ncell, N, M, K, rho, u);
        else if (t) k_macro_to_luma<2, Dual><<<gr, b, 0, g->stream>>>(g->rho, g->u, g->ncell, N, M, K, rho, u);
        else k_macro_to_luma<2><<<gr, b, 0, g->stream>>>(g->rho, g->u, g->ncell, N, M, K, rho, u);
    } else {
        if (t && tan) k_macro_to_luma<3, Dual, true><<<gr, b, 0, g->stream>>>(g->rho, g->u, g->ncell, N, M, K, rho, u);
        else if (t) k_macro_to_luma<3, Dual><<<gr, b, 0, g->stream>>>(g->rho, g->u, g->ncell, N, M, K, rho, u);
        else k_macro_to_luma<3><<<gr, b, 0, g->stream>>>(g->rho, g->u, g->ncell, N, M, K, rho, u);
    }
    return hipGetLastError();
}

hipError_t launch_extrap(const lumagpu_grid *g) {
    const StepArgs a = make_args(g);
    double *xr = const_cast<double *>(a.xr);
    switch (g->d.precision) {
        case LUMAGPU_FP64: return launch_aux_f64(kAuxExtrapRight, g->Q, a, nullptr, nullptr, xr, g->parity, 1, g->stream);
        case LUMAGPU_FP32S: return launch_aux_f32s(kAuxExtrapRight, g->Q, a, nullptr, nullptr, xr, g->parity, 1, g->stream);
        case LUMAGPU_FP64T: return launch_aux_dual(kAuxExtrapRight, g->Q, a, nullptr, nullptr, xr, g->parity, 1, g->stream);
        case LUMAGPU_FP32T: return launch_aux_dual32(kAuxExtrapRight, g->Q, a, nullptr, nullptr, xr, g->parity, 1, g->stream);
        default: return launch_aux_f16s(kAuxExtrapRight, g->Q, a, nullptr, nullptr, xr, g->parity, 1, g->stream);
    }
}

// The first half of one grid's LBM_multi_opt once its children have stepped: stream order
// against the child and parent grids, the outlet side buffer, and _LBM_resetForces on a level
// holding IB bodies (:53-57).  Sweep 1's stream, BC and macro happen in the step kernel, which
// end_step launches; what the IB step reads of sweep 1 is recomputed at its supports.
int begin_step(lumagpu_grid *g) {
    // the outlet side buffer reads this grid's own populations only (the children read them
    // too, but nobody writes them before this grid's step): it runs while the children step
    // the previous step wrote this step's outlet sources unless the state changed since
    if (g->has_outlet && !g->xr_ready) {
        hipError_t e = launch_extrap(g);
        if (e != hipSuccess) return fail(LUMAGPU_EHIP, std::string("outlet launch: ") + hipGetErrorString(e));
    }
    for (lumagpu_grid *c : g->kids) {
        // the child reads this grid's populations, so order the two streams (one stream, as
        // lumagpu_grid_add_child sets up, orders them already)
        if (c->stream != g->stream) {
            HIP_TRY(hipEventRecord(c->ev_done, c->stream));
            HIP_TRY(hipStreamWaitEvent(g->stream, c->ev_done, 0));
        }
        if (c->parity != 0) return fail(LUMAGPU_ESTATE, "child grid out of step with its parent");
    }
    if (g->parent && g->parent->stream != g->stream) {
        // this grid reads the parent's populations: wait for the parent's previous work
        HIP_TRY(hipEventRecord(g->parent->ev_ready, g->parent->stream));
        HIP_TRY(hipStreamWaitEvent(g->stream, g->parent->ev_ready, 0));
    }
    // _LBM_resetForces at the start of the step on a level holding IB bodies (:53-57); with
    // gravity it writes the gravity component from the previous step's rho, and every site
    // then carries a force
    const int has_bodies = ibm_marker_count(g) > 0;
    g->all_forced = (g->gravity_on && has_bodies) ? 1 : 0;
    g->sweep1_zero = (has_bodies && !g->gravity_on) ? 1 : 0;
    if (g->all_forced) {
        const StepArgs ga = make_args(g);
        const double rr = 1.0 / std::pow(2.0, (double)g->d.level);   // GridObj.cpp:54
        const unsigned nb = (unsigned)((g->ncell + 255) / 256);
        if (g->melem == sizeof(double))
            k_gravity_reset<double><<<nb, 256, 0, g->stream>>>(ga, g->force, g->gravity, rr, g->gravity_dir);
        else
            k_gravity_reset<Dual><<<nb, 256, 0, g->stream>>>(ga, g->force, g->gravity, rr, g->gravity_dir);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

// The second half: stream, BC, macro (with the IB force), Guo forcing and BGK collision in one
// pass, the edge/corner and pressure sites around it, then the swap (AA parity) and ++t.
int end_step(lumagpu_grid *g, int subcycle, hipEvent_t e0, hipEvent_t e1) {
    LaunchSpec ls{g->Q, g->masked != 0, g->parity == 0, g->store_macro != 0, g->mlist, g->nmlist};
    ls.forced = g->all_forced || g->ibm || g->host_forced;
    StepArgs a = make_args(g);
    a.subcycle = subcycle;
    if (g->ib_fold) ibm_fold_args(g, a);
    g->ib_fold = 0;
    if (g->nedge) {
        hipError_t e = launch_edges(g, a, 0);
        if (e != hipSuccess) return fail(LUMAGPU_EHIP, std::string("edge launch: ") + hipGetErrorString(e));
    }
    const bool halo = g->zslab && g->parity == 0;
    if (halo) {
        hipError_t e = launch_zslab(g, a, 0);
        if (e != hipSuccess) return fail(LUMAGPU_EHIP, std::string("halo launch: ") + hipGetErrorString(e));
    }
    if (e0) HIP_TRY(hipEventRecord(e0, g->stream));
    hipError_t e = launch_step(g, ls, a);
    if (e != hipSuccess) return fail(LUMAGPU_EHIP, std::string("step launch: ") + hipGetErrorString(e));
    if (e1) HIP_TRY(hipEventRecord(e1, g->stream));
    if (a.xr_next) {
        g->xr_sel ^= 1;
        g->xr_ready = 1;
    } else {
        g->xr_ready = 0;
    }
    if (halo) {
        e = launch_zslab(g, a, 1);
        if (e != hipSuccess) return fail(LUMAGPU_EHIP, std::string("halo launch: ") + hipGetErrorString(e));
    }
    if (g->nedge) {
        e = launch_edges(g, a, 1);
        if (e != hipSuccess) return fail(LUMAGPU_EHIP, std::string("edge launch: ") + hipGetErrorString(e));
    }
    g->parity ^= 1;
    g->t += 1;
    return 0;
}

int apply_step(lumagpu_grid *g, int subcycle, hipEvent_t e0, hipEvent_t e1) {
    if (!g->has_state) return fail(LUMAGPU_ESTATE, "step before the populations were initialised");
    if (g->half) return fail(LUMAGPU_ESTATE, "a split step is open on this grid: finish it with lumagpu_force_collide_swap");
    if (g->re_ramp > 0.0) update_reynolds(g);
    // GridObj_ops_lbm_optimised.cpp:45-49: two sub-steps of each child first, child by child
    for (lumagpu_grid *c : g->kids)
        for (int sc = 0; sc < 2; ++sc)
            if (int r = apply_step(c, sc, nullptr, nullptr)) return r;
    if (int r = begin_step(g)) return r;
    // ObjectManager::ibm_apply between sweep 1 and sweep 2 (:120-121)
    if (g->ibm)
        if (int r = ibm_apply(g, true)) return r;
    return end_step(g, subcycle, e0, e1);
}

}  // namespace

namespace {

// The two-step kernel applies to a grid that the unlabelled periodic kernel steps on its own,
// with the same omega for both steps.  It fuses the local (even) step with the following
// neighbour (odd) step, so it starts at parity 1 (the local step's input layout) and leaves the
// grid there.
bool fused_ready(const lumagpu_grid *g) {
    return g->fuse_shape >= 0 && g->f2 && !g->masked && !g->zslab && !g->parent && g->kids.empty() &&
           ibm_marker_count(g) == 0 && !(g->re_ramp > 0.0) && g->parity == 1 && g->has_state;
}

// two steps (local, then neighbour) in one launch from g->f into g->f2, then the arrays trade places
int fused_pair(lumagpu_grid *g, hipEvent_t e0, hipEvent_t e1) {
    const StepArgs a = make_args(g);
    const bool store = g->store_macro != 0;
    if (e0) HIP_TRY(hipEventRecord(e0, g->stream));
    hipError_t e;
    switch (g->d.precision) {
        case LUMAGPU_FP64: e = launch_fused2_f64(g->fuse_shape, a, g->f2, g->fuse_zc, store, g->stream); break;
        case LUMAGPU_FP32S: e = launch_fused2_f32s(g->fuse_shape, a, g->f2, g->fuse_zc, store, g->stream); break;
        case LUMAGPU_FP16S: e = launch_fused2_f16s(g->fuse_shape, a, g->f2, g->fuse_zc, store, g->stream); break;
        default: return fail(LUMAGPU_EUNSUPPORTED, "no two-step kernel for this precision");
    }
    if (e != hipSuccess) return fail(LUMAGPU_EHIP, std::string("two-step launch: ") + hipGetErrorString(e));
    if (e1) HIP_TRY(hipEventRecord(e1, g->stream));
    std::swap(g->f, g->f2);
    g->t += 2;
    return 0;
}

// Capture two two-step launches (f -> f2 -> f, four steps) into a graph and replay it reps
// times.  Returns 0 when the steps ran, 1 when the capture was not possible (nothing ran, host
// state restored), or an error code.
int run_fused_graph(lumagpu_grid *g, int reps) {
    const int t0 = g->t;
    void *const f0 = g->f, *const f20 = g->f2;
    if (hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    int rc = 0;
    for (int k = 0; k < 2 && !rc; ++k) rc = fused_pair(g, nullptr, nullptr);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(g->stream, &graph);
    g->t = t0;
    g->f = f0;
    g->f2 = f20;
    if (rc || ec != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        return rc ? rc : 1;
    }
    hipGraphExec_t exec = nullptr;
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        return 1;
    }
    hipError_t e = hipSuccess;
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = hipGraphLaunch(exec, g->stream);
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(LUMAGPU_EHIP, std::string("graph launch: ") + hipGetErrorString(e));
    g->t += 4 * reps;
    return 0;
}

// no host work inside a step anywhere in the hierarchy (flexible bodies download forces and
// run the FEM on the host every step)
bool graph_ok(const lumagpu_grid *g) {
    bool ok = !g->parent;
    for_each_level(g, [&](const lumagpu_grid *x) {
        if (ibm_has_flexible(x)) ok = false;
        if (x->ramp_time > 0.0) ok = false;   // the ramp factors change every step
        if (x->re_ramp > 0.0) ok = false;     // so does omega under a Reynolds ramp
        if (x->fuse_shape >= 0) ok = false;   // the two-step kernel alternates two arrays
    });
    return ok;
}

void add_time(lumagpu_grid *g, int steps) {
    g->t += steps;
    for (lumagpu_grid *k : g->kids) add_time(k, 2 * steps);
}

// Everything the two captured steps pass to their kernels, level by level: while it stays the
// same, the graph captured from them is the graph a new capture would give.
std::vector<uint8_t> graph_key(lumagpu_grid *g) {
    std::vector<uint8_t> key;
    auto put = [&](const void *p, size_t n) {
        const uint8_t *b = static_cast<const uint8_t *>(p);
        key.insert(key.end(), b, b + n);
    };
    for_each_level(g, [&](lumagpu_grid *x) {
        const StepArgs a = make_args(x);
        put(&a, sizeof a);
        const int misc[] = {x->Q, x->masked, x->store_macro, x->nmlist, x->nedge, x->zslab, x->has_outlet,
                            x->gravity_on, x->gravity_dir, (int)x->kids.size(), x->parity, (int)x->d.precision,
                            x->d.level, x->d.regularised, x->xr_sel, x->xr_ready, x->xr_fold,
                            x->host_forced, x->ibm != nullptr};
        put(misc, sizeof misc);
        const void *ptrs[] = {x->mlist, x->edge, x->erho, x->stream, x->force};
        put(ptrs, sizeof ptrs);
        put(&x->gravity, sizeof x->gravity);
        ibm_graph_key(x, key);
    });
    return key;
}

// Capture two steps of g (and its descendants) into a graph and launch it reps times; the
// executable graph is kept and launched again by later runs while its key holds.
// Returns 0 when the steps ran, 1 when the capture was not possible (nothing ran, host state
// restored), or an error code.
int run_graph(lumagpu_grid *g, int reps) {
    std::vector<int> t0;
    for_each_level(g, [&](lumagpu_grid *x) { t0.push_back(x->t); });
    auto restore = [&]() {
        size_t i = 0;
        for_each_level(g, [&](lumagpu_grid *x) {
            x->t = t0[i++];
            x->parity = 0;
        });
    };
    std::vector<uint8_t> key = graph_key(g);
    if (!g->gexec || key != g->gkey) {
        if (hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            (void)hipGetLastError();
            return 1;
        }
        int rc = 0;
        for (int k = 0; k < 2 && !rc; ++k) rc = apply_step(g, 0, nullptr, nullptr);
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(g->stream, &graph);
        restore();
        if (rc || ec != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            (void)hipGetLastError();
            return rc ? rc : 1;
        }
        hipGraphExec_t exec = nullptr;
        const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            (void)hipGetLastError();
            return 1;
        }
        if (g->gexec) (void)hipGraphExecDestroy(g->gexec);
        g->gexec = exec;
        // the key after the capture: its host-side effects (IB supports settling) are in it
        g->gkey = graph_key(g);
    }
    hipGraphExec_t exec = g->gexec;
    // the graph runs on g's stream: order it after whatever the descendants' streams hold,
    // and the descendants' later work after it
    hipError_t e = hipSuccess;
    for_each_level(g, [&](lumagpu_grid *x) {
        if (x->stream == g->stream || e != hipSuccess) return;
        e = hipEventRecord(x->ev_done, x->stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(g->stream, x->ev_done, 0);
    });
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = hipGraphLaunch(exec, g->stream);
    for_each_level(g, [&](lumagpu_grid *x) {
        if (x->stream == g->stream || e != hipSuccess) return;
        e = hipEventRecord(g->ev_ready, g->stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(x->stream, g->ev_ready, 0);
    });
    if (e != hipSuccess) return fail(LUMAGPU_EHIP, std::string("graph launch: ") + hipGetErrorString(e));
    add_time(g, 2 * reps);
    return 0;
}

}  // namespace

namespace {
// Per-block partial sums of lumagpu_moments_noalloc: device memory fixed at load time, one copy
// per GPU, at most kScratchBlocks row blocks.  Host code never hands the symbol itself to a
// kernel (in host code it is a stand-in, not a device address): it takes each GPU's address
// from hipGetSymbolAddress once, with that GPU current, and passes that.
constexpr size_t kScratchBlocks = size_t(1) << 18;
static __device__ __align__(256) double scratch[kScratchBlocks * kMomentCount];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

int scratch_on(int device, double **out) {
    if (device < 0 || device >= kMaxDevices) return fail(LUMAGPU_EINVAL, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(LUMAGPU_EHIP, "hipGetSymbolAddress returned no address for the reduction scratch");
        g_scratch_ptr[device] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[device];
    return 0;
}
}  // namespace

namespace lumagpu {
int step_grid(lumagpu_grid *g, int subcycle) { return apply_step(g, subcycle, nullptr, nullptr); }
// the device table of g's refined regions, rewritten whenever the list changes
int upload_kids(lumagpu_grid *g) {
    (void)hipStreamSynchronize(g->stream);   // no queued step reads the old table
    (void)hipFree(g->d_kids);
    g->d_kids = nullptr;
    if (g->kids.empty()) return 0;
    std::vector<ChildRef> h(g->kids.size());
    for (size_t i = 0; i < g->kids.size(); ++i) {
        const lumagpu_grid *c = g->kids[i];
        ChildRef &r = h[i];
        r.f = c->f;
        r.ps = c->ps;
        r.nx = c->d.N;
        r.ny = c->d.M;
        r.nz = c->d.K;
        for (int a = 0; a < 3; ++a) {
            r.lo[a] = c->d.coarse_lims[2 * a];
            r.hi[a] = c->d.coarse_lims[2 * a + 1];
        }
        r.rho = c->rho;
        r.u = c->u;
        r.ncell = c->ncell;
    }
    if (hipMalloc(&g->d_kids, h.size() * sizeof(ChildRef)) != hipSuccess)
        return fail(LUMAGPU_ENOMEM, "refined-region table allocation failed");
    HIP_TRY(hipMemcpy(g->d_kids, h.data(), h.size() * sizeof(ChildRef), hipMemcpyHostToDevice));
    return 0;
}
namespace {
// f_v of the listed cells from the post-local layout (population v of a cell in its slot
// opp(v)), into out[v * n + k]
template <class W, int Q>
__global__ void k_copy_slots(const W *f, size_t ps, const int *cells, int n, W *out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    for (int v = 0; v < Q; ++v) out[(size_t)v * n + k] = f[(size_t)Lattice<Q>::opp(v) * ps + (size_t)cells[k]];
}
struct Elem16 {
    uint64_t a, b;
};
template <class W>
void copy_slots(const lumagpu_grid *g) {
    const unsigned nb = (unsigned)((g->nstale + 127) / 128);
    const W *f = static_cast<const W *>(g->f);
    W *out = static_cast<W *>(g->stale);
    if (g->Q == 9) k_copy_slots<W, 9><<<nb, 128, 0, g->stream>>>(f, g->ps, g->stale_cells, g->nstale, out);
    else k_copy_slots<W, 19><<<nb, 128, 0, g->stream>>>(f, g->ps, g->stale_cells, g->nstale, out);
}
}  // namespace

int capture_stale(lumagpu_grid *g) {
    if (!g->nstale || !g->has_state) return 0;
    if (g->parity != 0) return fail(LUMAGPU_ESTATE, "refined-site capture needs the post-local layout");
    switch (g->elem) {
        case 2: copy_slots<uint16_t>(g); break;
        case 4: copy_slots<uint32_t>(g); break;
        case 8: copy_slots<uint64_t>(g); break;
        case 16: copy_slots<Elem16>(g); break;
        default: return fail(LUMAGPU_EUNSUPPORTED, "unexpected population element size");
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

// refined sites read by a site that does not coalesce (anything but TL2F, solid or refined),
// with the periodic wrap of the step kernels
static int build_stale(lumagpu_grid *g) {
    const int N = g->d.N, M = g->d.M, K = g->d.K;
    auto ty = [&](int i, int j, int k) {
        i = (i + N) % N; j = (j + M) % M; k = (k + K) % K;
        return g->hflags[(size_t)i + (size_t)N * ((size_t)j + (size_t)M * k)] & kTypeMask;
    };
    std::vector<int> idx(g->ncell, -1), cells;
    for (int k = 0; k < K; ++k)
        for (int j = 0; j < M; ++j)
            for (int i = 0; i < N; ++i) {
                const int t = ty(i, j, k);
                if (t == kTL2F || t == kRefined || t == kSolid) continue;
                // an outlet site's source two cells left (its side buffer, :246-251)
                if (t == kExtrapRight && i >= 2 && ty(i - 2, j, k) == kRefined) {
                    const size_t c = (size_t)(i - 2) + (size_t)N * ((size_t)j + (size_t)M * k);
                    if (idx[c] < 0) {
                        idx[c] = (int)cells.size();
                        cells.push_back((int)c);
                    }
                }
                for (int dk = (g->D == 3 ? -1 : 0); dk <= (g->D == 3 ? 1 : 0); ++dk)
                    for (int dj = -1; dj <= 1; ++dj)
                        for (int di = -1; di <= 1; ++di) {
                            if (ty(i + di, j + dj, k + dk) != kRefined) continue;
                            const int ii = (i + di + N) % N, jj = (j + dj + M) % M, kk = (k + dk + K) % K;
                            const size_t c = (size_t)ii + (size_t)N * ((size_t)jj + (size_t)M * kk);
                            if (idx[c] < 0) {
                                idx[c] = (int)cells.size();
                                cells.push_back((int)c);
                            }
                        }
            }
    // the same sites as before (a flag upload that changed only force bits): keep what was captured
    if (cells == g->stale_list) return 0;
    g->stale_list = cells;
    (void)hipStreamSynchronize(g->stream);
    (void)hipFree(g->stale);
    (void)hipFree(g->stale_idx);
    (void)hipFree(g->stale_cells);
    g->stale = nullptr;
    g->stale_idx = g->stale_cells = nullptr;
    g->nstale = 0;
    if (cells.empty()) return 0;
    if (hipMalloc(&g->stale, (size_t)g->Q * cells.size() * g->elem) != hipSuccess ||
        hipMalloc(&g->stale_idx, g->ncell * sizeof(int)) != hipSuccess ||
        hipMalloc(&g->stale_cells, cells.size() * sizeof(int)) != hipSuccess)
        return fail(LUMAGPU_ENOMEM, "refined-site buffer allocation failed");
    HIP_TRY(hipMemcpy(g->stale_idx, idx.data(), g->ncell * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(g->stale_cells, cells.data(), cells.size() * sizeof(int), hipMemcpyHostToDevice));
    g->nstale = (int)cells.size();
    return capture_stale(g);
}

// Large labelled grids step in two launches: k_aa_plain over the grid for the plain sites
// (fluid, every stencil neighbour fluid or TL2F: the k_plain_bits predicate) and the labelled
// kernel over the list of the other sites it steps.  Small grids, whose steps are launch-bound,
// keep one launch.
constexpr size_t kSplitMinCells = size_t(1) << 20;
static int build_mlist(lumagpu_grid *g) {
    if (g->ncell < kSplitMinCells || g->zslab) {
        g->nmlist = -1;
        return 0;
    }
    std::vector<uint8_t> types(g->ncell);
    for (size_t c = 0; c < g->ncell; ++c) types[c] = g->hflags[c] & kTypeMask;
    if (types == g->mlist_types && g->nmlist >= 0) return 0;   // only force bits changed
    const int N = g->d.N, M = g->d.M, K = g->d.K;
    auto ty = [&](int i, int j, int k) {
        i = (i + N) % N; j = (j + M) % M; k = (k + K) % K;
        return (int)types[(size_t)i + (size_t)N * ((size_t)j + (size_t)M * k)];
    };
    std::vector<int> list;
    for (int k = 0; k < K; ++k)
        for (int j = 0; j < M; ++j)
            for (int i = 0; i < N; ++i) {
                const size_t c = (size_t)i + (size_t)N * ((size_t)j + (size_t)M * k);
                const int t = types[c];
                const unsigned fl = g->hflags[c];
                // the sites k_aa_masked leaves alone
                if (t == kSolid || t == kRefined || (!g->d.regularised && t == kVelocity)) continue;
                if (t == kPressure || (t == kVelocity && ((fl >> kNormalShift) & 7) == kNormalEdge)) continue;
                bool plain = t == kFluid;
                for (int dk = (g->D == 3 ? -1 : 0); plain && dk <= (g->D == 3 ? 1 : 0); ++dk)
                    for (int dj = -1; plain && dj <= 1; ++dj)
                        for (int di = -1; plain && di <= 1; ++di) {
                            if (g->D == 3 && di != 0 && dj != 0 && dk != 0) continue;   // D3Q19: no corners
                            const int tn = ty(i + di, j + dj, k + dk);
                            plain = tn == kFluid || tn == kTL2F;
                        }
                if (!plain) list.push_back((int)c);
            }
    (void)hipStreamSynchronize(g->stream);
    (void)hipFree(g->mlist);
    g->mlist = nullptr;
    g->nmlist = -1;
    g->mlist_types.clear();
    if (!list.empty()) {
        if (hipMalloc(&g->mlist, list.size() * sizeof(int)) != hipSuccess)
            return fail(LUMAGPU_ENOMEM, "site list allocation failed");
        HIP_TRY(hipMemcpy(g->mlist, list.data(), list.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    g->nmlist = (int)list.size();
    g->mlist_types = std::move(types);
    return 0;
}

int upload_flags(lumagpu_grid *g) {
    if (g->hflags.size() != g->ncell) return fail(LUMAGPU_ESTATE, "host flag array does not cover the grid");
    if (!g->flags && hipMalloc(&g->flags, g->ncell) != hipSuccess)
        return fail(LUMAGPU_ENOMEM, "flag array allocation failed");
    const size_t nw = (g->ncell + 31) / 32;
    if (!g->plain && hipMalloc(&g->plain, nw * sizeof(uint32_t)) != hipSuccess)
        return fail(LUMAGPU_ENOMEM, "plain-site bitmap allocation failed");
    HIP_TRY(hipMemcpyAsync(g->flags, g->hflags.data(), g->ncell, hipMemcpyHostToDevice, g->stream));
    g->host_forced = std::any_of(g->hflags.begin(), g->hflags.end(), [](uint8_t f) { return (f & kForcedBit) != 0; });
    const unsigned nb = (unsigned)((nw + 255) / 256);
    if (g->Q == 9) k_plain_bits<9><<<nb, 256, 0, g->stream>>>(g->flags, g->d.N, g->d.M, g->d.K, g->ncell, g->plain);
    else k_plain_bits<19><<<nb, 256, 0, g->stream>>>(g->flags, g->d.N, g->d.M, g->d.K, g->ncell, g->plain);
    HIP_TRY(hipGetLastError());
    if (int r = build_mlist(g)) return r;
    return build_stale(g);
}
hipError_t outlet_prepass(const lumagpu_grid *g) { return launch_extrap(g); }
}  // namespace lumagpu

extern "C" {

const char *lumagpu_last_error(void) { return g_last_error.c_str(); }
int lumagpu_abi_version(void) { return LUMAGPU_ABI_VERSION; }

int lumagpu_grid_create(const lumagpu_grid_desc *desc, lumagpu_grid **out) {
    if (!desc || !out) return fail(LUMAGPU_EINVAL, "null descriptor or output pointer");
    *out = nullptr;
    const lumagpu_grid_desc &d = *desc;
    if (!((d.dims == 2 && d.Q == 9) || (d.dims == 3 && d.Q == 19)))
        return fail(LUMAGPU_EUNSUPPORTED, "lattice must be D2Q9 or D3Q19");
    if (d.N < 3 || d.M < 3 || d.K < 1 || (d.dims == 2 && d.K != 1) || (d.dims == 3 && d.K < 3))
        return fail(LUMAGPU_EINVAL, "grid extents too small (need >= 3 per active dimension; K = 1 in 2D)");
    if (d.precision < LUMAGPU_FP64 || d.precision > LUMAGPU_FP32T)
        return fail(LUMAGPU_EINVAL, "unknown precision");
    // GridObj_init_grids.cpp:353-356
    if (!(d.omega > 0.0) || d.omega >= 2.0)
        return fail(LUMAGPU_EPHYS, "LBM relaxation frequency omega too large (must be 0 < omega < 2)");
    if (d.level < 0) return fail(LUMAGPU_EINVAL, "negative refinement level");
    // the shifted forms store f - w with the reference density 1 (their collision is written
    // for it); the plain forms do not use rho0
    if ((d.precision == LUMAGPU_FP32S || d.precision == LUMAGPU_FP16S || d.precision == LUMAGPU_FP32T) &&
        d.rho0 != 1.0)
        return fail(LUMAGPU_EINVAL, "the shifted forms use the reference density rho0 = 1");

    lumagpu_grid *g = new lumagpu_grid();
    g->d = d;
    g->Q = d.Q;
    g->D = d.dims;
    g->ncell = (size_t)d.N * d.M * d.K;
    g->ps = (g->ncell + 255) / 256 * 256 + kPlanePad;
    const bool tan = d.precision == LUMAGPU_FP64T || d.precision == LUMAGPU_FP32T;
    g->elem = d.precision == LUMAGPU_FP64T                                  ? 16
              : (d.precision == LUMAGPU_FP64 || d.precision == LUMAGPU_FP32T) ? 8
              : d.precision == LUMAGPU_FP32S                                 ? 4
                                                                             : 2;
    g->melem = tan ? 16 : 8;
    auto cleanup = [&](int code, const std::string &m) {
        lumagpu_grid_destroy(g);
        return fail(code, m);
    };
    if (hipGetDevice(&g->device) != hipSuccess) return cleanup(LUMAGPU_EHIP, "hipGetDevice failed");
    if (hipStreamCreateWithFlags(&g->own_stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(LUMAGPU_EHIP, "hipStreamCreate failed");
    g->stream = g->own_stream;
    if (hipEventCreateWithFlags(&g->ev_done, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&g->ev_ready, hipEventDisableTiming) != hipSuccess)
        return cleanup(LUMAGPU_EHIP, "hipEventCreate failed");
    if (hipMalloc(&g->f, g->ps * g->Q * g->elem) != hipSuccess)
        return cleanup(LUMAGPU_ENOMEM, "population array allocation failed");
    if (hipMalloc(&g->rho, g->ncell * g->melem) != hipSuccess ||
        hipMalloc(&g->u, g->ncell * g->D * g->melem) != hipSuccess)
        return cleanup(LUMAGPU_ENOMEM, "macro array allocation failed");
    if (hipMalloc(&g->uin, 3 * (size_t)d.M * sizeof(double)) != hipSuccess)
        return cleanup(LUMAGPU_ENOMEM, "inlet array allocation failed");
    if (hipMemsetAsync(g->uin, 0, 3 * (size_t)d.M * sizeof(double), g->stream) != hipSuccess)
        return cleanup(LUMAGPU_EHIP, "inlet array clear failed");
    const dim3 gr = cell_grid(d.N, d.M, d.K);
    const size_t nblocks = (size_t)gr.x * gr.y * gr.z;
    if (hipMalloc(&g->partials, nblocks * kMomentCount * sizeof(double)) != hipSuccess ||
        hipMalloc(&g->moments_out, kMomentCount * sizeof(double)) != hipSuccess)
        return cleanup(LUMAGPU_ENOMEM, "reduction buffers allocation failed");
    if (d.forced) {
        if (hipMalloc(&g->force, g->ncell * g->D * g->melem) != hipSuccess)
            return cleanup(LUMAGPU_ENOMEM, "force array allocation failed");
        if (hipMemsetAsync(g->force, 0, g->ncell * g->D * g->melem, g->stream) != hipSuccess)
            return cleanup(LUMAGPU_EHIP, "force array clear failed");
    }
    *out = g;
    return 0;
}

int lumagpu_grid_destroy(lumagpu_grid *g) {
    if (!g) return 0;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    (void)hipFree(g->f);
    (void)hipFree(g->f2);
    (void)hipFree(g->flags);
    (void)hipFree(g->plain);
    (void)hipFree(g->stale);
    (void)hipFree(g->stale_idx);
    (void)hipFree(g->stale_cells);
    (void)hipFree(g->mlist);
    (void)hipFree(g->force);
    (void)hipFree(g->uin);
    (void)hipFree(g->rho);
    (void)hipFree(g->u);
    (void)hipFree(g->partials);
    (void)hipFree(g->stage);
    (void)hipFree(g->moments_out);
    (void)hipFree(g->xr);
    (void)hipFree(g->edge);
    (void)hipFree(g->erho);
    ibm_destroy(g);
    if (g->parent) {
        auto &pk = g->parent->kids;
        pk.erase(std::remove(pk.begin(), pk.end(), g), pk.end());
        (void)upload_kids(g->parent);
    }
    for (lumagpu_grid *k : g->kids) {
        if (k->parent == g) k->parent = nullptr;
        // a detached subtree goes back to its own streams
        for_each_level(k, [](lumagpu_grid *x) { x->stream = x->own_stream; });
    }
    (void)hipFree(g->d_kids);
    if (g->ev_done) (void)hipEventDestroy(g->ev_done);
    if (g->ev_ready) (void)hipEventDestroy(g->ev_ready);
    if (g->gexec) (void)hipGraphExecDestroy(g->gexec);
    if (g->own_stream) (void)hipStreamDestroy(g->own_stream);
    delete g;
    return 0;
}

int lumagpu_set_labels(lumagpu_grid *g, const int *lattyp, const double *xpos, const double *ypos,
                       const double *zpos) {
    if (int r = check_grid(g)) return r;
    if (!lattyp) return fail(LUMAGPU_EINVAL, "null label array");
    // the grid's labels change only once the whole array has been accepted
    if (xpos && ypos && zpos) {
        g->xpos.assign(xpos, xpos + g->d.N);
        g->ypos.assign(ypos, ypos + g->d.M);
        g->zpos.assign(zpos, zpos + g->d.K);
    }
    const int N = g->d.N, M = g->d.M, K = g->d.K;
    // keep the force bits a previous lumagpu_set_force call placed
    std::vector<uint8_t> prev = g->hflags;
    if (prev.size() != g->ncell) prev.assign(g->ncell, 0);
    std::vector<uint8_t> nf(g->ncell, 0);
    int any = 0, outlet = 0;
    std::vector<int> ecell, enrm;   // edge/corner velocity sites (device cell, packed normal)
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < M; ++j)
            for (int k = 0; k < K; ++k) {
                const size_t id = (size_t)k + (size_t)j * K + (size_t)i * K * M;
                const int ty = lattyp[id];
                uint8_t fl = 0;
                switch (ty) {
                    case kSolid:
                    case kFluid: fl = (uint8_t)ty; break;
                    case kVelocity: {
                        fl = (uint8_t)ty;
                        if (g->d.regularised) {
                            if (!xpos || !ypos || !zpos)
                                return fail(LUMAGPU_EINVAL, "positions are required for regularised velocity sites");
                            int code = 0, n[3];
                            const int cnt = wall_normal(g->d, xpos[i], ypos[j], zpos[k], &code, n);
                            if (cnt == 0)
                                return fail(LUMAGPU_EINVAL,
                                            "Velocity site not outside domain walls is currently not supported.");
                            if (cnt > 1) {
                                code = kNormalEdge;
                                ecell.push_back(i + N * (j + M * k));
                                enrm.push_back((n[0] + 1) | ((n[1] + 1) << 2) | ((n[2] + 1) << 4));
                            }
                            fl |= (uint8_t)(code << kNormalShift);
                        }
                        break;
                    }
                    case kPressure: {
                        // regularised pressure BC on a face (:389-402); LUMA has no other form
                        if (!g->d.regularised)
                            return fail(LUMAGPU_EUNSUPPORTED, "pressure sites need regularised boundaries");
                        if (!xpos || !ypos || !zpos)
                            return fail(LUMAGPU_EINVAL, "positions are required for regularised pressure sites");
                        int code = 0, n[3];
                        const int cnt = wall_normal(g->d, xpos[i], ypos[j], zpos[k], &code, n);
                        if (cnt == 0)
                            return fail(LUMAGPU_EINVAL,
                                        "Velocity site not outside domain walls is currently not supported.");
                        if (cnt > 1) return fail(LUMAGPU_EINVAL, "Pressure BC not supported at corner/edge site. Exiting.");
                        fl = (uint8_t)(ty | (code << kNormalShift));
                        ecell.push_back(i + N * (j + M * k));
                        enrm.push_back((n[0] + 1) | ((n[1] + 1) << 2) | ((n[2] + 1) << 4));
                        break;
                    }
                    case kRefined:
                    case kTL2C:
                    case kTL2F: fl = (uint8_t)ty; break;
                    case kExtrapRight:
                        // LUMA pulls from two sites to the left of the source (:246-251); the
                        // outlet must be the last x column for that to stay on the grid
                        if (i != N - 1 || N < 4)
                            return fail(LUMAGPU_EUNSUPPORTED, "ExtrapolateRight sites must form the last x column");
                        fl = (uint8_t)ty;
                        outlet = 1;
                        break;
                    default:
                        return fail(LUMAGPU_EUNSUPPORTED,
                                    "site label " + std::to_string(ty) + " is not supported by this build");
                }
                if (ty != kFluid) any = 1;
                const size_t cell = (size_t)i + (size_t)N * ((size_t)j + (size_t)M * k);
                nf[cell] = (uint8_t)(fl | (prev[cell] & kForcedBit));
            }
    // Edge/corner velocity sites extrapolate rho, pressure sites their tangential u, from x + n
    // and x + 2n (_LBM_updateAndExtrapolate, :1353-1412).  Those neighbours must be sites whose
    // sweep 1 computes rho and u (fluid, TL2F); for any other type LUMA reads a value left from
    // an earlier step or its initial state.
    const int ne = (int)ecell.size();
    std::vector<int> etab(4 * (size_t)ne);
    for (int e = 0; e < ne; ++e) {
        const int c = ecell[e];
        const int i = c % N, j = (c / N) % M, k = c / (N * M);
        const int n[3] = {(enrm[e] & 3) - 1, ((enrm[e] >> 2) & 3) - 1, ((enrm[e] >> 4) & 3) - 1};
        int nb[2];
        for (int s = 1; s <= 2; ++s) {
            const int ii = i + s * n[0], jj = j + s * n[1], kk = k + s * n[2];
            if (ii < 0 || ii >= N || jj < 0 || jj >= M || kk < 0 || kk >= K)
                return fail(LUMAGPU_EINVAL, "Extrapolation does not have enough available cells.");
            nb[s - 1] = ii + N * (jj + M * kk);
            const int nty = nf[nb[s - 1]] & kTypeMask;
            if (nty != kFluid && nty != kTL2F)
                return fail(LUMAGPU_EUNSUPPORTED,
                            "edge/corner velocity or pressure site whose extrapolation neighbours are not fluid "
                            "is not supported");
        }
        etab[e] = c;
        etab[ne + e] = enrm[e];
        etab[2 * ne + e] = nb[0];
        etab[3 * ne + e] = nb[1];
    }
    (void)hipFree(g->edge);
    (void)hipFree(g->erho);
    g->edge = nullptr;
    g->erho = nullptr;
    g->nedge = 0;
    if (ne) {
        if (hipMalloc(&g->edge, etab.size() * sizeof(int)) != hipSuccess ||
            hipMalloc(&g->erho, 8 * (size_t)ne * g->melem) != hipSuccess)
            return fail(LUMAGPU_ENOMEM, "edge site table allocation failed");
        HIP_TRY(hipMemcpyAsync(g->edge, etab.data(), etab.size() * sizeof(int), hipMemcpyHostToDevice, g->stream));
        g->nedge = ne;
    }
    g->hflags = std::move(nf);
    if (int r = upload_flags(g)) return r;
    if (outlet && !g->xr &&
        hipMalloc(&g->xr, 2 * (size_t)g->Q * g->d.M * g->d.K * g->melem) != hipSuccess)
        return fail(LUMAGPU_ENOMEM, "outlet buffer allocation failed");
    HIP_TRY(hipStreamSynchronize(g->stream));
    g->has_outlet = outlet;
    // the step kernels can write the next step's outlet sources when every outlet row's source
    // two cells left is a site they step (collided, or streamed only for TL2C)
    g->xr_fold = 0;
    g->xr_ready = 0;
    if (outlet && g->d.N >= 3) {
        bool ok = true;
        const int N = g->d.N, M = g->d.M, K = g->d.K;
        for (int k = 0; k < K && ok; ++k)
            for (int j = 0; j < M && ok; ++j) {
                const size_t row = (size_t)N * ((size_t)j + (size_t)M * k);
                if ((g->hflags[row + N - 1] & kTypeMask) != kExtrapRight) continue;
                const unsigned fl = g->hflags[row + N - 3];
                const int t = fl & kTypeMask;
                const bool stepped = t == kFluid || t == kTL2F || t == kTL2C ||
                                     (t == kVelocity && g->d.regularised && ((fl >> kNormalShift) & 7) != kNormalEdge);
                ok = stepped;
            }
        g->xr_fold = ok ? 1 : 0;
    }
    g->masked = any || g->d.forced;
    return 0;
}

}  // extern "C"

struct lumagpu_checkpoint {
    const lumagpu_grid *owner = nullptr;
    void *f = nullptr;
    double *rho = nullptr, *u = nullptr;
    size_t fbytes = 0;
    int parity = 0, t = 0, has_state = 0;
};

extern "C" {

int lumagpu_checkpoint_save(lumagpu_grid *g, lumagpu_checkpoint **out) {
    if (int r = check_grid(g)) return r;
    if (!out) return fail(LUMAGPU_EINVAL, "null output");
    if (g->parent || !g->kids.empty()) return fail(LUMAGPU_EUNSUPPORTED, "checkpoints cover single-level grids");
    if (ibm_has_flexible(g)) return fail(LUMAGPU_EUNSUPPORTED, "checkpoints do not hold flexible-body state");
    auto *ck = new lumagpu_checkpoint();
    ck->owner = g;
    ck->fbytes = g->ps * g->Q * g->elem;
    if (hipMalloc(&ck->f, ck->fbytes) != hipSuccess || hipMalloc(&ck->rho, g->ncell * g->melem) != hipSuccess ||
        hipMalloc(&ck->u, g->ncell * g->D * g->melem) != hipSuccess) {
        lumagpu_checkpoint_free(ck);
        return fail(LUMAGPU_ENOMEM, "checkpoint allocation failed");
    }
    if (hipMemcpyAsync(ck->f, g->f, ck->fbytes, hipMemcpyDeviceToDevice, g->stream) != hipSuccess ||
        hipMemcpyAsync(ck->rho, g->rho, g->ncell * g->melem, hipMemcpyDeviceToDevice, g->stream) != hipSuccess ||
        hipMemcpyAsync(ck->u, g->u, g->ncell * g->D * g->melem, hipMemcpyDeviceToDevice, g->stream) != hipSuccess ||
        hipStreamSynchronize(g->stream) != hipSuccess) {
        lumagpu_checkpoint_free(ck);
        return fail(LUMAGPU_EHIP, "checkpoint copy failed");
    }
    ck->parity = g->parity;
    ck->t = g->t;
    ck->has_state = g->has_state;
    *out = ck;
    return 0;
}

int lumagpu_checkpoint_restore(lumagpu_grid *g, const lumagpu_checkpoint *ck) {
    if (int r = check_grid(g)) return r;
    if (!ck) return fail(LUMAGPU_EINVAL, "null checkpoint");
    if (ck->owner != g) return fail(LUMAGPU_EINVAL, "checkpoint belongs to another grid");
    HIP_TRY(hipMemcpyAsync(g->f, ck->f, ck->fbytes, hipMemcpyDeviceToDevice, g->stream));
    g->xr_ready = 0;
    HIP_TRY(hipMemcpyAsync(g->rho, ck->rho, g->ncell * g->melem, hipMemcpyDeviceToDevice, g->stream));
    HIP_TRY(hipMemcpyAsync(g->u, ck->u, g->ncell * g->D * g->melem, hipMemcpyDeviceToDevice, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    g->parity = ck->parity;
    g->t = ck->t;
    g->has_state = ck->has_state;
    return 0;
}

int lumagpu_checkpoint_free(lumagpu_checkpoint *ck) {
    if (!ck) return 0;
    (void)hipFree(ck->f);
    (void)hipFree(ck->rho);
    (void)hipFree(ck->u);
    delete ck;
    return 0;
}

int lumagpu_set_zslab(lumagpu_grid *g, int on) {
    if (int r = check_grid(g)) return r;
    if (!on) {
        g->zslab = 0;
        return 0;
    }
    if (g->d.dims != 3 || g->Q != 19 || g->d.K < 3)
        return fail(LUMAGPU_EINVAL, "a z-slab needs a D3Q19 grid with at least one interior and two halo planes");
    if (g->masked || g->d.forced || g->parent || !g->kids.empty() || g->ibm)
        return fail(LUMAGPU_EUNSUPPORTED, "z-slab halo exchange is implemented for unlabelled, unforced single-level grids");
    if (g->melem != sizeof(double)) return fail(LUMAGPU_EUNSUPPORTED, "z-slabs run in the FP64 and shifted forms");
    if (g->parity != 0) return fail(LUMAGPU_ESTATE, "switch to a z-slab between step pairs");
    g->zslab = 1;
    return 0;
}

int lumagpu_grid_add_child(lumagpu_grid *parent, lumagpu_grid *child) {
    if (int r = check_grid(parent)) return r;
    if (int r = check_grid(child)) return r;
    if (child->parent) return fail(LUMAGPU_ESTATE, "grid already has a parent");
    if (parent->kids.size() >= kMaxKids)
        return fail(LUMAGPU_EUNSUPPORTED, "at most " + std::to_string(kMaxKids) + " refined regions per grid");
    const lumagpu_grid_desc &p = parent->d, &c = child->d;
    if (c.level != p.level + 1) return fail(LUMAGPU_EINVAL, "child level must be parent level + 1");
    if (c.precision != p.precision || c.Q != p.Q) return fail(LUMAGPU_EINVAL, "child and parent must share lattice and precision");
    // LBM_initSubGrid (GridObj_init_grids.cpp:416-425): N = 2*(CoarseLimsMax - CoarseLimsMin + 1)
    const int ext[3] = {c.N, c.M, c.K};
    const int pext[3] = {p.N, p.M, p.K};
    for (int a = 0; a < c.dims; ++a) {
        const int lo = c.coarse_lims[2 * a], hi = c.coarse_lims[2 * a + 1];
        if (lo < 0 || hi >= pext[a] || hi < lo || ext[a] != 2 * (hi - lo + 1))
            return fail(LUMAGPU_EINVAL, "child extents do not match its coarse limits on the parent");
    }
    // regions of one grid are disjoint (GridUtils::getSubGrid takes the first that holds a site)
    for (const lumagpu_grid *o : parent->kids) {
        bool apart = false;
        for (int a = 0; a < c.dims; ++a)
            apart = apart || c.coarse_lims[2 * a + 1] < o->d.coarse_lims[2 * a] ||
                    o->d.coarse_lims[2 * a + 1] < c.coarse_lims[2 * a];
        if (!apart) return fail(LUMAGPU_EINVAL, "refined regions of one grid overlap");
    }
    parent->kids.push_back(child);
    child->parent = parent;
    if (int r = upload_kids(parent)) {
        parent->kids.pop_back();
        child->parent = nullptr;
        return r;
    }
    // The hierarchy's levels step one after another (each child reads its parent and the parent
    // its children), so they share the root's stream: one in-order queue instead of an event
    // hop between queues at every level change.
    HIP_TRY(hipStreamSynchronize(child->stream));
    HIP_TRY(hipStreamSynchronize(parent->stream));
    for_each_level(child, [&](lumagpu_grid *x) { x->stream = parent->stream; });
    return 0;
}

int lumagpu_set_inlet(lumagpu_grid *g, const double *ux, const double *uy, const double *uz) {
    if (int r = check_grid(g)) return r;
    const size_t M = g->d.M;
    std::vector<double> h(3 * M, 0.0);
    for (size_t j = 0; j < M; ++j) {
        h[j] = ux ? ux[j] : 0.0;
        h[M + j] = uy ? uy[j] : 0.0;
        h[2 * M + j] = uz ? uz[j] : 0.0;
    }
    HIP_TRY(hipMemcpyAsync(g->uin, h.data(), 3 * M * sizeof(double), hipMemcpyHostToDevice, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    return 0;
}

int lumagpu_init_equilibrium(lumagpu_grid *g, const double *rho, const double *u) {
    if (int r = check_grid(g)) return r;
    if (!rho || !u) return fail(LUMAGPU_EINVAL, "null rho or u");
    const size_t nb = g->ncell * (1 + g->D) * sizeof(double);
    if (int r = ensure_stage(g, nb)) return r;
    double *drho = g->stage, *du = g->stage + g->ncell;
    HIP_TRY(hipMemcpyAsync(drho, rho, g->ncell * sizeof(double), hipMemcpyHostToDevice, g->stream));
    HIP_TRY(hipMemcpyAsync(du, u, g->ncell * g->D * sizeof(double), hipMemcpyHostToDevice, g->stream));
    HIP_TRY(launch_aux(g, kAuxInitFeq, drho, du, nullptr));
    HIP_TRY(macro_from_luma(g, drho, du));
    g->parity = 0;
    g->has_state = 1;
    g->xr_ready = 0;
    if (int r = capture_stale(g)) return r;
    HIP_TRY(hipStreamSynchronize(g->stream));
    return 0;
}

int lumagpu_upload_f(lumagpu_grid *g, const double *f) {
    if (int r = check_grid(g)) return r;
    if (!f) return fail(LUMAGPU_EINVAL, "null f");
    const size_t nb = g->ncell * g->Q * sizeof(double);
    if (int r = ensure_stage(g, nb)) return r;
    HIP_TRY(hipMemcpyAsync(g->stage, f, nb, hipMemcpyHostToDevice, g->stream));
    HIP_TRY(launch_aux(g, kAuxUpload, g->stage, nullptr, nullptr));
    g->parity = 0;
    g->has_state = 1;
    g->xr_ready = 0;
    if (int r = capture_stale(g)) return r;
    HIP_TRY(hipStreamSynchronize(g->stream));
    return 0;
}

int lumagpu_upload_macro(lumagpu_grid *g, const double *rho, const double *u) {
    if (int r = check_grid(g)) return r;
    if (!rho || !u) return fail(LUMAGPU_EINVAL, "null rho or u");
    const size_t nb = g->ncell * (1 + g->D) * sizeof(double);
    if (int r = ensure_stage(g, nb)) return r;
    double *drho = g->stage, *du = g->stage + g->ncell;
    HIP_TRY(hipMemcpyAsync(drho, rho, g->ncell * sizeof(double), hipMemcpyHostToDevice, g->stream));
    HIP_TRY(hipMemcpyAsync(du, u, g->ncell * g->D * sizeof(double), hipMemcpyHostToDevice, g->stream));
    HIP_TRY(macro_from_luma(g, drho, du));
    HIP_TRY(hipStreamSynchronize(g->stream));
    return 0;
}

int lumagpu_download_f(lumagpu_grid *g, double *f) {
    if (int r = check_grid(g)) return r;
    if (!f) return fail(LUMAGPU_EINVAL, "null f");
    if (!g->has_state) return fail(LUMAGPU_ESTATE, "no state to download");
    const size_t nb = g->ncell * g->Q * sizeof(double);
    if (int r = ensure_stage(g, nb)) return r;
    HIP_TRY(launch_aux(g, kAuxDownload, nullptr, nullptr, g->stage));
    HIP_TRY(hipMemcpyAsync(f, g->stage, nb, hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    return 0;
}

int lumagpu_download_macro(lumagpu_grid *g, double *rho, double *u) {
    if (int r = check_grid(g)) return r;
    if (!rho || !u) return fail(LUMAGPU_EINVAL, "null rho or u");
    const size_t nb = g->ncell * (1 + g->D) * sizeof(double);
    if (int r = ensure_stage(g, nb)) return r;
    double *drho = g->stage, *du = g->stage + g->ncell;
    HIP_TRY(macro_to_luma(g, drho, du, false));
    HIP_TRY(hipMemcpyAsync(rho, drho, g->ncell * sizeof(double), hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipMemcpyAsync(u, du, g->ncell * g->D * sizeof(double), hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    return 0;
}

int lumagpu_tangent_seed_omega(lumagpu_grid *g, double d_omega) {
    if (int r = check_grid(g)) return r;
    if (g->d.precision != LUMAGPU_FP64T)
        return fail(LUMAGPU_ESTATE, "an omega seed needs an FP64 tangent-form grid (LUMAGPU_FP64T)");
    g->d_omega = d_omega;
    return 0;
}

int lumagpu_download_tangent(lumagpu_grid *g, double *df, double *drho, double *du) {
    if (int r = check_grid(g)) return r;
    if (g->melem == sizeof(double)) return fail(LUMAGPU_ESTATE, "derivatives exist on tangent-form grids only");
    if (!g->has_state) return fail(LUMAGPU_ESTATE, "no state to download");
    const size_t nf = g->ncell * g->Q * sizeof(double), nm = g->ncell * (1 + g->D) * sizeof(double);
    if (int r = ensure_stage(g, nf > nm ? nf : nm)) return r;
    if (df) {
        HIP_TRY(launch_aux(g, kAuxDownloadTan, nullptr, nullptr, g->stage));
        HIP_TRY(hipMemcpyAsync(df, g->stage, nf, hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
    }
    if (drho || du) {
        double *sr = g->stage, *su = g->stage + g->ncell;
        HIP_TRY(macro_to_luma(g, sr, su, true));
        if (drho) HIP_TRY(hipMemcpyAsync(drho, sr, g->ncell * sizeof(double), hipMemcpyDeviceToHost, g->stream));
        if (du) HIP_TRY(hipMemcpyAsync(du, su, g->ncell * g->D * sizeof(double), hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
    }
    return 0;
}

int lumagpu_set_force(lumagpu_grid *g, const double *force_xyz) {
    if (int r = check_grid(g)) return r;
    if (!g->d.forced) return fail(LUMAGPU_ESTATE, "grid was created without forcing (desc.forced = 0)");
    if (!force_xyz) return fail(LUMAGPU_EINVAL, "null force");
    const int N = g->d.N, M = g->d.M, K = g->d.K, D = g->D;
    if (g->hflags.empty()) g->hflags.assign(g->ncell, (uint8_t)kFluid);
    const size_t w = g->melem / sizeof(double);   // (value, derivative) pairs in the tangent form
    std::vector<double> soa(g->ncell * D * w, 0.0);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < M; ++j)
            for (int k = 0; k < K; ++k) {
                const size_t id = (size_t)k + (size_t)j * K + (size_t)i * K * M;
                const size_t cell = (size_t)i + (size_t)N * ((size_t)j + (size_t)M * k);
                bool nz = false;
                for (int d = 0; d < D; ++d) {
                    const double v = force_xyz[d + D * id];
                    soa[w * (d * g->ncell + cell)] = v;
                    nz = nz || v != 0.0;
                }
                g->hflags[cell] = (uint8_t)((g->hflags[cell] & ~kForcedBit) | (nz ? kForcedBit : 0));
            }
    if (int r = upload_flags(g)) return r;
    HIP_TRY(hipMemcpyAsync(g->force, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    g->masked = 1;
    return 0;
}

int lumagpu_set_velocity_ramp(lumagpu_grid *g, double ramp_time) {
    if (int r = check_grid(g)) return r;
    if (!(ramp_time >= 0.0)) return fail(LUMAGPU_EINVAL, "ramp time must be >= 0");
    if (ramp_time > 0.0 && !(g->d.dt > 0.0)) return fail(LUMAGPU_EINVAL, "a velocity ramp needs desc.dt > 0");
    g->ramp_time = ramp_time;
    return 0;
}

int lumagpu_set_temporal_blocking(lumagpu_grid *g, int shape, int zchunk) {
    if (int r = check_grid(g)) return r;
    if (shape < 0) {
        if (g->f2) {
            HIP_TRY(hipStreamSynchronize(g->stream));
            HIP_TRY(hipFree(g->f2));
        }
        g->f2 = nullptr;
        g->fuse_shape = -1;
        return 0;
    }
    int tx, ty;
    if (!fused_tile(shape, tx, ty)) return fail(LUMAGPU_EINVAL, "unknown tile shape");
    if (g->Q != 19) return fail(LUMAGPU_EUNSUPPORTED, "the two-step kernel runs on D3Q19 grids");
    const int p = g->d.precision;
    if (p != LUMAGPU_FP64 && p != LUMAGPU_FP32S && p != LUMAGPU_FP16S)
        return fail(LUMAGPU_EUNSUPPORTED, "the two-step kernel runs in the FP64, FP32- and FP16-shifted forms");
    if ((shape & 8) && g->elem > 4)
        return fail(LUMAGPU_EINVAL, "double-buffered tiles need storage of at most 4 bytes (FP32- or FP16-shifted)");
    if (!fused_shape_fits(shape, g->d.N, g->d.M))
        return fail(LUMAGPU_EINVAL, "N and M must be whole numbers of the tile's x and y extents");
    if (zchunk < 1 || g->d.K < 3) return fail(LUMAGPU_EINVAL, "z chunk must be >= 1 and K >= 3");
    if (!g->f2 && hipMalloc(&g->f2, g->ps * g->Q * g->elem) != hipSuccess) {
        g->f2 = nullptr;
        return fail(LUMAGPU_ENOMEM, "second population array");
    }
    g->fuse_shape = shape;
    g->fuse_zc = zchunk;
    return 0;
}

int lumagpu_set_reynolds_ramp(lumagpu_grid *g, double re, double ramp_time) {
    if (int r = check_grid(g)) return r;
    if (!(ramp_time >= 0.0)) return fail(LUMAGPU_EINVAL, "Reynolds ramp time must be >= 0");
    if (ramp_time > 0.0 && !(re > 0.0)) return fail(LUMAGPU_EINVAL, "a Reynolds ramp needs L_RE > 0");
    bool dims_ok = true;
    for_each_level(g, [&](const lumagpu_grid *x) {
        if (ramp_time > 0.0 && !(x->d.dt > 0.0 && x->d.dh > 0.0)) dims_ok = false;
    });
    if (!dims_ok) return fail(LUMAGPU_EINVAL, "a Reynolds ramp needs desc.dt and desc.dh > 0 on every level");
    for_each_level(g, [&](lumagpu_grid *x) {
        x->re = re;
        x->re_ramp = ramp_time;
    });
    return 0;
}

int lumagpu_get_omega(const lumagpu_grid *g, double *omega) {
    if (int r = check_grid(g)) return r;
    if (omega) *omega = g->d.omega;
    return 0;
}

int lumagpu_set_pressure_delta(lumagpu_grid *g, double p) {
    if (int r = check_grid(g)) return r;
    if (!std::isfinite(p)) return fail(LUMAGPU_EINVAL, "pressure delta must be finite");
    if (p != 0.0 && !(g->d.dm > 0.0 && g->d.dt > 0.0 && g->d.dh > 0.0))
        return fail(LUMAGPU_EINVAL, "a pressure delta needs desc.dh, dt and dm > 0");
    // GridUnits::pd2dlbm (GridUnits.h:177-180)
    const double cs = 1.0 / std::sqrt(3.0);
    g->pressure_drho = p == 0.0 ? 0.0 : (p * g->d.dh * (g->d.dt * g->d.dt) / g->d.dm) / (cs * cs);
    return 0;
}

int lumagpu_set_gravity(lumagpu_grid *g, double gravity, int direction) {
    if (int r = check_grid(g)) return r;
    if (!g->d.forced) return fail(LUMAGPU_ESTATE, "gravity needs a grid created with desc.forced = 1");
    if (direction < 0 || direction >= g->D) return fail(LUMAGPU_EINVAL, "gravity direction out of range");
    g->gravity_on = 1;
    g->gravity = gravity;
    g->gravity_dir = direction;
    g->store_macro = 1;   // the reset reads every site's rho of the previous step
    return 0;
}

int lumagpu_set_store_macro(lumagpu_grid *g, int on) {
    if (int r = check_grid(g)) return r;
    g->store_macro = on ? 1 : 0;
    return 0;
}

int lumagpu_step(lumagpu_grid *g, int subcycle) {
    if (int r = check_grid(g)) return r;
    if (subcycle < 0 || subcycle > 1) return fail(LUMAGPU_EINVAL, "subcycle must be 0 or 1");
    return apply_step(g, subcycle, nullptr, nullptr);
}

int lumagpu_stream_bc_macro(lumagpu_grid *g, int subcycle) {
    if (int r = check_grid(g)) return r;
    if (subcycle < 0 || subcycle > 1) return fail(LUMAGPU_EINVAL, "subcycle must be 0 or 1");
    if (!g->has_state) return fail(LUMAGPU_ESTATE, "step before the populations were initialised");
    if (g->half) return fail(LUMAGPU_ESTATE, "a split step is already open on this grid");
    if (g->re_ramp > 0.0) update_reynolds(g);
    if (int r = begin_step(g)) return r;
    // sweep 1's rho, u at every fluid, TL2F and regularised-face site, as LUMA's arrays hold them
    // between the sweeps (edge/corner and pressure sites get theirs in the second half)
    const StepArgs a = make_args(g);
    hipError_t e = launch_sites(g, a, g->parity == 0, (int)g->ncell, nullptr, 0);
    if (e != hipSuccess) return fail(LUMAGPU_EHIP, std::string("macro launch: ") + hipGetErrorString(e));
    g->half = 1;
    g->half_sub = subcycle;
    return 0;
}

int lumagpu_force_collide_swap(lumagpu_grid *g) {
    if (int r = check_grid(g)) return r;
    if (!g->half) return fail(LUMAGPU_ESTATE, "no split step open: call lumagpu_stream_bc_macro first");
    g->half = 0;
    return end_step(g, g->half_sub, nullptr, nullptr);
}

int lumagpu_reset_forces(lumagpu_grid *g) {
    if (int r = check_grid(g)) return r;
    if (!g->d.forced) return fail(LUMAGPU_ESTATE, "grid was created without forcing (desc.forced = 0)");
    const StepArgs a = make_args(g);
    const unsigned nb = (unsigned)((g->ncell + 255) / 256);
    if (g->gravity_on) {
        // _LBM_resetForces (:1327-1337) with L_GRAVITY_ON: the gravity component from rho
        const double rr = 1.0 / std::pow(2.0, (double)g->d.level);
        if (g->melem == sizeof(double))
            k_gravity_reset<double><<<nb, 256, 0, g->stream>>>(a, g->force, g->gravity, rr, g->gravity_dir);
        else
            k_gravity_reset<Dual><<<nb, 256, 0, g->stream>>>(a, g->force, g->gravity, rr, g->gravity_dir);
        HIP_TRY(hipGetLastError());
        return 0;
    }
    // every component of every site zeroed, the force bits with it
    if (g->melem == sizeof(double))
        k_force_clear<double><<<nb, 256, 0, g->stream>>>(a, g->force, g->flags, g->D);
    else
        k_force_clear<Dual><<<nb, 256, 0, g->stream>>>(a, g->force, g->flags, g->D);
    HIP_TRY(hipGetLastError());
    for (auto &f : g->hflags) f &= (uint8_t)~kForcedBit;
    ibm_forget_flags(g);
    return 0;
}

int lumagpu_run(lumagpu_grid *g, int nsteps) {
    if (int r = check_grid(g)) return r;
    if (nsteps < 0) return fail(LUMAGPU_EINVAL, "negative step count");
    if (g->half) return fail(LUMAGPU_ESTATE, "a split step is open on this grid: finish it with lumagpu_force_collide_swap");
    int s = 0;
    if (g->fuse_shape >= 0) {
        if (g->parity == 0 && nsteps >= 3) {
            if (int r = apply_step(g, 0, nullptr, nullptr)) return r;
            ++s;
        }
        // long runs replay a captured pair of two-step launches (the two arrays trade places
        // twice, so the graph's pointers hold for every replay)
        const int reps = fused_ready(g) ? (nsteps - s) / 4 : 0;
        if (reps >= 2) {
            const int r = run_fused_graph(g, reps);
            if (r < 0) return r;
            if (r == 0) s += 4 * reps;
        }
        while (nsteps - s >= 2 && fused_ready(g)) {
            if (int r = fused_pair(g, nullptr, nullptr)) return r;
            s += 2;
        }
    }
    // Long runs of a hierarchy without host work per step replay a captured pair of steps: the
    // step alternates between the two AA kernels, so two steps are the period (the child's
    // four sub-steps included).  One eager step first settles the IB support flags.
    if (nsteps >= 8 && graph_ok(g)) {
        while (s < nsteps && (s == 0 || g->parity != 0)) {
            if (int r = apply_step(g, 0, nullptr, nullptr)) return r;
            ++s;
        }
        const int reps = (nsteps - s) / 2;
        if (reps >= 2) {
            int r = run_graph(g, reps);
            if (r < 0) return r;
            if (r == 0) s += 2 * reps;   // r > 0: not capturable here, steps continue eagerly
        }
    }
    for (; s < nsteps; ++s)
        if (int r = apply_step(g, 0, nullptr, nullptr)) return r;
    return 0;
}

int lumagpu_run_timed(lumagpu_grid *g, int nsteps, double *ms_total, double *ms_kernel_mean) {
    if (int r = check_grid(g)) return r;
    if (nsteps <= 0) return fail(LUMAGPU_EINVAL, "step count must be positive");
    std::vector<hipEvent_t> ev(2 * (size_t)nsteps);
    for (auto &e : ev) HIP_TRY(hipEventCreate(&e));
    int rc = 0;
    // launches: event pairs in use, each covering one step or (two-step kernel) two
    int nl = 0;
    for (int s = 0; s < nsteps && !rc;) {
        if (nsteps - s >= 2 && fused_ready(g)) {
            rc = fused_pair(g, ev[2 * nl], ev[2 * nl + 1]);
            s += 2;
        } else {
            rc = apply_step(g, 0, ev[2 * nl], ev[2 * nl + 1]);
            s += 1;
        }
        ++nl;
    }
    if (!rc) {
        HIP_TRY(hipEventSynchronize(ev[2 * nl - 1]));
        float tot = 0.f;
        HIP_TRY(hipEventElapsedTime(&tot, ev.front(), ev[2 * nl - 1]));
        double sum = 0.0;
        for (int l = 0; l < nl; ++l) {
            float k = 0.f;
            HIP_TRY(hipEventElapsedTime(&k, ev[2 * l], ev[2 * l + 1]));
            sum += k;
        }
        if (ms_total) *ms_total = tot;
        // per step: a two-step launch counts half its duration for each of its steps
        if (ms_kernel_mean) *ms_kernel_mean = sum / nsteps;
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    return rc;
}

int lumagpu_get_time(const lumagpu_grid *g, int *t) {
    if (int r = check_grid(g)) return r;
    if (t) *t = g->t;
    return 0;
}

int lumagpu_sync(lumagpu_grid *g) {
    if (int r = check_grid(g)) return r;
    HIP_TRY(hipStreamSynchronize(g->stream));
    return 0;
}

int lumagpu_moments(lumagpu_grid *g, double out[LUMAGPU_NUM_MOMENTS]) {
    if (int r = check_grid(g)) return r;
    if (!out) return fail(LUMAGPU_EINVAL, "null output");
    if (!g->has_state) return fail(LUMAGPU_ESTATE, "no state");
    HIP_TRY(launch_aux(g, kAuxMoments, nullptr, nullptr, g->partials));
    const dim3 gr = cell_grid(g->d.N, g->d.M, g->d.K);
    const size_t nblocks = (size_t)gr.x * gr.y * gr.z;
    k_sum_partials<<<1, 256, 0, g->stream>>>(g->partials, nblocks, g->moments_out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, g->moments_out, kMomentCount * sizeof(double), hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    return 0;
}

int lumagpu_moments_noalloc(lumagpu_grid *g, double *out_dev) {
    if (int r = check_grid(g)) return r;
    if (!out_dev) return fail(LUMAGPU_EINVAL, "null output");
    if (!g->has_state) return fail(LUMAGPU_ESTATE, "no state");
    const dim3 gr = cell_grid(g->d.N, g->d.M, g->d.K);
    const size_t nblocks = (size_t)gr.x * gr.y * gr.z;
    if (nblocks > kScratchBlocks)
        return fail(LUMAGPU_EINVAL, "grid too large for lumagpu_moments_noalloc: " + std::to_string(nblocks) +
                                        " row blocks of 256 cells, at most " + std::to_string(kScratchBlocks));
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != g->device) return fail(LUMAGPU_EINVAL, "call lumagpu_moments_noalloc with the grid's GPU current");
    double *partials = nullptr;
    if (int r = scratch_on(dev, &partials)) return r;
    // every block writes its own partials and the sum writes every output: nothing to clear
    HIP_TRY(launch_aux(g, kAuxMoments, nullptr, nullptr, partials));
    k_sum_partials<<<1, 256, 0, g->stream>>>(partials, nblocks, out_dev);
    HIP_TRY(hipGetLastError());
    return 0;
}

int lumagpu_footprint(const lumagpu_grid *g, size_t *bytes_per_cell, size_t *f_bytes) {
    if (int r = check_grid(g)) return r;
    if (bytes_per_cell) *bytes_per_cell = g->elem * g->Q;
    if (f_bytes) *f_bytes = g->ps * g->Q * g->elem;
    return 0;
}

}  // extern "C"
