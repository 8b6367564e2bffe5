// launch.hpp -- host-side launchers for the step kernels, one set per storage form.
//
// The FP64 launchers live in lbm_f64.hip (built with -ffp-contract=off so the exact
// arithmetic keeps LUMA's rounding); the shifted forms in lbm_f32.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "aa_kernels.hpp"
#include "fused_kernels.hpp"
#include "ibm_kernels.hpp"

namespace lumagpu {

struct LaunchSpec {
    int Q;          // 9 or 19
    bool masked;    // labelled kernel
    bool odd;       // neighbour-access step
    bool store;     // write rho, u
    // labelled grid stepped as plain sites (k_aa_plain) plus a list of the others
    const int *mlist = nullptr;
    int nmlist = -1;   // < 0: one k_aa_masked launch over the grid
    bool forced = true;   // some site may carry a body force (IB bodies, gravity, a set force)
};

// Each returns hipSuccess or the launch error.
hipError_t launch_step_f64(const LaunchSpec &ls, const StepArgs &a, hipStream_t s);
hipError_t launch_step_f32s(const LaunchSpec &ls, const StepArgs &a, hipStream_t s);
hipError_t launch_step_f16s(const LaunchSpec &ls, const StepArgs &a, hipStream_t s);
hipError_t launch_step_dual(const LaunchSpec &ls, const StepArgs &a, hipStream_t s);
hipError_t launch_step_dual32(const LaunchSpec &ls, const StepArgs &a, hipStream_t s);

// two AA steps per launch on a periodic D3Q19 box, src = a.f -> dst (fused_kernels.hpp)
hipError_t launch_fused2_f64(int shape, const StepArgs &a, void *dst, int zc, bool store, hipStream_t s);
hipError_t launch_fused2_f32s(int shape, const StepArgs &a, void *dst, int zc, bool store, hipStream_t s);
hipError_t launch_fused2_f16s(int shape, const StepArgs &a, void *dst, int zc, bool store, hipStream_t s);

hipError_t launch_aux_f64(int op, int Q, const StepArgs &a, const double *p0, const double *p1,
                          double *out, int parity, int masked, hipStream_t s);
hipError_t launch_aux_f32s(int op, int Q, const StepArgs &a, const double *p0, const double *p1,
                           double *out, int parity, int masked, hipStream_t s);
hipError_t launch_aux_f16s(int op, int Q, const StepArgs &a, const double *p0, const double *p1,
                           double *out, int parity, int masked, hipStream_t s);
hipError_t launch_aux_dual(int op, int Q, const StepArgs &a, const double *p0, const double *p1,
                           double *out, int parity, int masked, hipStream_t s);
hipError_t launch_aux_dual32(int op, int Q, const StepArgs &a, const double *p0, const double *p1,
                             double *out, int parity, int masked, hipStream_t s);

// k_site_macro over a list of cells (IB support sites; every cell when cells is null), in the
// layout the next step reads
hipError_t launch_sites_f64(int Q, const StepArgs &a, bool odd, int n, const int *cells, int ib_force, hipStream_t s);
hipError_t launch_sites_f32s(int Q, const StepArgs &a, bool odd, int n, const int *cells, int ib_force, hipStream_t s);
hipError_t launch_sites_f16s(int Q, const StepArgs &a, bool odd, int n, const int *cells, int ib_force, hipStream_t s);
hipError_t launch_sites_dual(int Q, const StepArgs &a, bool odd, int n, const int *cells, int ib_force, hipStream_t s);
hipError_t launch_sites_dual32(int Q, const StepArgs &a, bool odd, int n, const int *cells, int ib_force, hipStream_t s);

template <class P>
hipError_t launch_sites_p(int Q, const StepArgs &a, bool odd, int n, const int *cells, int ib_force, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const dim3 b(128), g((n + 127) / 128);
    if (Q == 9) {
        if (odd) k_site_macro<9, P, true><<<g, b, 0, s>>>(a, n, cells, ib_force);
        else k_site_macro<9, P, false><<<g, b, 0, s>>>(a, n, cells, ib_force);
    } else {
        if (odd) k_site_macro<19, P, true><<<g, b, 0, s>>>(a, n, cells, ib_force);
        else k_site_macro<19, P, false><<<g, b, 0, s>>>(a, n, cells, ib_force);
    }
    return hipGetLastError();
}

// ibm_interpolate (:498-591) with the sweep-1 macro of each support computed where it is
// needed: one lane per (marker, support), a marker's S supports in consecutive lanes of one
// wave (7 markers per wave in 2D, 2 in 3D).  Each lane computes its site's rho, u from the
// populations exactly as the site-macro pass does (and stores them, so the macro arrays hold
// what that pass writes; lanes sharing a site store the same bits), then the group sums its
// products in support order through lane shuffles and the first lane writes the interpolated
// values and, for mode 0, the marker force (ibm_computeForce, :598-614).  Every memory access
// of a marker is issued in parallel across its lanes instead of one support after another.
__device__ __forceinline__ double shfl_m(double v, int src) { return __shfl(v, src, 64); }
__device__ __forceinline__ Dual shfl_m(const Dual &v, int src) {
    return Dual(__shfl(v.v, src, 64), __shfl(v.d, src, 64));
}
template <int Q, class P, bool ODD>
__global__ __launch_bounds__(256) void k_ibm_interp_sites(StepArgs a, IbmArgs b, int mode) {
    using L = Lattice<Q>;
    using M = typename P::M;
    constexpr int D = L::D, S = (D == 3) ? 27 : 9, MPW = 64 / S;
    const int lane = threadIdx.x & 63;
    const int grp = lane / S, k = lane - grp * S;
    const int m = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * MPW + grp;
    const bool live = grp < MPW && m < b.nm;
    const int n = live ? b.ns[m] : 0;
    const M area = 1.0;
    M pr = M(0.0), pm[3] = {M(0.0), M(0.0), M(0.0)};
    if (k < n) {
        const size_t cell = (size_t)b.scell[m * S + k];
        const M dl = as<M>(b.sdelta)[m * S + k];
        M r, u[3];
        if (site_macro_value<Q, P, ODD>(a, cell, 0, r, u)) {
            store_macro<D>(a, cell, r, u);
        } else {   // sites of other types keep the rho, u their arrays hold
            r = as<M>(b.rho)[cell];
            for (int d = 0; d < D; ++d) u[d] = as<M>(b.u)[d * b.ncell + cell];
        }
        pr = r * dl * area;
        for (int d = 0; d < D; ++d) pm[d] = r * u[d] * dl * area;
    }
    M rho = 0.0, mom[3] = {M(0.0), M(0.0), M(0.0)};
    const int base = grp * S;
    for (int kk = 0; kk < S; ++kk) {
        const M tr = shfl_m(pr, base + kk);
        M tm[3];
        for (int d = 0; d < D; ++d) tm[d] = shfl_m(pm[d], base + kk);
        if (kk < n) {
            rho += tr;
            for (int d = 0; d < D; ++d) mom[d] += tm[d];
        }
    }
    if (!live || k != 0) return;
    M *interp = as<M>(b.interp);
    interp[4 * m] = rho;
    for (int d = 0; d < 3; ++d) interp[4 * m + 1 + d] = mom[d];
    if (mode == 0) {
        const M *vel = as<M>(b.vel);
        M *force = as<M>(b.force);
        for (int d = 0; d < D; ++d) force[3 * m + d] = M(2.0) * (mom[d] - rho * vel[3 * m + d]) / M(1.0);
    }
}

hipError_t launch_ibm_interp_f64(int Q, const StepArgs &a, const IbmArgs &b, bool odd, int mode, hipStream_t s);
hipError_t launch_ibm_interp_f32s(int Q, const StepArgs &a, const IbmArgs &b, bool odd, int mode, hipStream_t s);
hipError_t launch_ibm_interp_f16s(int Q, const StepArgs &a, const IbmArgs &b, bool odd, int mode, hipStream_t s);
hipError_t launch_ibm_interp_dual(int Q, const StepArgs &a, const IbmArgs &b, bool odd, int mode, hipStream_t s);
hipError_t launch_ibm_interp_dual32(int Q, const StepArgs &a, const IbmArgs &b, bool odd, int mode, hipStream_t s);

template <class P>
hipError_t launch_ibm_interp_p(int Q, const StepArgs &a, const IbmArgs &b, bool odd, int mode, hipStream_t s) {
    if (b.nm <= 0) return hipSuccess;
    const int mpb = 4 * (64 / (Q == 9 ? 9 : 27));   // markers per 256-thread block
    const dim3 g((unsigned)((b.nm + mpb - 1) / mpb)), blk(256);
    if (Q == 9) {
        if (odd) k_ibm_interp_sites<9, P, true><<<g, blk, 0, s>>>(a, b, mode);
        else k_ibm_interp_sites<9, P, false><<<g, blk, 0, s>>>(a, b, mode);
    } else {
        if (odd) k_ibm_interp_sites<19, P, true><<<g, blk, 0, s>>>(a, b, mode);
        else k_ibm_interp_sites<19, P, false><<<g, blk, 0, s>>>(a, b, mode);
    }
    return hipGetLastError();
}

// Edge/corner regularised velocity sites: phase 0 = neighbour densities (before the step
// kernel), phase 1 = the sites themselves (after it)
hipError_t launch_edges_f64(int Q, const StepArgs &a, const EdgeArgs &e, bool odd, int phase, int store, hipStream_t s);
hipError_t launch_edges_f32s(int Q, const StepArgs &a, const EdgeArgs &e, bool odd, int phase, int store, hipStream_t s);
hipError_t launch_edges_f16s(int Q, const StepArgs &a, const EdgeArgs &e, bool odd, int phase, int store, hipStream_t s);
hipError_t launch_edges_dual(int Q, const StepArgs &a, const EdgeArgs &e, bool odd, int phase, int store, hipStream_t s);
hipError_t launch_edges_dual32(int Q, const StepArgs &a, const EdgeArgs &e, bool odd, int phase, int store, hipStream_t s);

template <int Q, class P, bool ODD>
void launch_edges_qo(const StepArgs &a, const EdgeArgs &e, int phase, int store, hipStream_t s) {
    if (phase == 0) k_edge_rho<Q, P, ODD><<<(2 * e.n + 127) / 128, 128, 0, s>>>(a, e);
    else k_edge_sites<Q, P, ODD><<<(e.n + 127) / 128, 128, 0, s>>>(a, e, store);
}
template <class P>
hipError_t launch_edges_p(int Q, const StepArgs &a, const EdgeArgs &e, bool odd, int phase, int store, hipStream_t s) {
    if (e.n <= 0) return hipSuccess;
    if (Q == 9) {
        if (odd) launch_edges_qo<9, P, true>(a, e, phase, store, s);
        else launch_edges_qo<9, P, false>(a, e, phase, store, s);
    } else {
        if (odd) launch_edges_qo<19, P, true>(a, e, phase, store, s);
        else launch_edges_qo<19, P, false>(a, e, phase, store, s);
    }
    return hipGetLastError();
}

// z-slab halo exchange (D3Q19 periodic grids with halo planes), phase 0 before / 1 after the odd step
hipError_t launch_zslab_f64(const StepArgs &a, int phase, hipStream_t s);
hipError_t launch_zslab_f32s(const StepArgs &a, int phase, hipStream_t s);
hipError_t launch_zslab_f16s(const StepArgs &a, int phase, hipStream_t s);

template <class P>
hipError_t launch_zslab_p(const StepArgs &a, int phase, hipStream_t s) {
    const size_t plane = (size_t)a.nx * a.ny;
    k_zslab_exchange<P><<<dim3((unsigned)((plane + 255) / 256), 10), 256, 0, s>>>(a, phase);
    return hipGetLastError();
}

// aux ops
enum AuxOp : int {
    kAuxInitFeq = 0,     // p0 = rho (LUMA id), p1 = u (LUMA id)
    kAuxUpload = 1,      // p0 = f AoS
    kAuxDownload = 2,    // out = f AoS
    kAuxMoments = 3,     // out = per-block partials
    kAuxExtrapRight = 4, // out = ExtrapolateRight side buffer
    kAuxDownloadTan = 5, // out = derivative part of f AoS (tangent form)
};

// Cells between consecutive population planes beyond the grid (rounded to 256): the planes'
// bases then do not all fall on one large power-of-two stride, which the 19 streams of a step
// kernel read and write in step.  Measured at 256^3 (planes of 64 MiB in FP32) against no
// padding: one-step kernels FP64 +4.7%, FP32-shifted +1.8%, FP16-shifted +4.9%, two-step
// kernel +0.5-0.9%; 3D channel 384 x 128 x 128 +2-3% in FP32-shifted.  (4096 cells lost 1.7% on
// the two-step kernel; 256 and 8192 were between.)
constexpr size_t kPlanePad = 2048;

// One block per row segment: 256 threads, or the widest of 192 and 128 that divides the row
// when 256 does not (a 384-cell row in two blocks of 192 rather than 256 + a half-idle 256).
inline dim3 cell_block(int nx) {
    if (nx < 256) return dim3(((nx + 63) / 64) * 64, 1, 1);
    for (int bx : {256, 192, 128})
        if (nx % bx == 0) return dim3(bx, 1, 1);
    return dim3(256, 1, 1);
}
inline dim3 cell_grid(int nx, int ny, int nz) {
    const dim3 b = cell_block(nx);
    return dim3((nx + b.x - 1) / b.x, ny, nz);
}

template <int Q, class P>
hipError_t launch_step_q(const LaunchSpec &ls, const StepArgs &a, hipStream_t s) {
    const dim3 b = cell_block(a.nx), g = cell_grid(a.nx, a.ny, a.nz - 2 * a.zoff);
    if (ls.masked && ls.nmlist >= 0) {
        // AA steps touch only a site's own set of slots, so the two disjoint site sets can run
        // one after the other
        const dim3 lg((unsigned)((ls.nmlist + 255) / 256));
        if (ls.odd) {
            if (ls.forced) {
                if (ls.store) k_aa_plain<Q, P, true, true, true><<<g, b, 0, s>>>(a);
                else k_aa_plain<Q, P, true, false, true><<<g, b, 0, s>>>(a);
            } else {
                if (ls.store) k_aa_plain<Q, P, true, true, false><<<g, b, 0, s>>>(a);
                else k_aa_plain<Q, P, true, false, false><<<g, b, 0, s>>>(a);
            }
            if (ls.nmlist > 0) {
                if (ls.store) k_aa_masked_list<Q, P, true, true><<<lg, 256, 0, s>>>(a, ls.mlist, ls.nmlist);
                else k_aa_masked_list<Q, P, true, false><<<lg, 256, 0, s>>>(a, ls.mlist, ls.nmlist);
            }
        } else {
            if (ls.forced) {
                if (ls.store) k_aa_plain<Q, P, false, true, true><<<g, b, 0, s>>>(a);
                else k_aa_plain<Q, P, false, false, true><<<g, b, 0, s>>>(a);
            } else {
                if (ls.store) k_aa_plain<Q, P, false, true, false><<<g, b, 0, s>>>(a);
                else k_aa_plain<Q, P, false, false, false><<<g, b, 0, s>>>(a);
            }
            if (ls.nmlist > 0) {
                if (ls.store) k_aa_masked_list<Q, P, false, true><<<lg, 256, 0, s>>>(a, ls.mlist, ls.nmlist);
                else k_aa_masked_list<Q, P, false, false><<<lg, 256, 0, s>>>(a, ls.mlist, ls.nmlist);
            }
        }
    } else if (ls.masked) {
        if (ls.odd) {
            if (ls.store) k_aa_masked<Q, P, true, true><<<g, b, 0, s>>>(a);
            else k_aa_masked<Q, P, true, false><<<g, b, 0, s>>>(a);
        } else {
            if (ls.store) k_aa_masked<Q, P, false, true><<<g, b, 0, s>>>(a);
            else k_aa_masked<Q, P, false, false><<<g, b, 0, s>>>(a);
        }
    } else if (!P::kShift && a.ncell * sizeof(typename P::S) < (size_t(1) << 32)) {
        // FP64 planes below 4 GiB: 32-bit offsets from the plane bases (measured at 256^3:
        // 920 us against 947 us; the shifted forms are faster with 64-bit addresses)
        if (ls.odd) {
            if (ls.store) k_aa_periodic<Q, P, true, true, true><<<g, b, 0, s>>>(a);
            else k_aa_periodic<Q, P, true, false, true><<<g, b, 0, s>>>(a);
        } else {
            if (ls.store) k_aa_periodic<Q, P, false, true, true><<<g, b, 0, s>>>(a);
            else k_aa_periodic<Q, P, false, false, true><<<g, b, 0, s>>>(a);
        }
    } else {
        if (ls.odd) {
            if (ls.store) k_aa_periodic<Q, P, true, true><<<g, b, 0, s>>>(a);
            else k_aa_periodic<Q, P, true, false><<<g, b, 0, s>>>(a);
        } else {
            if (ls.store) k_aa_periodic<Q, P, false, true><<<g, b, 0, s>>>(a);
            else k_aa_periodic<Q, P, false, false><<<g, b, 0, s>>>(a);
        }
    }
    return hipGetLastError();
}

template <class P>
hipError_t launch_step_p(const LaunchSpec &ls, const StepArgs &a, hipStream_t s) {
    if (ls.Q == 9) return launch_step_q<9, P>(ls, a, s);
    if (ls.Q == 19) return launch_step_q<19, P>(ls, a, s);
    return hipErrorInvalidValue;
}

template <int Q, class P>
hipError_t launch_aux_q(int op, const StepArgs &a, const double *p0, const double *p1, double *out,
                        int parity, int masked, hipStream_t s) {
    const dim3 b = cell_block(a.nx), g = cell_grid(a.nx, a.ny, a.nz);
    switch (op) {
        case kAuxInitFeq: k_init_feq<Q, P><<<g, b, 0, s>>>(a, p0, p1); break;
        case kAuxUpload: k_upload_aos<Q, P><<<g, b, 0, s>>>(a, p0); break;
        case kAuxDownload: k_download_aos<Q, P><<<g, b, 0, s>>>(a, out, parity, masked); break;
        case kAuxDownloadTan:
            if constexpr (!std::is_same_v<typename P::M, Dual>) return hipErrorInvalidValue;
            else k_download_aos<Q, P, true><<<g, b, 0, s>>>(a, out, parity, masked);
            break;
        case kAuxMoments: k_moments<Q, P><<<g, b, 0, s>>>(a, parity, masked, out); break;
        case kAuxExtrapRight: {
            const dim3 bb(64), gg((a.ny + 63) / 64, a.nz);
            k_extrap_right<Q, P><<<gg, bb, 0, s>>>(a, parity, out);
            break;
        }
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class P>
hipError_t launch_aux_p(int op, int Q, const StepArgs &a, const double *p0, const double *p1, double *out,
                        int parity, int masked, hipStream_t s) {
    if (Q == 9) return launch_aux_q<9, P>(op, a, p0, p1, out, parity, masked, s);
    if (Q == 19) return launch_aux_q<19, P>(op, a, p0, p1, out, parity, masked, s);
    return hipErrorInvalidValue;
}

}  // namespace lumagpu
