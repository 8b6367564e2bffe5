// ibm_kernels.hpp -- immersed-boundary device kernels (FP64 in every storage form; Duals in the
// tangent form).
//
// Markers of all bodies on one grid are stored back to back in body order, each with up to S
// support slots (S = 9 in 2D, 27 in 3D).  The spread is a gather: every distinct support site
// owns the list of (marker, slot) entries that touch it, kept in LUMA's accumulation order
// (body, marker, support), so each site sums its contributions in exactly the order of
// ObjectManager::ibm_spread (ObjectManager_ops_ibm.cpp:621-689) -- no atomics, no colouring
// passes, and bit-for-bit the reference's sum.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "dual.hpp"
#include "lattice.hpp"

namespace lumagpu {

// The double arrays hold elements of the kernels' number type M: FP64, or (value, derivative)
// pairs on a tangent-form grid (support indices and counts stay integers).
struct IbmArgs {
    int nm, S, D;
    const int *scell;      // [nm*S] support cell index (device cell order)
    const double *sdelta;  // [nm*S] delta value of each support
    const int *ns;         // [nm] number of supports
    const double *vel;     // [nm*3] markerVel
    const double *eps;     // [nm] epsilon
    const double *ds;      // [nm] ds
    double *interp;        // [nm*4] interpRho, interpMom[3]
    double *force;         // [nm*3] marker force_xyz
    const double *rho;     // grid rho [cell]
    const double *u;       // grid u [d*ncell + cell]
    double *fxyz;          // grid force_xyz [d*ncell + cell]
    uint8_t *flags;        // grid flag bytes
    size_t ncell;
    int nsites;
    const int *sites;      // [nsites] distinct support cells
    const int *sptr;       // [nsites+1] entry ranges
    const int *sent;       // entries: m*S + slot
    int accumulate;        // spread onto the force field's current values (gravity reset) instead of 0
};

template <class M, class A>
__device__ __forceinline__ M *as(A *p) {
    return reinterpret_cast<M *>(p);
}
template <class M, class A>
__device__ __forceinline__ const M *as(const A *p) {
    return reinterpret_cast<const M *>(p);
}

// ibm_interpolate (:498-591) and ibm_computeForce (:598-614); one thread per marker, sums in
// support order.  local_area = 1 (IBMarker.cpp:61).  mode 0: both; 1: interpolation only;
// 2: the force from the interpolated values held since the last interpolation.  The loads of a
// chunk of supports are all issued before its ordered sum, so a marker waits on two memory
// round trips per chunk instead of two per support.  DD: the dimension when known at compile
// time (0: b.D).
template <class M = double, int DD = 0>
__device__ __forceinline__ void interp_force_one(const IbmArgs &b, int m, int mode) {
    constexpr int CH = sizeof(M) == sizeof(double) ? 9 : 3;
    const int D = DD ? DD : b.D;
    const M *sdelta = as<M>(b.sdelta), *grho = as<M>(b.rho), *gu = as<M>(b.u), *vel = as<M>(b.vel);
    M *interp = as<M>(b.interp), *force = as<M>(b.force);
    const M area = 1.0;
    M rho = 0.0, mom[3] = {0.0, 0.0, 0.0};
    if (mode != 2) {
        const int n = b.ns[m];
        const int *sc = b.scell + (size_t)m * b.S;
        const M *sd = sdelta + (size_t)m * b.S;
        for (int k0 = 0; k0 < n; k0 += CH) {
            int cell[CH];
            M dl[CH], r[CH], uv[CH][3];
#pragma unroll
            for (int j = 0; j < CH; ++j) {
                const int k = (k0 + j < n) ? k0 + j : k0;   // slots past the count repeat the first
                cell[j] = sc[k];
                dl[j] = sd[k];
            }
#pragma unroll
            for (int j = 0; j < CH; ++j) {
                r[j] = grho[cell[j]];
#pragma unroll
                for (int d = 0; d < 3; ++d) uv[j][d] = (d < D) ? gu[d * b.ncell + (size_t)cell[j]] : M(0.0);
            }
#pragma unroll
            for (int j = 0; j < CH; ++j) {
                if (k0 + j < n) {
                    rho += r[j] * dl[j] * area;
#pragma unroll
                    for (int d = 0; d < 3; ++d)
                        if (d < D) mom[d] += r[j] * uv[j][d] * dl[j] * area;
                }
            }
        }
        interp[4 * m] = rho;
        for (int d = 0; d < 3; ++d) interp[4 * m + 1 + d] = mom[d];
    } else {
        rho = interp[4 * m];
        for (int d = 0; d < 3; ++d) mom[d] = interp[4 * m + 1 + d];
    }
    if (mode != 1)
        for (int d = 0; d < D; ++d) force[3 * m + d] = M(2.0) * (mom[d] - rho * vel[3 * m + d]) / M(1.0);
}
template <class M = double>
__global__ void k_ibm_interp_force(IbmArgs b, int mode) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= b.nm) return;
    interp_force_one<M>(b, m, mode);
}

// The sum of ibm_spread (:621-689) at support site s: acc starts from the site's reset value
// and takes every (marker, slot) entry of the site in LUMA's accumulation order.  Entries are
// read in chunks whose loads are issued together, then summed in entry order.
template <class M, int D>
__device__ __forceinline__ void spread_sum(int s, int S, const int *sptr, const int *sent, const M *sdelta,
                                           const M *force, const M *eps, const M *ds, M acc[3]) {
    constexpr int CH = sizeof(M) == sizeof(double) ? 6 : 3;
    const int e0 = sptr[s], e1 = sptr[s + 1];
    for (int eb = e0; eb < e1; eb += CH) {
        int code[CH];
#pragma unroll
        for (int j = 0; j < CH; ++j) code[j] = sent[(eb + j < e1) ? eb + j : eb];
        M dl[CH], fm[CH][3], vw[CH], dsm[CH];
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            const int m = code[j] / S;
            dl[j] = sdelta[code[j]];
            vw[j] = eps[m];
            dsm[j] = ds[m];
#pragma unroll
            for (int d = 0; d < 3; ++d) fm[j][d] = (d < D) ? force[3 * m + d] : M(0.0);
        }
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            if (eb + j < e1) {
                const M volDepth = (D == 3) ? dsm[j] : M(1.0);
#pragma unroll
                for (int d = 0; d < 3; ++d)
                    if (d < D) acc[d] -= dl[j] * fm[j][d] * vw[j] * volDepth * dsm[j];
            }
        }
    }
}

// ibm_spread as a per-site gather; the site's force starts from the reset value 0
// (_LBM_resetForces, GridObj_ops_lbm_optimised.cpp:1327-1337) and the site is flagged so the
// step kernel adds it in the macro (+F/2) and the Guo term.
template <class M = double, int DD = 0>
__device__ __forceinline__ void spread_one(const IbmArgs &b, int s) {
    const size_t cell = (size_t)b.sites[s];
    M *fxyz = as<M>(b.fxyz);
    M acc[3] = {0.0, 0.0, 0.0};
    const int D = DD ? DD : b.D;
    if (b.accumulate)
        for (int d = 0; d < D; ++d) acc[d] = fxyz[d * b.ncell + cell];
    if (D == 3)
        spread_sum<M, 3>(s, b.S, b.sptr, b.sent, as<M>(b.sdelta), as<M>(b.force), as<M>(b.eps), as<M>(b.ds), acc);
    else
        spread_sum<M, 2>(s, b.S, b.sptr, b.sent, as<M>(b.sdelta), as<M>(b.force), as<M>(b.eps), as<M>(b.ds), acc);
    for (int d = 0; d < D; ++d) fxyz[d * b.ncell + cell] = acc[d];
    b.flags[cell] |= kForcedBit;
}
template <class M = double>
__global__ void k_ibm_spread(IbmArgs b) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= b.nsites) return;
    spread_one<M>(b, s);
}

// Clear the force and flag of sites that stop being supports.
template <class M = double>
__global__ void k_ibm_clear(IbmArgs b, int n, const int *cells) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const size_t cell = (size_t)cells[s];
    M *fxyz = as<M>(b.fxyz);
    for (int d = 0; d < b.D; ++d) fxyz[d * b.ncell + cell] = M(0.0);
    b.flags[cell] &= (uint8_t)~kForcedBit;
}

// FP64 sum over one body's markers of force * epsilon * volDepth * ds, the per-marker values
// LUMA writes to Body_<id>_LD (ObjectManager_ops_io.cpp:98-121).  One 64-wide wave: strided
// lane sums, then a fixed xor-shuffle tree, so the result does not depend on timing.  out[0..2]
// receives the sums; on a tangent-form grid out[3..5] their derivatives.
template <class M = double>
__global__ void k_ibm_body_force(IbmArgs b, int m0, int m1, double *out) {
    const int lane = threadIdx.x;
    const M *force = as<M>(b.force), *eps = as<M>(b.eps), *ds = as<M>(b.ds);
    M acc[3] = {0.0, 0.0, 0.0};
    for (int m = m0 + lane; m < m1; m += 64) {
        const M w = eps[m] * ((b.D == 3) ? ds[m] : M(1.0)) * ds[m];
        for (int d = 0; d < 3; ++d) acc[d] += force[3 * m + d] * w;
    }
    for (int d = 0; d < 3; ++d) {
        double v = value_of(acc[d]), t = deriv_of(acc[d]);
        for (int off = 32; off > 0; off >>= 1) {
            v += __shfl_xor(v, off, 64);
            t += __shfl_xor(t, off, 64);
        }
        if (lane == 0) {
            out[d] = v;
            if constexpr (std::is_same_v<M, Dual>) out[3 + d] = t;
        }
    }
}

}  // namespace lumagpu
