// aa_kernels.hpp -- fused stream-collide kernels for the LBM time step, AA-pattern in place.
//
// One population array A (SoA: A[v*ps + cell], cell = x + nx*(y + ny*z), x = LUMA's i) is
// updated in place.  Two kernels alternate:
//   odd  kernel (array in "post-local" layout):  incoming f_v(x) = A[opp v][x - c_v];
//                 outgoing f*_v(x) -> A[v][x + c_v]
//   even kernel (array in "post-neighbour" layout): incoming f_v(x) = A[v][x];
//                 outgoing f*_v(x) -> A[opp v][x]
// Each population is read once and written once per step; no second array exists.  The
// host uploads LUMA's post-collision f (GridObj::f after the swap) in the post-local layout
// so the first step is an odd step.  Halfway bounce-back (LUMA's HWBB pull,
// GridObj_ops_lbm_optimised.cpp:238-243) is folded into the odd kernel: a population whose
// destination is solid is written to the cell's own opposite slot, where both the next even
// step and the odd step's own read find it.
//
// Per-site arithmetic has two forms:
//   * exact (FP64 storage): the expression trees of LUMA's _LBM_macro_opt (:800-893),
//     _LBM_equilibrium_opt (:674-705), _LBM_forceGrid_opt (:928-990), _LBM_collide_opt
//     (:765-790) and _LBM_regularised_opt (:313-510), operand for operand.  The translation
//     unit that instantiates it is compiled with -ffp-contract=off, so results match a
//     non-FMA x86-64 build of LUMA bit for bit.
//   * shifted (FP32 / FP16 storage): populations are stored as f_v - w_v*rho0 (rho0 = 1) and
//     the collision works on those deviations in FP32; moments and equilibrium are the same
//     polynomials rearranged around rho0, with every fused multiply-add written out.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <type_traits>
#include <utility>

#include "dual.hpp"
#include "ibm_kernels.hpp"
#include "lattice.hpp"

namespace lumagpu {

// S: population storage type; T: per-site arithmetic type; M: type of the macro arrays,
// force field and side buffers (FP64 in every form, a Dual in the tangent form).
struct P64 {
    using S = double;
    using T = double;
    using M = double;
    static constexpr bool kShift = false;
};
struct P32S {
    using S = float;
    using T = float;
    using M = double;
    static constexpr bool kShift = true;
};
struct P16S {
    using S = _Float16;
    using T = float;
    using M = double;
    static constexpr bool kShift = true;
};
// Tangent form: the FP64 exact arithmetic on Duals (value and derivative along one
// parameter direction); values equal the FP64 form bit for bit.
struct PDual {
    using S = Dual;
    using T = Dual;
    using M = Dual;
    static constexpr bool kShift = false;
};
// Shifted tangent form: the FP32-shifted storage and collision on float Duals, FP64 Duals for
// the macro arrays and force field.
struct P32T {
    using S = DualF;
    using T = DualF;
    using M = Dual;
    static constexpr bool kShift = true;
};

// One refined region of a grid (GridObj::subGrid[i]): its populations (coalesce source, read
// in the post-local layout), extents, CoarseLims{X,Y,Z}[eMinimum/eMaximum] on this grid, and
// macro arrays (TL2F copy, :850-891).
struct ChildRef {
    const void *f;
    size_t ps;
    int nx, ny, nz;
    int lo[3], hi[3];
    double *rho, *u;
    size_t ncell;
};

struct StepArgs {
    void *f;                      // population array
    size_t ps;                    // population stride in elements
    int nx, ny, nz;               // x = LUMA i (N_lim), y = j (M_lim), z = k (K_lim)
    const uint8_t *flags;         // per-cell flag byte (masked kernels only)
    const uint32_t *plain;        // plain-site bitmap of a labelled grid (k_plain_bits), or null
    // The double arrays below hold P::M elements: FP64, or (value, derivative) pairs in the
    // tangent form; kernels address them through mv<M>.
    const double *force;          // force_xyz, SoA [d*ncell + cell], read where flagged
    const double *uin;            // inlet profile [a*ny + j]: ux_in, uy_in, uz_in
    double *rho_out;              // macro output [cell] (store kernels)
    double *u_out;                // macro output [d*ncell + cell]
    size_t ncell;
    double omega, rho0, rho_in;
    double rho_p;                 // pressure sites: L_RHOIN + pd2dlbm(L_PRESSURE_DELTA)
    double d_omega;               // derivative of omega along the seeded direction (FP64 tangent form)
    double cs2, den2;             // cs*cs and 2*cs^2*cs^2, evaluated as LUMA does
    float om_hi, om_lo;           // omega as a float pair (shifted forms)
    int regularised;              // L_REGULARISED_BOUNDARIES
    int fz_compat;                // 3D macro quirk (:833): F_z/2 added to x-momentum
    int all_forced;               // every site carries force_xyz (gravity reset on this level)
    int sweep1_zero;              // the IB reset zeroes force_xyz at the start of the step
    // L_VELOCITY_RAMP: inlet factors of this step, (1 - cos(pi t / T)) / 2 at t = (step+1) dt for
    // the regularised sites and t = step dt for forced-equilibrium sources (1 when off)
    int ramp_on;
    double ramp_reg, ramp_feq;
    // --- refinement coupling (GridObj::parentGrid / subGrid) -----------------------------
    const void *pf;               // parent population array (explode source)
    size_t pps;
    int pnx, pny, pnz, pparity;
    const uint8_t *pflags;
    int clx, cly, clz;            // this grid's CoarseLims{X,Y,Z}[eMinimum] on the parent
    const ChildRef *kids;         // GridObj::subGrid: the refined regions (coalesce sources), device table
    int nkids;
    int subcycle;
    // --- ExtrapolateRight sources: xr[v*(ny*nz) + y + ny*z] = f*_v two cells left of (N-1, y, z)
    const double *xr;
    // the same buffer for the next step, written by this step's sites at x = N-3 with their
    // post-collision populations as stored (null: the next step's pre-pass rebuilds it)
    double *xr_next;
    // --- z-slab with halo planes: the step updates planes zoff .. nz-1-zoff only ------------
    int zoff;
    // --- refined sites read by sites other than TL2F (a region touching a velocity, pressure or
    // outlet boundary): LUMA never updates a refined site, so those reads see its initial
    // populations, kept here in storage form at stale[v * nstale + stale_idx[cell]]
    const void *stale;
    const int *stale_idx;
    int nstale;
    // --- ibm_spread done by the step kernel (rigid bodies whose supports did not move): a
    // forced site with ib_site_of[cell] >= 0 sums its spread force (spread_sum, k_ibm_spread's
    // order), writes it to force_xyz and uses it.  Marker arrays hold P::M elements.
    const int *ib_site_of;
    const int *ib_sptr, *ib_sent;
    const double *ib_sdelta, *ib_force, *ib_eps, *ib_ds;
    int ib_S, ib_accumulate;
};


// ----------------------------------------------------------------------------------------
// storage helpers
// Population traffic is streamed: every slot is read once and written once per step.  Row-
// aligned accesses (c_x = 0, and every access of the even kernel) use non-temporal loads and
// stores; the x-shifted accesses of the odd kernel share cache lines between neighbouring
// waves and stay cached, since a non-temporal fetch there is fetched again by the next wave.
// Measured on gfx950 at 256^3 FP32: all cached 5.3 TB/s, all non-temporal 5.2 TB/s, this
// split 6.0 TB/s.
// The cache policy is a template argument: a run-time flag, even one that folds after
// unrolling, lets the compiler merge both branches into one load and drop the hint.
template <class P, bool NT = true>
__device__ __forceinline__ typename P::T ld(const typename P::S *p) {
    if constexpr (std::is_same_v<typename P::S, Dual> || std::is_same_v<typename P::S, DualF>) return *p;
    else if constexpr (NT) return static_cast<typename P::T>(__builtin_nontemporal_load(p));
    else return static_cast<typename P::T>(*p);
}
template <class P, bool NT = true>
__device__ __forceinline__ void st(typename P::S *p, typename P::T v) {
    if constexpr (std::is_same_v<typename P::S, Dual> || std::is_same_v<typename P::S, DualF>) *p = v;
    else if constexpr (NT) __builtin_nontemporal_store(static_cast<typename P::S>(v), p);
    else *p = static_cast<typename P::S>(v);
}

// omega in the exact arithmetic type: a Dual carries the seeded derivative d_omega
template <class T>
__device__ __forceinline__ T omega_of(const StepArgs &a) {
    if constexpr (std::is_same_v<T, Dual>) return Dual(a.omega, a.d_omega);
    else return T(a.omega);
}

// an FP64 array of StepArgs viewed as elements of the form's macro type M
template <class M>
__device__ __forceinline__ M *mv(double *p) {
    return reinterpret_cast<M *>(p);
}
template <class M>
__device__ __forceinline__ const M *mv(const double *p) {
    return reinterpret_cast<const M *>(p);
}

// The body force of a forced site: spread here (see StepArgs::ib_site_of) or read from force_xyz
template <int D, class M>
__device__ __forceinline__ void site_force(const StepArgs &a, size_t cell, unsigned fl, M Fv[3]) {
    const int sidx = (a.ib_site_of && (fl & kForcedBit)) ? a.ib_site_of[cell] : -1;
    M *fx = mv<M>(const_cast<double *>(a.force));
    if (sidx >= 0) {
        M acc[3] = {M(0.0), M(0.0), M(0.0)};
        if (a.ib_accumulate)
            for (int d = 0; d < D; ++d) acc[d] = fx[d * a.ncell + cell];
        spread_sum<M, D>(sidx, a.ib_S, a.ib_sptr, a.ib_sent, mv<M>(a.ib_sdelta), mv<M>(a.ib_force), mv<M>(a.ib_eps),
                         mv<M>(a.ib_ds), acc);
        for (int d = 0; d < D; ++d) fx[d * a.ncell + cell] = acc[d];
        Fv[0] = acc[0];
        Fv[1] = acc[1];
        Fv[2] = (D == 3) ? acc[2] : M(0.0);
    } else {
        Fv[0] = fx[cell];
        Fv[1] = fx[a.ncell + cell];
        Fv[2] = (D == 3) ? fx[2 * a.ncell + cell] : M(0.0);
    }
}

// Compile-time loop: f(std::integral_constant<int, i>) for i = 0..N-1.
template <class F, int... Is>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, Is...>) {
    (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F &&f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

__device__ __forceinline__ size_t cell_of(int x, int y, int z, int nx, int ny) {
    return (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * (size_t)z);
}

// Periodic neighbour coordinates of (x,y,z): n[a][c+1] for displacement c in {-1,0,1}.
struct Nbr {
    int xs[3], ys[3], zs[3];
    __device__ __forceinline__ void init(int x, int y, int z, int nx, int ny, int nz) {
        xs[0] = (x == 0) ? nx - 1 : x - 1;
        xs[1] = x;
        xs[2] = (x == nx - 1) ? 0 : x + 1;
        ys[0] = (y == 0) ? ny - 1 : y - 1;
        ys[1] = y;
        ys[2] = (y == ny - 1) ? 0 : y + 1;
        zs[0] = (z == 0) ? nz - 1 : z - 1;
        zs[1] = z;
        zs[2] = (z == nz - 1) ? 0 : z + 1;
    }
    __device__ __forceinline__ size_t at(int dx, int dy, int dz, int nx, int ny) const {
        return cell_of(xs[dx + 1], ys[dy + 1], zs[dz + 1], nx, ny);
    }
    __device__ __forceinline__ uint32_t at32(int dx, int dy, int dz, int nx, int ny) const {
        return (uint32_t)xs[dx + 1] + (uint32_t)nx * ((uint32_t)ys[dy + 1] + (uint32_t)ny * (uint32_t)zs[dz + 1]);
    }
};

// Element idx of a population plane.  With O32 the byte offset is a 32-bit value added to the
// plane's (wave-uniform) base, which lets every load and store of a site share one offset
// register (scalar base + vector offset addressing); valid while a plane is below 4 GiB.
template <class S, bool O32>
__device__ __forceinline__ S *elem_at(S *plane, size_t idx) {
    if constexpr (O32)
        return reinterpret_cast<S *>(reinterpret_cast<char *>(plane) + (uint32_t)((uint32_t)idx * (uint32_t)sizeof(S)));
    else
        return plane + idx;
}

// ----------------------------------------------------------------------------------------
// exact (LUMA-order) per-site arithmetic

// _LBM_equilibrium_opt, GridObj_ops_lbm_optimised.cpp:674-705
template <int Q, class T>
__device__ __forceinline__ T feq_exact(int v, T rho, const T *u, T cs2, T den2) {
    using L = Lattice<Q>;
    const int c0 = L::c(v, 0), c1 = L::c(v, 1), c2 = L::c(v, 2);
    T A, B;
    if constexpr (L::D == 3) {
        A = (T(c0) * u[0]) + (T(c1) * u[1]) + (T(c2) * u[2]);
        B = (T(c0 * c0) - cs2) * (u[0] * u[0]) + (T(c1 * c1) - cs2) * (u[1] * u[1]) +
            (T(c2 * c2) - cs2) * (u[2] * u[2]) + T(2 * c0 * c1) * u[0] * u[1] +
            T(2 * c0 * c2) * u[0] * u[2] + T(2 * c1 * c2) * u[1] * u[2];
    } else {
        A = (T(c0) * u[0]) + (T(c1) * u[1]);
        B = (T(c0 * c0) - cs2) * (u[0] * u[0]) + (T(c1 * c1) - cs2) * (u[1] * u[1]) +
            T(2 * c0 * c1) * u[0] * u[1];
    }
    return rho * T(L::w(v)) * (T(1.0) + (A / cs2) + (B / den2));
}

// _LBM_macro_opt, :800-847
template <int Q, class T, class FT>
__device__ __forceinline__ void macro_exact(const T *fin, const FT *F, bool fz_compat, T &rho,
                                            T *u) {
    using L = Lattice<Q>;
    T r = T(0), rx = T(0), ry = T(0), rz = T(0);
#pragma unroll
    for (int v = 0; v < Q; ++v) {
        r += fin[v];
        rx += T(L::c(v, 0)) * fin[v];
        ry += T(L::c(v, 1)) * fin[v];
        if constexpr (L::D == 3) rz += T(L::c(v, 2)) * fin[v];
    }
    if (F) {
        rx += T(0.5) * T(F[0]);
        ry += T(0.5) * T(F[1]);
        if constexpr (L::D == 3) {
            if (fz_compat) rx += T(0.5) * T(F[2]);
            else rz += T(0.5) * T(F[2]);
        }
    }
    u[0] = rx / r;
    u[1] = ry / r;
    u[2] = T(0);
    if constexpr (L::D == 3) u[2] = rz / r;
    rho = r;
}

// _LBM_forceGrid_opt, :928-990 (one direction)
template <int Q, class T>
__device__ __forceinline__ T guo_exact(int v, const T *u, const T *F, T omega, T cs2) {
    using L = Lattice<Q>;
    T beta = T(0);
    const T lambda = (T(1) - T(0.5) * omega) * (T(L::w(v)) / cs2);
#pragma unroll
    for (int d = 0; d < L::D; ++d) beta += (T(L::c(v, d)) * u[d]);
    beta = beta * (T(1) / cs2);
    T fi = T(0);
#pragma unroll
    for (int d = 0; d < L::D; ++d) fi += T(F[d]) * (T(L::c(v, d)) * (T(1) + beta) - u[d]);
    fi *= lambda;
    return fi;
}

// _LBM_collide_opt (BGK), :765-790, one population at a time; F == nullptr means no force
// term.  Kernels that store each population as soon as it is collided keep only the incoming
// set live.
template <int Q, class T>
struct CollideExact {
    T rho, u[3], omega, cs2, den2;
    const T *F;
    __device__ __forceinline__ T operator()(int v, T fin) const {
        const T feq = feq_exact<Q, T>(v, rho, u, cs2, den2);
        if (F) return fin + (omega * (feq - fin) + guo_exact<Q, T>(v, u, F, omega, cs2));
        return fin + omega * (feq - fin);
    }
    // v and its opposite o = v + 1 (LUMA lists opposites in pairs), bit for bit the two
    // single-population results: A(o) = -A(v) and B(o) = B(v) exactly (the same products,
    // negated or not), and division is sign-symmetric, so the pair shares its two divisions;
    // a zero of either sign added to 1.0 gives 1.0.
    template <int V>
    __device__ __forceinline__ void pair(T fin_v, T fin_o, T &out_v, T &out_o) const {
        using L = Lattice<Q>;
        static_assert(L::opp(V) == V + 1, "opposites must be adjacent");
        constexpr int c0 = L::c(V, 0), c1 = L::c(V, 1), c2 = L::c(V, 2);
        T A, B;
        if constexpr (L::D == 3) {
            A = (T(c0) * u[0]) + (T(c1) * u[1]) + (T(c2) * u[2]);
            B = (T(c0 * c0) - cs2) * (u[0] * u[0]) + (T(c1 * c1) - cs2) * (u[1] * u[1]) +
                (T(c2 * c2) - cs2) * (u[2] * u[2]) + T(2 * c0 * c1) * u[0] * u[1] +
                T(2 * c0 * c2) * u[0] * u[2] + T(2 * c1 * c2) * u[1] * u[2];
        } else {
            A = (T(c0) * u[0]) + (T(c1) * u[1]);
            B = (T(c0 * c0) - cs2) * (u[0] * u[0]) + (T(c1 * c1) - cs2) * (u[1] * u[1]) +
                T(2 * c0 * c1) * u[0] * u[1];
        }
        const T a = A / cs2, b = B / den2;
        const T feq_v = rho * T(L::w(V)) * (T(1.0) + a + b);
        const T feq_o = rho * T(L::w(V + 1)) * (T(1.0) + (-a) + b);
        if (F) {
            out_v = fin_v + (omega * (feq_v - fin_v) + guo_exact<Q, T>(V, u, F, omega, cs2));
            out_o = fin_o + (omega * (feq_o - fin_o) + guo_exact<Q, T>(V + 1, u, F, omega, cs2));
        } else {
            out_v = fin_v + omega * (feq_v - fin_v);
            out_o = fin_o + omega * (feq_o - fin_o);
        }
    }
};

template <int Q, class T>
__device__ __forceinline__ void collide_exact(const T *fin, T rho, const T *u, const T *F,
                                              T omega, T cs2, T den2, T *fout) {
    const CollideExact<Q, T> c{rho, {u[0], u[1], u[2]}, omega, cs2, den2, F};
#pragma unroll
    for (int v = 0; v < Q; ++v) fout[v] = c(v, fin[v]);
}

// The reconstruction shared by the face conditions (:432-508): unknown populations (those
// entering through the face) by non-equilibrium bounce-back, then every population rebuilt from
// feq(rho, u) and the non-equilibrium stress of the result.
template <int Q, class T>
__device__ __forceinline__ void regularise_face_set(T *fa, int ndir, int nsgn, T rho, const T *u, T cs2, T den2) {
    using L = Lattice<Q>;
    T Sxx = 0, Syy = 0, Sxy = 0, Szz = 0, Sxz = 0, Syz = 0;
    // Unknown populations first (non-equilibrium bounce-back), in v order as LUMA does:
    // each v's update reads fa[opp v], which for an unknown v is a known population.
#pragma unroll
    for (int v = 0; v < Q; ++v) {
        const int c0 = L::c(v, 0), c1 = L::c(v, 1), c2 = L::c(v, 2);
        const int o = L::opp(v);
        if (L::c(v, ndir) == nsgn)
            fa[v] = feq_exact<Q, T>(v, rho, u, cs2, den2) + (fa[o] - feq_exact<Q, T>(o, rho, u, cs2, den2));
        const T fneq = fa[v] - feq_exact<Q, T>(v, rho, u, cs2, den2);
        Sxx += T(c0 * c0) * fneq;
        Syy += T(c1 * c1) * fneq;
        Sxy += T(c0 * c1) * fneq;
        if constexpr (L::D == 3) {
            Szz += T(c2 * c2) * fneq;
            Sxz += T(c0 * c2) * fneq;
            Syz += T(c1 * c2) * fneq;
        }
    }
#pragma unroll
    for (int v = 0; v < Q; ++v) {
        const int c0 = L::c(v, 0), c1 = L::c(v, 1), c2 = L::c(v, 2);
        fa[v] = feq_exact<Q, T>(v, rho, u, cs2, den2) +
                (T(L::w(v)) / den2) *
                    (((T(c0 * c0) - cs2) * Sxx) + ((T(c1 * c1) - cs2) * Syy) +
                     ((T(c2 * c2) - cs2) * Szz) + (T(2.0) * T(c0) * T(c1) * Sxy) +
                     (T(2.0) * T(c0) * T(c2) * Sxz) + (T(2.0) * T(c1) * T(c2) * Syz));
    }
}

// _LBM_regularised_opt for a velocity site on a face (edgeCount == 1), :313-510.
// fa holds absolute populations after streaming; on return it holds the regularised set and
// rho/u the site's macroscopic values.
template <int Q, class T>
__device__ __forceinline__ void regularised_face(T *fa, int ncode, const T *vel_in, T rho_in,
                                                 T cs2, T den2, T &rho, T *u) {
    using L = Lattice<Q>;
    const int ndir = (ncode - 1) >> 1;              // 0:x 1:y 2:z
    const int nsgn = ((ncode - 1) & 1) ? -1 : 1;    // inward normal component
    T f_plus = T(0), f_zero = T(0);
#pragma unroll
    for (int v = 0; v < Q; ++v) {
        const int cn = L::c(v, ndir);
        if (cn == -nsgn) f_plus += fa[v];
        else if (cn == 0) f_zero += fa[v];
    }
    T vel[3] = {vel_in[0], vel_in[1], vel_in[2]};
    T nv = vel[ndir];
    if (nsgn == -1) nv *= T(-1.0);
    (void)rho_in;
    const T dens = (T(1.0) / (T(1.0) - nv)) * (T(2.0) * f_plus + f_zero);
    rho = dens;
    u[0] = vel[0];
    u[1] = vel[1];
    u[2] = (L::D == 3) ? vel[2] : T(0);
    regularise_face_set<Q, T>(fa, ndir, nsgn, rho, u, cs2, den2);
}

// _LBM_regularised_opt for a pressure site on a face (:389-402): rho is the prescribed density
// (L_RHOIN), the tangential velocities are extrapolated from the two inward neighbours by the
// caller (vel_t), the wall-normal one follows from the known populations.
template <int Q, class T>
__device__ __forceinline__ void regularised_pressure(T *fa, int ncode, const T *vel_t, T rho_in, T cs2, T den2,
                                                     T &rho, T *u) {
    using L = Lattice<Q>;
    const int ndir = (ncode - 1) >> 1;
    const int nsgn = ((ncode - 1) & 1) ? -1 : 1;
    T f_plus = T(0), f_zero = T(0);
#pragma unroll
    for (int v = 0; v < Q; ++v) {
        const int cn = L::c(v, ndir);
        if (cn == -nsgn) f_plus += fa[v];
        else if (cn == 0) f_zero += fa[v];
    }
    const T dens = rho_in;
    T vel[3] = {vel_t[0], vel_t[1], vel_t[2]};
    vel[ndir] = T(1.0) - ((T(1.0) / dens) * (T(2.0) * f_plus + f_zero));
    if (nsgn == -1) vel[ndir] *= T(-1.0);
    rho = dens;
    u[0] = vel[0];
    u[1] = vel[1];
    u[2] = (L::D == 3) ? vel[2] : T(0);
    regularise_face_set<Q, T>(fa, ndir, nsgn, rho, u, cs2, den2);
}

// _LBM_regularised_opt for a velocity site on an edge or corner (edgeCount > 1), :346-364 and
// :432-508: rho is the first-order extrapolation from the two inward neighbours along the
// combined normal n (GridUtils::extrapolate, GridUtils.h:231-280, done by the caller); every
// population sharing a normal component with n is unknown -- buried links (c.n = 0, |c| > 1)
// take feq, the others non-equilibrium bounce-back -- updated in place in v order, so a later
// v may read an opposite that an earlier v already replaced, as LUMA's loop does.
template <int Q, class T>
__device__ __forceinline__ void regularised_edge(T *fa, const int *n, const T *vel_in, T dens,
                                                 T cs2, T den2, T &rho, T *u) {
    using L = Lattice<Q>;
    rho = dens;
    u[0] = vel_in[0];
    u[1] = vel_in[1];
    u[2] = (L::D == 3) ? vel_in[2] : T(0);
    T Sxx = 0, Syy = 0, Sxy = 0, Szz = 0, Sxz = 0, Syz = 0;
#pragma unroll
    for (int v = 0; v < Q; ++v) {
        const int c0 = L::c(v, 0), c1 = L::c(v, 1), c2 = L::c(v, 2);
        const int o = L::opp(v);
        const bool unknown = c0 == n[0] || c1 == n[1] || (L::D == 3 && c2 == n[2]);
        if (unknown) {
            const int dp = c0 * n[0] + c1 * n[1] + (L::D == 3 ? c2 * n[2] : 0);
            const int mag2 = c0 * c0 + c1 * c1 + (L::D == 3 ? c2 * c2 : 0);
            if (dp == 0 && mag2 > 1)
                fa[v] = feq_exact<Q, T>(v, rho, u, cs2, den2);
            else
                fa[v] = feq_exact<Q, T>(v, rho, u, cs2, den2) + (fa[o] - feq_exact<Q, T>(o, rho, u, cs2, den2));
        }
        const T fneq = fa[v] - feq_exact<Q, T>(v, rho, u, cs2, den2);
        Sxx += T(c0 * c0) * fneq;
        Syy += T(c1 * c1) * fneq;
        Sxy += T(c0 * c1) * fneq;
        if constexpr (L::D == 3) {
            Szz += T(c2 * c2) * fneq;
            Sxz += T(c0 * c2) * fneq;
            Syz += T(c1 * c2) * fneq;
        }
    }
#pragma unroll
    for (int v = 0; v < Q; ++v) {
        const int c0 = L::c(v, 0), c1 = L::c(v, 1), c2 = L::c(v, 2);
        fa[v] = feq_exact<Q, T>(v, rho, u, cs2, den2) +
                (T(L::w(v)) / den2) *
                    (((T(c0 * c0) - cs2) * Sxx) + ((T(c1 * c1) - cs2) * Syy) +
                     ((T(c2 * c2) - cs2) * Szz) + (T(2.0) * T(c0) * T(c1) * Sxy) +
                     (T(2.0) * T(c0) * T(c2) * Sxz) + (T(2.0) * T(c1) * T(c2) * Syz));
    }
}

// ----------------------------------------------------------------------------------------
// shifted (deviation) per-site arithmetic, FP32 compute (TF = float, or DualF in the shifted
// tangent form)

// Every fused operation of the shifted arithmetic below is written out (fmac, fma3, mul_split)
// and the units that instantiate it are built without contraction, so the same site update
// rounds the same way in every kernel that runs it (one- and two-step kernels, boundary and
// IB kernels) whatever the compiler's scheduling.

// c_V . x over the non-zero components of c_V only (c is -1, 0 or 1): the products with zero
// would otherwise cost instructions, since 0 * x is not folded under IEEE semantics
template <int Q, int V, class TF>
__device__ __forceinline__ TF cdot(const TF *x) {
    using L = Lattice<Q>;
    constexpr int c0 = L::c(V, 0), c1 = L::c(V, 1), c2 = L::c(V, 2);
    TF s = TF(0);
    bool any = false;
    if constexpr (c0 != 0) {
        s = c0 > 0 ? x[0] : -x[0];
        any = true;
    }
    if constexpr (c1 != 0) {
        s = any ? (c1 > 0 ? s + x[1] : s - x[1]) : (c1 > 0 ? x[1] : -x[1]);
        any = true;
    }
    if constexpr (c2 != 0) s = any ? (c2 > 0 ? s + x[2] : s - x[2]) : (c2 > 0 ? x[2] : -x[2]);
    return s;
}

template <int Q, class TF, class FT>
__device__ __forceinline__ void macro_shift(const TF *fin, const FT *F, bool fz_compat,
                                            TF &drho, TF *j) {
    using L = Lattice<Q>;
    // sums start from their first term (0 + x is not folded under IEEE semantics)
    TF r = fin[0], jx = TF(0.f), jy = TF(0.f), jz = TF(0.f);
    bool ax = false, ay = false, az = false;
#pragma unroll
    for (int v = 0; v < Q; ++v) {
        if (v > 0) r += fin[v];
        auto acc = [&](TF &s, bool &any, int c) {
            if (c == 0) return;
            s = any ? (c > 0 ? s + fin[v] : s - fin[v]) : (c > 0 ? fin[v] : -fin[v]);
            any = true;
        };
        acc(jx, ax, L::c(v, 0));
        acc(jy, ay, L::c(v, 1));
        acc(jz, az, L::c(v, 2));
    }
    if (F) {
        jx = fmac(0.5f, TF(F[0]), jx);
        jy = fmac(0.5f, TF(F[1]), jy);
        if constexpr (L::D == 3) {
            if (fz_compat) jx = fmac(0.5f, TF(F[2]), jx);
            else jz = fmac(0.5f, TF(F[2]), jz);
        }
    }
    drho = r;
    j[0] = jx;
    j[1] = jy;
    j[2] = jz;
}

// A constant carried as an unevaluated sum hi + lo of two floats.  Multiplying through
// fmaf(hi, x, lo*x) rounds once, so the product is as accurate as with the exact constant.
// This matters for the lattice weights and omega: rounding w_v to a float breaks the
// equilibrium's moment identities (sum w = 1, sum w c c = I/3) by ~1e-8 the same way every
// step, and near tau = 1/2 a rounded omega shifts the viscosity by ~1e-6 relative; both
// errors then accumulate coherently over thousands of steps.
struct Split {
    float hi, lo;
};
__host__ __device__ constexpr Split split_of(double x) {
    return Split{float(x), float(x - double(float(x)))};
}
__device__ __forceinline__ float mul_split(Split s, float x) { return fmaf(s.hi, x, s.lo * x); }
__device__ __forceinline__ DualF mul_split(Split s, const DualF &x) {
    return DualF(fmaf(s.hi, x.v, s.lo * x.v), fmaf(s.hi, x.d, s.lo * x.d));
}

// BGK on deviations from w_v (the storage of the shifted forms, reference density 1), with
// rho = 1 + drho:
//   feq_v - w_v = w_v s_v,  s_v = drho + (1 + drho) P_v,  P_v = 3 c.u + 4.5 (c.u)^2 - 1.5 u.u,
// so rho itself is never rounded, and the relaxation f* = f + omega (feq - f), with w_v and
// omega carried as float pairs (the constants enter exactly; the difference feq - f is small
// near equilibrium, so the result rounds once at its own magnitude);
// Guo term (1 - omega/2) w_v (3 (c - u).F + 9 (c.u)(c.F)).
struct ShiftK {
    Split om;
    float gpre;
};
__device__ __forceinline__ ShiftK shift_k(const StepArgs &a) {
    ShiftK k;
    k.om = Split{a.om_hi, a.om_lo};
    k.gpre = 1.0f - 0.5f * (a.om_hi + a.om_lo);
    return k;
}

template <int Q, class TF = float>
struct CollideShift {
    TF drho, u[3], nusq15, F[3], uF;
    ShiftK k;
    bool forced;
    template <class FT>
    __device__ __forceinline__ CollideShift(TF drho_, const TF *u_, const FT *Fs, const ShiftK &k_) {
        using L = Lattice<Q>;
        drho = drho_;
        u[0] = u_[0];
        u[1] = u_[1];
        u[2] = u_[2];
        k = k_;
        // -1.5 u.u
        nusq15 = -1.5f * fma3(u[0], u[0], fma3(u[1], u[1], u[2] * u[2]));
        forced = Fs != nullptr;
        F[0] = F[1] = F[2] = uF = 0.f;
        if (forced) {
            F[0] = TF(Fs[0]);
            F[1] = TF(Fs[1]);
            if constexpr (L::D == 3) F[2] = TF(Fs[2]);
            uF = fma3(u[0], F[0], fma3(u[1], F[1], u[2] * F[2]));
        }
    }
    // population V from its streamed value, with cu = c_V . u
    template <int V>
    __device__ __forceinline__ TF one(TF fin, TF cu) const {
        // P = (4.5 cu + 3) cu - 1.5 u.u
        const TF P = fma3(fmac(4.5f, cu, TF(3.0f)), cu, nusq15);
        const TF s = fma3(drho, P, drho + P);
        constexpr Split w = split_of(Lattice<Q>::w(V));
        const TF dlt = mul_split(w, s) - fin;
        TF o = fmac(k.om.hi, dlt, fmac(k.om.lo, dlt, fin));
        if (forced) {
            const TF cF = cdot<Q, V>(F);
            o = fmac(k.gpre, mul_split(w, fma3(9.0f * cu, cF, 3.0f * (cF - uF))), o);
        }
        return o;
    }
    // V and its opposite V + 1 (c_{V+1} = -c_V): one c.u for both
    template <int V>
    __device__ __forceinline__ void pair(TF fin_v, TF fin_o, TF &out_v, TF &out_o) const {
        const TF cu = cdot<Q, V>(u);
        out_v = one<V>(fin_v, cu);
        out_o = one<V + 1>(fin_o, -cu);
    }
    // single population (the rest population, and callers that loop over v)
    template <int V>
    __device__ __forceinline__ TF at(TF fin) const {
        return one<V>(fin, cdot<Q, V>(u));
    }
    __device__ __forceinline__ TF operator()(int v, TF fin) const {
        TF out = fin;
        static_for<Q>([&](auto vc) {
            constexpr int V = decltype(vc)::value;
            if (v == V) out = at<V>(fin);
        });
        return out;
    }
};

template <int Q, class TF, class FT>
__device__ __forceinline__ void collide_shift(const TF *fin, TF drho, const TF *u, const FT *F, const ShiftK &k,
                                              TF *fout) {
    const CollideShift<Q, TF> c(drho, u, F, k);
    static_for<Q>([&](auto vc) {
        constexpr int V = decltype(vc)::value;
        fout[V] = c.template at<V>(fin[V]);
    });
}

// equilibrium of (rho, u) in the storage convention of P (absolute or shifted)
template <int Q, class P>
__device__ __forceinline__ typename P::T feq_store(int v, double rho, const double *u,
                                                   const StepArgs &a) {
    double fe = feq_exact<Q, double>(v, rho, u, a.cs2, a.den2);
    if constexpr (P::kShift) fe -= Lattice<Q>::w(v) * a.rho0;
    return static_cast<typename P::T>(fe);
}

// ----------------------------------------------------------------------------------------
// one-site update shared by all step kernels: macro, force, collide.
// make_collider computes the site's macroscopic values (rho_o, u_o in FP64 for the macro
// arrays) and returns the collision of one population at a time.
template <int Q, class P>
__device__ __forceinline__ auto make_collider(const typename P::T *fin, const typename P::M *F, const StepArgs &a,
                                              typename P::M &rho_o, typename P::M *u_o) {
    using T = typename P::T;
    if constexpr (!P::kShift) {
        T rho, u[3];
        macro_exact<Q, T>(fin, F, a.fz_compat != 0, rho, u);
        rho_o = rho;
        u_o[0] = u[0];
        u_o[1] = u[1];
        u_o[2] = u[2];
        return CollideExact<Q, T>{rho, {u[0], u[1], u[2]}, omega_of<T>(a), T(a.cs2), T(a.den2), F};
    } else {
        using M = typename P::M;
        T drho, j[3], u[3];
        macro_shift<Q>(fin, F, a.fz_compat != 0, drho, j);
        // u = j / (1 + drho) = j - j q with q = drho / (1 + drho): the rounding of 1 + drho
        // and of the reciprocal (one ulp) only reach the small correction term q
        const T q = drho * recip(T(1.0f) + drho);
        u[0] = fma3(-j[0], q, j[0]);
        u[1] = fma3(-j[1], q, j[1]);
        u[2] = fma3(-j[2], q, j[2]);
        rho_o = M(1.0f + drho);
        u_o[0] = M(u[0]);
        u_o[1] = M(u[1]);
        u_o[2] = M(u[2]);
        return CollideShift<Q, T>(drho, u, F, shift_k(a));
    }
}

// all populations through a collider, an opposite pair at a time (shared FP64 divisions)
template <int Q, class C, class T>
__device__ __forceinline__ void collide_pairs(const C &col, const T *fin, T *fout) {
    static_for<(Q - 1) / 2>([&](auto pc) {
        constexpr int v = 2 * decltype(pc)::value;
        col.template pair<v>(fin[v], fin[v + 1], fout[v], fout[v + 1]);
    });
    fout[Q - 1] = col(Q - 1, fin[Q - 1]);
}

template <int Q, class P>
__device__ __forceinline__ void site_update(typename P::T *fin, typename P::T *fout,
                                            const typename P::M *F, const StepArgs &a, typename P::M &rho_o,
                                            typename P::M *u_o, bool collide) {
    const auto col = make_collider<Q, P>(fin, F, a, rho_o, u_o);
    if (collide) collide_pairs<Q>(col, fin, fout);
    else
        for (int v = 0; v < Q; ++v) fout[v] = fin[v];
}

template <int D, class M>
__device__ __forceinline__ void store_macro(const StepArgs &a, size_t cell, M rho, const M *u) {
    M *ro = mv<M>(a.rho_out), *uo = mv<M>(a.u_out);
    ro[cell] = rho;
    uo[cell] = u[0];
    uo[a.ncell + cell] = u[1];
    if constexpr (D == 3) uo[2 * a.ncell + cell] = u[2];
}

// ----------------------------------------------------------------------------------------
// periodic box, no labels: the metric kernel
// Minimum waves per SIMD asked of the register allocator: the FP64 local (even) step fits five
// without spilling; every other variant is left to the compiler.
template <class P, bool ODD, bool STORE>
constexpr int periodic_waves() {
    return (std::is_same_v<P, P64> && !ODD && !STORE) ? 5 : 1;
}

template <int Q, class P, bool ODD, bool STORE, bool O32 = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(periodic_waves<P, ODD, STORE>())))
void k_aa_periodic(StepArgs a) {
    using L = Lattice<Q>;
    using S = typename P::S;
    using T = typename P::T;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y, z = blockIdx.z + a.zoff;
    if (x >= a.nx) return;
    S *__restrict__ f = static_cast<S *>(a.f);
    const size_t ps = a.ps;
    const int nx = a.nx, ny = a.ny;
    const size_t cell = O32 ? (size_t)((uint32_t)x + (uint32_t)nx * ((uint32_t)y + (uint32_t)ny * (uint32_t)z))
                            : cell_of(x, y, z, nx, ny);
    Nbr nb;
    if constexpr (ODD) nb.init(x, y, z, nx, ny, a.nz);
    auto nbr = [&](int dx, int dy, int dz) -> size_t {
        if constexpr (O32) return nb.at32(dx, dy, dz, nx, ny);
        else return nb.at(dx, dy, dz, nx, ny);
    };

    T fin[Q];
    static_for<Q>([&](auto vc) {
        constexpr int v = decltype(vc)::value;
        constexpr int cx = L::c(v, 0), cy = L::c(v, 1), cz = L::c(v, 2);
        if constexpr (ODD)
            fin[v] = ld<P, cx == 0>(elem_at<S, O32>(f + (size_t)L::opp(v) * ps, nbr(-cx, -cy, -cz)));
        else
            fin[v] = ld<P>(elem_at<S, O32>(f + (size_t)v * ps, cell));
    });
    typename P::M rho, u[3];
    const auto col = make_collider<Q, P>(fin, nullptr, a, rho, u);
    if constexpr (STORE) store_macro<L::D>(a, cell, rho, u);
    // collide and store an opposite pair at a time: only the incoming set stays live
    auto put = [&](auto vc, T o) {
        constexpr int v = decltype(vc)::value;
        constexpr int cx = L::c(v, 0), cy = L::c(v, 1), cz = L::c(v, 2);
        if constexpr (ODD)
            st<P, cx == 0>(elem_at<S, O32>(f + (size_t)v * ps, nbr(cx, cy, cz)), o);
        else
            st<P>(elem_at<S, O32>(f + (size_t)L::opp(v) * ps, cell), o);
    };
    static_for<(Q - 1) / 2>([&](auto pc) {
        constexpr int v = 2 * decltype(pc)::value;
        T ov, oo;
        col.template pair<v>(fin[v], fin[v + 1], ov, oo);
        put(std::integral_constant<int, v>{}, ov);
        put(std::integral_constant<int, v + 1>{}, oo);
    });
    put(std::integral_constant<int, Q - 1>{}, col(Q - 1, fin[Q - 1]));
}

// Halo exchange of a z-slab (planes 1..n interior, 0 and n+1 halo; n = nz - 2) with itself as
// both neighbours, i.e. what one rank of a z-sharded periodic box does with RCCL, done as a
// local copy.  Only the populations with c_z != 0 cross a face (5 of 19 each way), and each one
// is a contiguous nx*ny block of one population plane, so nothing is packed.
//   phase 0, before the odd step: halo slots the odd gather reads
//     A[opp v][0]   <- A[opp v][n]   (c_z(v) = +1: pulled from the lower neighbour)
//     A[opp v][n+1] <- A[opp v][1]   (c_z(v) = -1: pulled from the upper neighbour)
//   phase 1, after the odd step: halo slots the odd scatter wrote
//     A[v][1] <- A[v][n+1]           (c_z(v) = +1: pushed to the upper neighbour)
//     A[v][n] <- A[v][0]             (c_z(v) = -1: pushed to the lower neighbour)
template <class P>
__global__ __launch_bounds__(256) void k_zslab_exchange(StepArgs a, int phase) {
    using L = Lattice<19>;
    using S = typename P::S;
    const size_t plane = (size_t)a.nx * a.ny;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= plane) return;
    const int n = a.nz - 2;
    const int k = blockIdx.y;           // 0..9: five velocities with c_z = +1, then five with -1
    int cnt = 0, v = 0;
    const int want = (k < 5) ? 1 : -1;
    for (int w = 0; w < 19; ++w)
        if (L::c(w, 2) == want) {
            if (cnt == (k % 5)) v = w;
            ++cnt;
        }
    int slot, src, dst;
    if (phase == 0) {
        slot = L::opp(v);
        src = (want == 1) ? n : 1;
        dst = (want == 1) ? 0 : n + 1;
    } else {
        slot = v;
        src = (want == 1) ? n + 1 : 0;
        dst = (want == 1) ? 1 : n;
    }
    S *f = static_cast<S *>(a.f);
    f[(size_t)slot * a.ps + (size_t)dst * plane + i] = f[(size_t)slot * a.ps + (size_t)src * plane + i];
}

// ----------------------------------------------------------------------------------------
// labelled grids

// Post-collision f*_v at (x,y,z) of a grid stored in layout `parity` (0: post-local,
// 1: post-neighbour with solid redirect), in storage convention.
template <int Q, class P>
__device__ __forceinline__ typename P::T post_of(const void *fv, size_t ps, int nx, int ny, int nz,
                                                 const uint8_t *flags, int parity, int x, int y, int z,
                                                 int v) {
    using L = Lattice<Q>;
    const typename P::S *f = static_cast<const typename P::S *>(fv);
    const size_t cell = cell_of(x, y, z, nx, ny);
    if (parity == 0) return ld<P, false>(f + (size_t)L::opp(v) * ps + cell);
    int dx = x + L::c(v, 0), dy = y + L::c(v, 1), dz = z + L::c(v, 2);
    dx = dx < 0 ? nx - 1 : (dx >= nx ? 0 : dx);
    dy = dy < 0 ? ny - 1 : (dy >= ny ? 0 : dy);
    dz = dz < 0 ? nz - 1 : (dz >= nz ? 0 : dz);
    const size_t dst = cell_of(dx, dy, dz, nx, ny);
    if (flags && (flags[dst] & kTypeMask) == kSolid) return ld<P, false>(f + (size_t)L::opp(v) * ps + cell);
    return ld<P, false>(f + (size_t)v * ps + dst);
}

// GridUtils::getSubGrid (GridUtils.cpp:1761-1799): the refined region whose coarse limits hold
// site (x, y, z), or null
template <int D>
__device__ __forceinline__ const ChildRef *child_at(const StepArgs &a, int x, int y, int z) {
    for (int k = 0; k < a.nkids; ++k) {
        const ChildRef &c = a.kids[k];
        if (x >= c.lo[0] && x <= c.hi[0] && y >= c.lo[1] && y <= c.hi[1] && (D == 2 || (z >= c.lo[2] && z <= c.hi[2])))
            return &c;
    }
    return nullptr;
}

// Incoming (streamed, pre-collision) populations of one site: _LBM_stream_opt (:206-297) with
// LUMA's source precedence: solid (HWBB), ExtrapolateRight, forced-equilibrium velocity,
// explode from the parent (subcycle 0), coalesce from the child, regular pull.
template <int Q, class P, bool ODD>
__device__ __forceinline__ void gather_in(const StepArgs &a, int x, int y, int z, size_t cell, const Nbr &nb,
                                          int ty, typename P::T *fin) {
    using L = Lattice<Q>;
    using S = typename P::S;
    using T = typename P::T;
    const S *__restrict__ f = static_cast<const S *>(a.f);
    const size_t ps = a.ps;
    const int nx = a.nx, ny = a.ny;
    // the region a TL2F site coalesces from
    const ChildRef *kid = (a.nkids && ty == kTL2F) ? child_at<L::D>(a, x, y, z) : nullptr;
    static_for<Q>([&](auto vc) {
        constexpr int v = decltype(vc)::value;
        constexpr int cx = L::c(v, 0), cy = L::c(v, 1), cz = L::c(v, 2);
        const size_t src = nb.at(-cx, -cy, -cz, nx, ny);
        const int sty = a.flags[src] & kTypeMask;
        if (sty == kSolid) {
            // HWBB: f*_{opp v}(x) of the previous step, in this cell's slot v in either layout
            fin[v] = ld<P>(f + (size_t)v * ps + cell);
        } else if (sty == kExtrapRight) {
            fin[v] = T(mv<typename P::M>(a.xr)[(size_t)v * ny * a.nz + nb.ys[1 - cy] + (size_t)ny * nb.zs[1 - cz]]);
        } else if (!a.regularised && sty == kVelocity) {
            // forced-equilibrium velocity source, :254-270: its rho never changes; its u is the
            // source row's profile, or with the ramp the ramped profile of this site's row y
            // (the reference writes u[src] = ux_in[j] * ramp with j the pulling site's row)
            const int js = a.ramp_on ? y : nb.ys[1 - cy];
            const double r = a.ramp_feq;
            const double us[3] = {a.uin[js] * r, a.uin[ny + js] * r, (L::D == 3) ? a.uin[2 * ny + js] * r : 0.0};
            fin[v] = feq_store<Q, P>(v, a.rho_in, us, a);
        } else if (a.pf && sty == kTL2C && a.subcycle == 0) {
            // explode, :646-663: the parent's post-collision value at the coarse cell of the source
            const int sx = nb.xs[1 - cx], sy = nb.ys[1 - cy], sz = nb.zs[1 - cz];
            fin[v] = post_of<Q, P>(a.pf, a.pps, a.pnx, a.pny, a.pnz, a.pflags, a.pparity, (sx >> 1) + a.clx,
                                   (sy >> 1) + a.cly, (L::D == 3) ? (sz >> 1) + a.clz : 0, v);
        } else if (kid && sty == kRefined) {
            // coalesce, :591-636: average of the child's values over this cell's 2^D children
            const S *cfp = static_cast<const S *>(kid->f);
            const int fx = 2 * (x - kid->lo[0]), fy = 2 * (y - kid->lo[1]), fz = (L::D == 3) ? 2 * (z - kid->lo[2]) : 0;
            T acc = T(0);
            for (int ii = 0; ii < 2; ++ii)
                for (int jj = 0; jj < 2; ++jj)
                    for (int kk = 0; kk < (L::D == 3 ? 2 : 1); ++kk)
                        acc += ld<P, false>(cfp + (size_t)L::opp(v) * kid->ps +
                                            cell_of(fx + ii, fy + jj, fz + kk, kid->nx, kid->ny));
            if constexpr (L::D == 3) acc /= T(8.0);
            else acc /= T(4.0);
            fin[v] = acc;
        } else if (sty == kRefined && a.stale && a.stale_idx[src] >= 0) {
            // a refined source of any other site: the populations it was initialised with
            fin[v] = ld<P, false>(static_cast<const S *>(a.stale) + (size_t)v * a.nstale + a.stale_idx[src]);
        } else if constexpr (ODD) {
            fin[v] = ld<P, cx == 0>(f + (size_t)L::opp(v) * ps + src);
        } else {
            fin[v] = ld<P>(f + (size_t)v * ps + cell);
        }
    });
}

// gather_in for a plain site: every source is a regular pull, so no neighbour flag is read.
template <int Q, class P, bool ODD>
__device__ __forceinline__ void gather_plain(const StepArgs &a, size_t cell, const Nbr &nb, typename P::T *fin) {
    using L = Lattice<Q>;
    using S = typename P::S;
    const S *__restrict__ f = static_cast<const S *>(a.f);
    const size_t ps = a.ps;
    static_for<Q>([&](auto vc) {
        constexpr int v = decltype(vc)::value;
        constexpr int cx = L::c(v, 0), cy = L::c(v, 1), cz = L::c(v, 2);
        if constexpr (ODD) fin[v] = ld<P, cx == 0>(f + (size_t)L::opp(v) * ps + nb.at(-cx, -cy, -cz, a.nx, a.ny));
        else fin[v] = ld<P>(f + (size_t)v * ps + cell);
    });
}

// Plain-site bitmap of a labelled grid: bit (cell & 31) of word cell >> 5 is set when the site
// is fluid and every stencil neighbour (with the periodic wrap the step kernels use) is fluid
// or TL2F.  Each pull and push of such a site is then the regular one in every configuration
// (no bounce-back, outlet, velocity, explode or coalesce source, no solid destination), and the
// step kernel takes it without reading the neighbours' flag bytes.  One thread per word.
template <int Q>
__global__ void k_plain_bits(const uint8_t *flags, int nx, int ny, int nz, size_t ncell, uint32_t *plain) {
    using L = Lattice<Q>;
    const size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w * 32 >= ncell) return;
    uint32_t bits = 0;
    for (int b = 0; b < 32; ++b) {
        const size_t cell = w * 32 + b;
        if (cell >= ncell) break;
        if ((flags[cell] & kTypeMask) != kFluid) continue;
        const int x = (int)(cell % nx), y = (int)((cell / nx) % ny), z = (int)(cell / ((size_t)nx * ny));
        Nbr nb;
        nb.init(x, y, z, nx, ny, nz);
        bool ok = true;
        static_for<Q>([&](auto vc) {
            constexpr int v = decltype(vc)::value;
            const int t = flags[nb.at(L::c(v, 0), L::c(v, 1), L::c(v, 2), nx, ny)] & kTypeMask;
            ok = ok && (t == kFluid || t == kTL2F);
        });
        if (ok) bits |= 1u << b;
    }
    plain[w] = bits;
}

// Regularised velocity face for a gathered site: rho, u of the BC and the regularised
// populations in storage convention.
template <int Q, class P>
__device__ __forceinline__ void velocity_site(const StepArgs &a, unsigned fl, int y, typename P::T *fin,
                                              typename P::T &rr, typename P::T *uu) {
    using L = Lattice<Q>;
    using T = typename P::T;
    const int ny = a.ny;
    const int ncode = (fl >> kNormalShift) & 7;
    const double r = a.ramp_reg;
    T vin[3] = {T(a.uin[y] * r), T(a.uin[ny + y] * r), (L::D == 3) ? T(a.uin[2 * ny + y] * r) : T(0)};
    T fa[Q];
#pragma unroll
    for (int v = 0; v < Q; ++v) fa[v] = P::kShift ? fin[v] + T(L::w(v) * a.rho0) : fin[v];
    regularised_face<Q, T>(fa, ncode, vin, T(a.rho_in), T(a.cs2), T(a.den2), rr, uu);
#pragma unroll
    for (int v = 0; v < Q; ++v) fin[v] = P::kShift ? fa[v] - T(L::w(v) * a.rho0) : fa[v];
}

// labelled grid: solid (HWBB), velocity (regularised or forced equilibrium), ExtrapolateRight
// sources, refinement transition layers, body force.  SPEC: issue the regular pulls before the
// site's flags are known (most sites are plain, and their step then waits on one memory round
// trip less); sites that turn out otherwise gather again.
template <int Q, class P, bool ODD, bool STORE, bool SPEC = false>
__device__ __forceinline__ void masked_site(const StepArgs &a, int x, int y, int z) {
    using L = Lattice<Q>;
    using S = typename P::S;
    using T = typename P::T;
    const int nx = a.nx, ny = a.ny;
    const size_t cell = cell_of(x, y, z, nx, ny);
    Nbr nb;
    nb.init(x, y, z, nx, ny, a.nz);
    T fin[Q], fout[Q];
    if constexpr (SPEC) gather_plain<Q, P, ODD>(a, cell, nb, fin);
    const unsigned fl = a.flags[cell];
    const int ty = fl & kTypeMask;
    // sites LUMA's sweep 1 skips (:91-95)
    if (ty == kSolid || ty == kRefined || (!a.regularised && ty == kVelocity)) return;
    // edge and corner velocity sites and pressure sites run in k_edge_sites after this kernel
    if (ty == kPressure || (ty == kVelocity && ((fl >> kNormalShift) & 7) == kNormalEdge)) return;
    S *__restrict__ f = static_cast<S *>(a.f);
    const size_t ps = a.ps;

    const bool plain = a.plain && ((a.plain[cell >> 5] >> (cell & 31)) & 1u);
    if (!plain) gather_in<Q, P, ODD>(a, x, y, z, cell, nb, ty, fin);
    else if constexpr (!SPEC) gather_plain<Q, P, ODD>(a, cell, nb, fin);

    using M = typename P::M;
    const M *F = nullptr;
    M Fv[3];
    if (a.all_forced || (fl & kForcedBit)) {
        site_force<L::D>(a, cell, fl, Fv);
        F = Fv;
    }

    M rho, u[3];
    if (ty == kVelocity) {
        // regularised velocity face: the BC gives rho, u; no macro on this site (:101-107)
        T rr, uu[3];
        velocity_site<Q, P>(a, fl, y, fin, rr, uu);
        if constexpr (!P::kShift) {
            const CollideExact<Q, T> col{rr, {uu[0], uu[1], uu[2]}, omega_of<T>(a), T(a.cs2), T(a.den2), F};
            collide_pairs<Q>(col, fin, fout);
        } else {
            const CollideShift<Q, T> col(rr - 1.0f, uu, F, shift_k(a));
            collide_pairs<Q>(col, fin, fout);
        }
        rho = M(rr);
        u[0] = M(uu[0]);
        u[1] = M(uu[1]);
        u[2] = M(uu[2]);
    } else if (ty == kExtrapRight) {
        // outlet site: streamed (sources right of it are the x = 0 column, through the periodic
        // wrap; sources on the outlet itself give the value two cells left, :246-251) and
        // collided, but _LBM_macro_opt (:803-807) leaves its rho, u as they are: the collision
        // uses the values its macro arrays hold (set at initialisation or by a restart)
        rho = mv<M>(a.rho_out)[cell];
        u[0] = mv<M>(a.u_out)[cell];
        u[1] = mv<M>(a.u_out)[a.ncell + cell];
        u[2] = (L::D == 3) ? mv<M>(a.u_out)[2 * a.ncell + cell] : M(0.0);
        const T uu[3] = {T(u[0]), T(u[1]), T(u[2])};
        if constexpr (!P::kShift) {
            const CollideExact<Q, T> col{T(rho), {uu[0], uu[1], uu[2]}, omega_of<T>(a), T(a.cs2), T(a.den2), F};
            collide_pairs<Q>(col, fin, fout);
        } else {
            const CollideShift<Q, T> col(T(rho - M(1.0)), uu, F, shift_k(a));
            collide_pairs<Q>(col, fin, fout);
        }
    } else {
        const auto col = make_collider<Q, P>(fin, F, a, rho, u);
        // TL2C sites are streamed but not collided (:144)
        if (ty != kTL2C) collide_pairs<Q>(col, fin, fout);
        else
            for (int v = 0; v < Q; ++v) fout[v] = fin[v];
    }
    if constexpr (STORE) {
        if (ty == kFluid || ty == kTL2F || ty == kVelocity) store_macro<L::D>(a, cell, rho, u);
        const ChildRef *kid = (ty == kTL2F && a.nkids) ? child_at<L::D>(a, x, y, z) : nullptr;
        if (kid && kid->rho) {
            // child TL cluster gets this site's rho, u (:850-891)
            const int fx = 2 * (x - kid->lo[0]), fy = 2 * (y - kid->lo[1]), fz = (L::D == 3) ? 2 * (z - kid->lo[2]) : 0;
            for (int ii = 0; ii < 2; ++ii)
                for (int jj = 0; jj < 2; ++jj)
                    for (int kk = 0; kk < (L::D == 3 ? 2 : 1); ++kk) {
                        const size_t cc = cell_of(fx + ii, fy + jj, fz + kk, kid->nx, kid->ny);
                        mv<M>(kid->rho)[cc] = rho;
                        for (int d = 0; d < L::D; ++d) mv<M>(kid->u)[d * kid->ncell + cc] = u[d];
                    }
        }
    }

    if (a.xr_next && x == nx - 3) {
        const size_t row = (size_t)y + (size_t)ny * z, plane = (size_t)ny * a.nz;
        for (int v = 0; v < Q; ++v) mv<M>(a.xr_next)[(size_t)v * plane + row] = M(static_cast<T>(static_cast<S>(fout[v])));
    }
    static_for<Q>([&](auto vc) {
        constexpr int v = decltype(vc)::value;
        constexpr int cx = L::c(v, 0), cy = L::c(v, 1), cz = L::c(v, 2);
        if constexpr (ODD) {
            const size_t dst = nb.at(cx, cy, cz, nx, ny);
            if (!plain && (a.flags[dst] & kTypeMask) == kSolid)
                st<P>(f + (size_t)L::opp(v) * ps + cell, fout[v]);
            else
                st<P, cx == 0>(f + (size_t)v * ps + dst, fout[v]);
        } else {
            st<P>(f + (size_t)L::opp(v) * ps + cell, fout[v]);
        }
    });
}

template <int Q, class P, bool ODD, bool STORE>
__global__ __launch_bounds__(256) void k_aa_masked(StepArgs a) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= a.nx) return;
    masked_site<Q, P, ODD, STORE, true>(a, x, blockIdx.y, blockIdx.z);
}

// The same step for a list of sites: on large labelled grids the plain sites run in
// k_aa_plain and this kernel takes the rest (boundary, transition and bounce-back neighbours).
template <int Q, class P, bool ODD, bool STORE>
__global__ __launch_bounds__(256) void k_aa_masked_list(StepArgs a, const int *cells, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int cell = cells[i];
    const int x = cell % a.nx, y = (cell / a.nx) % a.ny, z = cell / (a.nx * a.ny);
    masked_site<Q, P, ODD, STORE>(a, x, y, z);
}

// The step of the plain sites of a labelled grid (k_plain_bits), every other site left to
// k_aa_masked_list: the unlabelled kernel's pulls and pushes plus the body force where the
// site carries one, with the arithmetic of masked_site.  It keeps the register use (and so the
// occupancy) of the unlabelled kernel, which the boundary branches of masked_site would not.
template <int Q, class P, bool ODD, bool STORE, bool FORCED>
__global__ __launch_bounds__(256) void k_aa_plain(StepArgs a) {
    using L = Lattice<Q>;
    using S = typename P::S;
    using T = typename P::T;
    using M = typename P::M;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y, z = blockIdx.z;
    if (x >= a.nx) return;
    const int nx = a.nx, ny = a.ny;
    const size_t cell = cell_of(x, y, z, nx, ny);
    if (!((a.plain[cell >> 5] >> (cell & 31)) & 1u)) return;
    S *__restrict__ f = static_cast<S *>(a.f);
    const size_t ps = a.ps;
    Nbr nb;
    nb.init(x, y, z, nx, ny, a.nz);
    T fin[Q];
    gather_plain<Q, P, ODD>(a, cell, nb, fin);
    // the outlet side buffer of the next step (see StepArgs::xr_next)
    M *xr_row = (a.xr_next && x == nx - 3) ? mv<M>(a.xr_next) + (size_t)y + (size_t)ny * z : nullptr;
    const size_t xr_plane = (size_t)ny * a.nz;
    auto put = [&](auto vc, T o) {
        constexpr int v = decltype(vc)::value;
        constexpr int cx = L::c(v, 0), cy = L::c(v, 1), cz = L::c(v, 2);
        if constexpr (ODD) st<P, cx == 0>(f + (size_t)v * ps + nb.at(cx, cy, cz, nx, ny), o);
        else st<P>(f + (size_t)L::opp(v) * ps + cell, o);
        if (xr_row) xr_row[(size_t)v * xr_plane] = M(static_cast<T>(static_cast<S>(o)));
    };
    auto step = [&](const M *F) {
        M rho, u[3];
        const auto col = make_collider<Q, P>(fin, F, a, rho, u);
        if constexpr (STORE) store_macro<L::D>(a, cell, rho, u);
        static_for<(Q - 1) / 2>([&](auto pc) {
            constexpr int v = 2 * decltype(pc)::value;
            T ov, oo;
            col.template pair<v>(fin[v], fin[v + 1], ov, oo);
            put(std::integral_constant<int, v>{}, ov);
            put(std::integral_constant<int, v + 1>{}, oo);
        });
        put(std::integral_constant<int, Q - 1>{}, col(Q - 1, fin[Q - 1]));
    };
    // FORCED: grids that may carry a body force (IB bodies, gravity, a set force field).  The
    // force path (with the spread sum of settled rigid bodies) doubles the kernel's registers,
    // so grids without a force take the variant without it, at the unlabelled kernel's
    // registers and occupancy.
    if constexpr (FORCED) {
        const unsigned fl = a.flags[cell];
        if (a.all_forced || (fl & kForcedBit)) {
            M Fv[3];
            site_force<L::D>(a, cell, fl, Fv);
            step(Fv);
            return;
        }
    }
    step(nullptr);
}

// Macroscopic values at a list of sites (every cell when cells is null), without touching the
// populations, in the layout the next step kernel reads.  ib_force = 0: sweep 1's values
// (stream + BC + macro, :72-107) with the force the step started from -- zero after an IB
// reset, the field itself when the grid is not reset or gravity rewrote it; the IB
// interpolation reads them.  ib_force = 1: ibm_updateMacroscopic (ObjectManager_ops_ibm.cpp
// :696-766), _LBM_macro_opt again with the spread force at the flagged sites.  Sites of other
// types keep the rho, u their arrays hold, as LUMA's arrays do.
template <int Q, class P, bool ODD>
__device__ __forceinline__ bool site_macro_value(const StepArgs &a, size_t cell, int ib_force, typename P::M &rho,
                                                 typename P::M *u) {
    using L = Lattice<Q>;
    using T = typename P::T;
    using M = typename P::M;
    const int x = (int)(cell % a.nx), y = (int)((cell / a.nx) % a.ny), z = (int)(cell / ((size_t)a.nx * a.ny));
    Nbr nb;
    nb.init(x, y, z, a.nx, a.ny, a.nz);
    // regular pulls issued before the flags arrive (see masked_site's SPEC)
    T fin[Q];
    gather_plain<Q, P, ODD>(a, cell, nb, fin);
    const unsigned fl = a.flags[cell];
    const int ty = fl & kTypeMask;
    const bool macro_type = ty == kFluid || ty == kTL2F;
    const bool vel_type = !ib_force && a.regularised && ty == kVelocity && ((fl >> kNormalShift) & 7) != kNormalEdge;
    if (!macro_type && !vel_type) return false;
    if (!(a.plain && ((a.plain[cell >> 5] >> (cell & 31)) & 1u))) gather_in<Q, P, ODD>(a, x, y, z, cell, nb, ty, fin);
    if (vel_type) {
        T rr, uu[3];
        velocity_site<Q, P>(a, fl, y, fin, rr, uu);
        rho = M(rr);
        u[0] = M(uu[0]);
        u[1] = M(uu[1]);
        u[2] = M(uu[2]);
    } else {
        // sweep 1's macro adds F/2 of the force the step started from: zero after an IB reset,
        // the gravity field when the reset writes gravity, the set field on a grid never reset
        M Fv[3];
        const M *F = nullptr;
        const bool flagged = (fl & kForcedBit) != 0;
        if (a.all_forced || (flagged && (ib_force || !a.sweep1_zero))) {
            const M *fx = mv<M>(a.force);
            Fv[0] = fx[cell];
            Fv[1] = fx[a.ncell + cell];
            Fv[2] = (L::D == 3) ? fx[2 * a.ncell + cell] : M(0.0);
            F = Fv;
        }
        T fout[Q];
        site_update<Q, P>(fin, fout, F, a, rho, u, false);
    }
    return true;
}
template <int Q, class P, bool ODD>
__device__ __forceinline__ void site_macro_one(const StepArgs &a, size_t cell, int ib_force) {
    typename P::M rho, u[3];
    if (site_macro_value<Q, P, ODD>(a, cell, ib_force, rho, u)) store_macro<Lattice<Q>::D>(a, cell, rho, u);
}
template <int Q, class P, bool ODD>
__global__ __launch_bounds__(256) void k_site_macro(StepArgs a, int nsites, const int *cells, int ib_force) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nsites) return;
    site_macro_one<Q, P, ODD>(a, cells ? (size_t)cells[i] : (size_t)i, ib_force);
}

// Regularised sites that extrapolate from their inward neighbours: velocity sites on edges and
// corners, and pressure sites (faces).  Each entry e holds the site's cell, its packed inward
// normal (n_a + 1 in bits 2a..2a+1) and the cells of the two inward neighbours x + n and x + 2n,
// which are fluid or TL2F sites.
struct EdgeArgs {
    int n;
    const int *cell, *nrm, *n1, *n2;
    double *erho;   // [2n][4] (M elements): sweep-1 rho, u of the two neighbours
};

// Sweep-1 density and velocity of the extrapolation neighbours, read from the layout the step
// kernel is about to consume.  They equal LUMA's rho[id1], u[id1], ... whether the neighbour
// was visited before the site or is streamed and macro-updated on the fly (:1376-1404): both
// give the neighbour's own stream + macro of this step, with the force the step started from.
template <int Q, class P, bool ODD>
__global__ __launch_bounds__(256) void k_edge_rho(StepArgs a, EdgeArgs e) {
    using T = typename P::T;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * e.n) return;
    const size_t cell = (size_t)((i & 1) ? e.n2[i >> 1] : e.n1[i >> 1]);
    const int x = (int)(cell % a.nx), y = (int)((cell / a.nx) % a.ny), z = (int)(cell / ((size_t)a.nx * a.ny));
    const int ty = a.flags[cell] & kTypeMask;
    Nbr nb;
    nb.init(x, y, z, a.nx, a.ny, a.nz);
    T fin[Q], fout[Q];
    gather_in<Q, P, ODD>(a, x, y, z, cell, nb, ty, fin);
    using M = typename P::M;
    M Fv[3];
    const M *F = nullptr;
    if (!a.sweep1_zero && (a.all_forced || (a.flags[cell] & kForcedBit))) {
        const M *fx = mv<M>(a.force);
        Fv[0] = fx[cell];
        Fv[1] = fx[a.ncell + cell];
        Fv[2] = (Lattice<Q>::D == 3) ? fx[2 * a.ncell + cell] : M(0.0);
        F = Fv;
    }
    M rho, u[3];
    site_update<Q, P>(fin, fout, F, a, rho, u, false);
    M *er = mv<M>(e.erho) + 4 * (size_t)i;
    er[0] = rho;
    er[1] = u[0];
    er[2] = u[1];
    er[3] = u[2];
}

// Edge/corner velocity sites: stream, regularise with the extrapolated density
// rho = 2 rho(x+n) - rho(x+2n), Guo force and BGK collision, then the same in-place AA store
// as the step kernel.  A site's AA slots are its own, so running after the step kernel reads
// and writes exactly what running inside it would.
template <int Q, class P, bool ODD>
__global__ __launch_bounds__(256) void k_edge_sites(StepArgs a, EdgeArgs e, int store) {
    using L = Lattice<Q>;
    using S = typename P::S;
    using T = typename P::T;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= e.n) return;
    const size_t cell = (size_t)e.cell[i];
    const int nx = a.nx, ny = a.ny;
    const int x = (int)(cell % nx), y = (int)((cell / nx) % ny), z = (int)(cell / ((size_t)nx * ny));
    const unsigned fl = a.flags[cell];
    const int pk = e.nrm[i];
    const int n[3] = {(pk & 3) - 1, ((pk >> 2) & 3) - 1, ((pk >> 4) & 3) - 1};
    S *__restrict__ f = static_cast<S *>(a.f);
    const size_t ps = a.ps;
    Nbr nb;
    nb.init(x, y, z, nx, ny, a.nz);
    T fin[Q], fout[Q];
    gather_in<Q, P, ODD>(a, x, y, z, cell, nb, kVelocity, fin);

    using M = typename P::M;
    const M *F = nullptr;
    M Fv[3];
    if (a.all_forced || (fl & kForcedBit)) {
        const M *fx = mv<M>(a.force);
        Fv[0] = fx[cell];
        Fv[1] = fx[a.ncell + cell];
        Fv[2] = (L::D == 3) ? fx[2 * a.ncell + cell] : M(0.0);
        F = Fv;
    }
    const double r = a.ramp_reg;
    const T vin[3] = {T(a.uin[y] * r), T(a.uin[ny + y] * r), (L::D == 3) ? T(a.uin[2 * ny + y] * r) : T(0)};
    const M *er1 = mv<M>(e.erho) + 8 * (size_t)i, *er2 = er1 + 4;
    T fa[Q], rr, uu[3];
#pragma unroll
    for (int v = 0; v < Q; ++v) fa[v] = P::kShift ? fin[v] + T(L::w(v) * a.rho0) : fin[v];
    if ((fl & kTypeMask) == kPressure) {
        // tangential velocities: 2 u(x+n) - u(x+2n) (GridUtils::extrapolate, order 1)
        const T vt[3] = {T(2.0) * T(er1[1]) - T(er2[1]), T(2.0) * T(er1[2]) - T(er2[2]),
                         (L::D == 3) ? T(2.0) * T(er1[3]) - T(er2[3]) : T(0)};
        regularised_pressure<Q, T>(fa, (fl >> kNormalShift) & 7, vt, T(a.rho_p), T(a.cs2), T(a.den2), rr, uu);
    } else {
        const T dens = T(2.0) * T(er1[0]) - T(er2[0]);
        regularised_edge<Q, T>(fa, n, vin, dens, T(a.cs2), T(a.den2), rr, uu);
    }
#pragma unroll
    for (int v = 0; v < Q; ++v) fin[v] = P::kShift ? fa[v] - T(L::w(v) * a.rho0) : fa[v];
    if constexpr (!P::kShift) {
        collide_exact<Q, T>(fin, rr, uu, F, omega_of<T>(a), T(a.cs2), T(a.den2), fout);
    } else {
        collide_shift<Q>(fin, rr - 1.0f, uu, F, shift_k(a), fout);
    }
    if (store) {
        const M u[3] = {M(uu[0]), M(uu[1]), M(uu[2])};
        store_macro<L::D>(a, cell, M(rr), u);
    }
    static_for<Q>([&](auto vc) {
        constexpr int v = decltype(vc)::value;
        constexpr int cx = L::c(v, 0), cy = L::c(v, 1), cz = L::c(v, 2);
        if constexpr (ODD) {
            const size_t dst = nb.at(cx, cy, cz, nx, ny);
            if ((a.flags[dst] & kTypeMask) == kSolid)
                st<P>(f + (size_t)L::opp(v) * ps + cell, fout[v]);
            else
                st<P>(f + (size_t)v * ps + dst, fout[v]);
        } else {
            st<P>(f + (size_t)L::opp(v) * ps + cell, fout[v]);
        }
    });
}

// _LBM_resetForces with L_GRAVITY_ON (:1327-1337): the gravity component of force_xyz at every
// site becomes rho * gravity * refinement_ratio from the rho of the previous step; the other
// components are left as they are, as the reference leaves them.
template <class M>
__global__ void k_gravity_reset(StepArgs a, double *force, double gravity, double rr, int dir) {
    const size_t cell = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= a.ncell) return;
    mv<M>(force)[(size_t)dir * a.ncell + cell] = mv<M>(a.rho_out)[cell] * M(gravity) * M(rr);
}

// _LBM_resetForces (:1327-1337) without gravity: force_xyz = 0 at every site, and the force
// bits (which make the step kernel read the field) cleared with it
template <class M>
__global__ void k_force_clear(StepArgs a, double *force, uint8_t *flags, int D) {
    const size_t cell = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= a.ncell) return;
    for (int d = 0; d < D; ++d) mv<M>(force)[(size_t)d * a.ncell + cell] = M(0.0);
    if (flags) flags[cell] &= (uint8_t)~kForcedBit;
}

// Side buffer for ExtrapolateRight sources (:246-251): xr[v][y, z] = f*_v(N-3, y, z), taken
// from the current layout before the step kernel overwrites it.
template <int Q, class P>
__global__ void k_extrap_right(StepArgs a, int parity, double *xr) {
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y;
    if (y >= a.ny) return;
    const size_t plane = (size_t)a.ny * a.nz;
    using M = typename P::M;
    const size_t src = cell_of(a.nx - 3, y, z, a.nx, a.ny);
    // a refined site there holds the populations it was initialised with (see StepArgs::stale)
    const int si = (a.stale && (a.flags[src] & kTypeMask) == kRefined) ? a.stale_idx[src] : -1;
    for (int v = 0; v < Q; ++v)
        mv<M>(xr)[(size_t)v * plane + y + (size_t)a.ny * z] =
            si >= 0 ? M(ld<P, false>(static_cast<const typename P::S *>(a.stale) + (size_t)v * a.nstale + si))
                  : M(post_of<Q, P>(a.f, a.ps, a.nx, a.ny, a.nz, a.flags, parity, a.nx - 3, y, z, v));
}

// ----------------------------------------------------------------------------------------
// layout conversion between LUMA's AoS f[v + Q*id] (id = k + j*K + i*K*M) and the device.

__device__ __forceinline__ size_t luma_id(int x, int y, int z, int ny, int nz) {
    return (size_t)z + (size_t)y * nz + (size_t)x * nz * ny;
}

// f = feq(rho, u) from LUMA-layout rho[id], u[d + D*id]   (GridObj_init_grids.cpp:315-333)
template <int Q, class P>
__global__ void k_init_feq(StepArgs a, const double *rho, const double *u) {
    using L = Lattice<Q>;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y, z = blockIdx.z;
    if (x >= a.nx) return;
    const size_t cell = cell_of(x, y, z, a.nx, a.ny);
    const size_t id = luma_id(x, y, z, a.ny, a.nz);
    const double us[3] = {u[L::D * id], u[L::D * id + 1], (L::D == 3) ? u[L::D * id + 2] : 0.0};
    typename P::S *f = static_cast<typename P::S *>(a.f);
    for (int v = 0; v < Q; ++v)
        st<P, false>(f + (size_t)L::opp(v) * a.ps + cell, feq_store<Q, P>(v, rho[id], us, a));
}

// upload post-collision f (AoS) into the post-local layout: A[opp v][x] = f_v(x)
template <int Q, class P>
__global__ void k_upload_aos(StepArgs a, const double *faos) {
    using L = Lattice<Q>;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y, z = blockIdx.z;
    if (x >= a.nx) return;
    const size_t cell = cell_of(x, y, z, a.nx, a.ny);
    const size_t id = luma_id(x, y, z, a.ny, a.nz);
    typename P::S *f = static_cast<typename P::S *>(a.f);
    for (int v = 0; v < Q; ++v) {
        double fv = faos[v + Q * id];
        if constexpr (P::kShift) fv -= L::w(v) * a.rho0;
        st<P, false>(f + (size_t)L::opp(v) * a.ps + cell, typename P::T(fv));
    }
}

// the post-collision f*_v(x) from the current layout: its value, or with TAN its derivative
// (tangent form)
template <int Q, class P, bool TAN = false>
__device__ __forceinline__ double gather_post(const StepArgs &a, int x, int y, int z, int v, int parity,
                                              bool masked) {
    using L = Lattice<Q>;
    const typename P::S *f = static_cast<const typename P::S *>(a.f);
    const size_t cell = cell_of(x, y, z, a.nx, a.ny);
    typename P::T fe;
    if (parity == 0) {
        fe = ld<P, false>(f + (size_t)L::opp(v) * a.ps + cell);
    } else {
        Nbr nb;
        nb.init(x, y, z, a.nx, a.ny, a.nz);
        const size_t dst = nb.at(L::c(v, 0), L::c(v, 1), L::c(v, 2), a.nx, a.ny);
        if (masked && (a.flags[dst] & kTypeMask) == kSolid)
            fe = ld<P, false>(f + (size_t)L::opp(v) * a.ps + cell);
        else
            fe = ld<P, false>(f + (size_t)v * a.ps + dst);
    }
    if constexpr (TAN) return deriv_of(fe);
    double fv = value_of(fe);
    if constexpr (P::kShift) fv += L::w(v) * a.rho0;
    return fv;
}

template <int Q, class P, bool TAN = false>
__global__ void k_download_aos(StepArgs a, double *faos, int parity, int masked) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y, z = blockIdx.z;
    if (x >= a.nx) return;
    const size_t id = luma_id(x, y, z, a.ny, a.nz);
    for (int v = 0; v < Q; ++v) faos[v + Q * id] = gather_post<Q, P, TAN>(a, x, y, z, v, parity, masked != 0);
}

// Macro arrays (device SoA of M) -> LUMA layout: values, or with TAN derivatives
template <int D, class M = double, bool TAN = false>
__global__ void k_macro_to_luma(const double *rho_s, const double *u_s, size_t ncell, int nx, int ny,
                                int nz, double *rho_l, double *u_l) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y, z = blockIdx.z;
    if (x >= nx) return;
    const size_t cell = cell_of(x, y, z, nx, ny);
    const size_t id = luma_id(x, y, z, ny, nz);
    const M *r = mv<M>(rho_s), *u = mv<M>(u_s);
    rho_l[id] = TAN ? deriv_of(r[cell]) : value_of(r[cell]);
    for (int d = 0; d < D; ++d) u_l[d + D * id] = TAN ? deriv_of(u[d * ncell + cell]) : value_of(u[d * ncell + cell]);
}

template <int D, class M = double>
__global__ void k_macro_from_luma(double *rho_s, double *u_s, size_t ncell, int nx, int ny, int nz,
                                  const double *rho_l, const double *u_l) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y, z = blockIdx.z;
    if (x >= nx) return;
    const size_t cell = cell_of(x, y, z, nx, ny);
    const size_t id = luma_id(x, y, z, ny, nz);
    M *r = mv<M>(rho_s), *u = mv<M>(u_s);
    r[cell] = M(rho_l[id]);
    for (int d = 0; d < D; ++d) u[d * ncell + cell] = M(u_l[d + D * id]);
}

// ----------------------------------------------------------------------------------------
// FP64 moments of the post-collision state: mass, momentum (3), kinetic energy.
// Deterministic: fixed-shape wave/block trees, per-block partials summed in block order.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

constexpr int kMomentCount = 5;

template <int Q, class P>
__global__ __launch_bounds__(256) void k_moments(StepArgs a, int parity, int masked, double *partials) {
    using L = Lattice<Q>;
    __shared__ double sh[4][kMomentCount];
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y, z = blockIdx.z;
    double m[kMomentCount] = {0, 0, 0, 0, 0};
    if (x < a.nx) {
        const size_t cell = cell_of(x, y, z, a.nx, a.ny);
        const int ty = masked ? (a.flags[cell] & kTypeMask) : kFluid;
        if (ty != kSolid && ty != kRefined) {
            double r = 0, j[3] = {0, 0, 0};
            for (int v = 0; v < Q; ++v) {
                const double fv = gather_post<Q, P>(a, x, y, z, v, parity, masked != 0);
                r += fv;
                j[0] += L::c(v, 0) * fv;
                j[1] += L::c(v, 1) * fv;
                j[2] += L::c(v, 2) * fv;
            }
            m[0] = r;
            m[1] = j[0];
            m[2] = j[1];
            m[3] = j[2];
            m[4] = 0.5 * (j[0] * j[0] + j[1] * j[1] + j[2] * j[2]) / r;
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < kMomentCount; ++k) {
        const double s = wave_sum(m[k]);
        if (lane == 0) sh[wave][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < kMomentCount) {
        double s = 0;
        const int nw = (blockDim.x + 63) >> 6;
        for (int w = 0; w < nw; ++w) s += sh[w][threadIdx.x];
        const size_t b = blockIdx.x + (size_t)gridDim.x * (blockIdx.y + (size_t)gridDim.y * blockIdx.z);
        partials[b * kMomentCount + threadIdx.x] = s;
    }
}

static __global__ void k_sum_partials(const double *partials, size_t nblocks, double *out) {
    // one block of 256: strided partial sums, then a fixed tree
    __shared__ double sh[4][kMomentCount];
    double acc[kMomentCount] = {0, 0, 0, 0, 0};
    for (size_t b = threadIdx.x; b < nblocks; b += blockDim.x)
        for (int k = 0; k < kMomentCount; ++k) acc[k] += partials[b * kMomentCount + k];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = 0; k < kMomentCount; ++k) {
        const double s = wave_sum(acc[k]);
        if (lane == 0) sh[wave][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < kMomentCount) {
        double s = 0;
        for (int w = 0; w < 4; ++w) s += sh[w][threadIdx.x];
        out[threadIdx.x] = s;
    }
}

}  // namespace lumagpu
