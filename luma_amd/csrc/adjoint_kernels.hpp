// adjoint_kernels.hpp -- discrete adjoint of one fused stream / IB / collide step.
//
// Forward step t (post-collision f of step t-1  ->  post-collision f' of step t):
//   g_v(x)  = stream of f (periodic pull, halfway bounce-back, ExtrapolateRight source,
//             forced-equilibrium velocity source), _LBM_stream_opt :206-297;
//   IB      : rho_s, j_s at support sites from g; (rho u)_m = sum_s j_s d_ms, rho_m = sum_s
//             rho_s d_ms; F_m = 2((rho u)_m - rho_m U_m); F_s = -sum_m d_ms F_m eps_m ds_m vd_m
//             (ObjectManager_ops_ibm.cpp:498-689);
//   collide : rho = sum g, u = (sum c g + B F/2)/rho, f'_v = g_v + omega (feq_v - g_v) + Guo_v
//             (:765-990; B adds F_z to the x momentum under the :833 quirk).
// Objective: J = (1/T) sum_t dir . sum_m F_m eps_m ds_m vd_m (the body force LUMA writes to
// Body_<id>_LD, ObjectManager_ops_io.cpp:98-121, averaged over the T steps).
//
// The adjoint runs the step backwards: the collision's transposed Jacobian per site
// (k_adj_collide; regularised faces in k_adj_faces, edges and corners in k_adj_edges), then spread and interpolation swapped (the spread's adjoint gathers the
// site force adjoint to the markers, k_adj_ibm_marker; the interpolation's adjoint spreads the
// marker adjoint to the sites through the same per-site gather lists, k_adj_ibm_site), then
// reverse streaming (k_adj_stream: every population adjoint is gathered from the site that
// pulled it).  Adjoint arrays are plain [v][cell] (no AA layout): lam = dJ/df, gbar = dJ/dg.
#pragma once

#include "aa_kernels.hpp"
#include "ibm_kernels.hpp"

namespace lumagpu {

template <class T>
struct AdjArgs {
    T *lam;          // [Q][ncell] adjoint of the post-collision populations
    T *gbar;         // [Q][ncell] adjoint of the streamed populations
    double *fbar;    // [3][ncell] adjoint of the body-force field (flagged sites)
    double *part;    // per-block partial sums: omega [0, nblk), inlet scale [nblk, 2 nblk)
    int nblk;
};

__device__ __forceinline__ int block_linear() {
    return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
}

// deterministic block sum (blocks of 64 to 1024 threads, a multiple of 64): a fixed tree over
// each wave's lanes, then the waves' sums in wave order; every thread gets the sum
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double sh[16];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sh[w];
    __syncthreads();
    return s;
}

// A regularised velocity face site (not an edge or corner) of a grid stepped with
// regularised boundaries; with the edge and corner sites and the pressure faces, every
// regularised boundary site (each streams like a fluid site and is pulled like one).
__device__ __forceinline__ bool regularised_face_site(const StepArgs &a, unsigned fl) {
    return a.regularised && (fl & kTypeMask) == kVelocity && ((fl >> kNormalShift) & 7) != kNormalEdge;
}
__device__ __forceinline__ bool regularised_boundary_site(const StepArgs &a, unsigned fl) {
    const int ty = fl & kTypeMask;
    return a.regularised && (ty == kVelocity || ty == kPressure);
}

// Transposed Jacobian of a regularised velocity face site's update (regularised_face, :313-510,
// then BGK, :765-790) at its streamed populations g (absolute): one forward-mode (Dual) pass per
// population, one for omega and one for the inlet profile scale (u_in -> s u_in at s = 1).
// gbar_v = sum_w lam_w d f'_w / d g_v; omb and inb the two parameter derivatives.  The site
// carries no body force (the adjoint refuses IB supports on velocity sites).
template <int Q>
__device__ void adj_regularised_face(const StepArgs &a, unsigned fl, int y, const double *g, const double *lam,
                                     double *gbar, double &omb, double &inb) {
    using L = Lattice<Q>;
    const int ncode = (fl >> kNormalShift) & 7;
    const double r = a.ramp_reg;
    const double vin0[3] = {a.uin[y] * r, a.uin[a.ny + y] * r, (L::D == 3) ? a.uin[2 * a.ny + y] * r : 0.0};
    for (int s = 0; s < Q + 2; ++s) {
        Dual ga[Q];
        for (int v = 0; v < Q; ++v) ga[v] = Dual(g[v], v == s ? 1.0 : 0.0);
        Dual vin[3];
        for (int d = 0; d < 3; ++d) vin[d] = Dual(vin0[d], s == Q + 1 ? vin0[d] : 0.0);
        Dual rr, uu[3];
        regularised_face<Q, Dual>(ga, ncode, vin, Dual(a.rho_in), Dual(a.cs2), Dual(a.den2), rr, uu);
        const CollideExact<Q, Dual> col{rr, {uu[0], uu[1], uu[2]}, Dual(a.omega, s == Q ? 1.0 : 0.0),
                                        Dual(a.cs2), Dual(a.den2), nullptr};
        double acc = 0.0;
        for (int w = 0; w < Q; ++w) acc += lam[w] * col(w, ga[w]).d;
        if (s < Q) gbar[s] = acc;
        else if (s == Q) omb = acc;
        else inb = acc;
    }
}

// Transposed collision Jacobian at every fluid site: gbar = (d f'/d g)^T lam, fbar =
// (d f'/d F)^T lam at forced sites, and the omega derivative sum_v lam_v d f'_v / d omega
// (part[block]).  Regularised velocity faces run in k_adj_faces: their forward-mode passes
// take several times this path's registers, and one kernel holding both ran at one wave per
// SIMD with spills.
// FP64 D3Q19: at most 256 VGPRs, two waves per SIMD (it needs 286 for one: 1.83 ms against
// 1.21 ms with 28 B of spills on 6.3 M cells)
template <int Q, class P, bool ODD>
__global__ __launch_bounds__(256, (sizeof(typename P::T) == 8 && Q == 19) ? 2 : 1) void k_adj_collide(
    StepArgs a, AdjArgs<typename P::T> b) {
    using L = Lattice<Q>;
    using T = typename P::T;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y, z = blockIdx.z;
    double om_site = 0.0;
    if (x < a.nx) {
        const size_t cell = cell_of(x, y, z, a.nx, a.ny);
        // the flag byte and the plain-site bit in one round trip (plain sites are fluid sites
        // whose pulls read no neighbour flag)
        const bool pl = (a.plain[cell >> 5] >> (cell & 31)) & 1u;   // never null in the adjoint
        const unsigned fl = a.flags[cell];
        if (pl || (fl & kTypeMask) == kFluid) {   // a plain site is a fluid site
            // the site's adjoint first: its loads need no flag, so they are in flight with the pulls
            T lam[Q];
#pragma unroll
            for (int v = 0; v < Q; ++v) lam[v] = b.lam[(size_t)v * a.ncell + cell];
            Nbr nb;
            nb.init(x, y, z, a.nx, a.ny, a.nz);
            T fin[Q];
            if (pl) gather_plain<Q, P, ODD>(a, cell, nb, fin);
            else gather_in<Q, P, ODD>(a, x, y, z, cell, nb, kFluid, fin);
            T g[Q];
#pragma unroll
            for (int v = 0; v < Q; ++v) g[v] = P::kShift ? fin[v] + T(L::w(v) * a.rho0) : fin[v];
            const bool forced = (fl & kForcedBit) != 0;
            T F[3] = {T(0), T(0), T(0)};
            if (forced) {
                F[0] = T(a.force[cell]);
                F[1] = T(a.force[a.ncell + cell]);
                if constexpr (L::D == 3) F[2] = T(a.force[2 * a.ncell + cell]);
            }
            T rho = T(0), j[3] = {T(0), T(0), T(0)};
#pragma unroll
            for (int v = 0; v < Q; ++v) {
                rho += g[v];
                j[0] += T(L::c(v, 0)) * g[v];
                j[1] += T(L::c(v, 1)) * g[v];
                j[2] += T(L::c(v, 2)) * g[v];
            }
            j[0] += T(0.5) * F[0];
            j[1] += T(0.5) * F[1];
            if constexpr (L::D == 3) {
                if (a.fz_compat) j[0] += T(0.5) * F[2];
                else j[2] += T(0.5) * F[2];
            }
            const T u[3] = {j[0] / rho, j[1] / rho, j[2] / rho};
            const T usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
            const T uF = u[0] * F[0] + u[1] * F[1] + u[2] * F[2];
            const T om = T(a.omega), guo = T(1) - T(0.5) * om;
            T S0 = T(0), Ua[3] = {T(0), T(0), T(0)}, Ga[3] = {T(0), T(0), T(0)}, Fb[3] = {T(0), T(0), T(0)};
            T omb = T(0);
#pragma unroll
            for (int v = 0; v < Q; ++v) {
                const T cx = T(L::c(v, 0)), cy = T(L::c(v, 1)), cz = T(L::c(v, 2)), w = T(L::w(v));
                const T cu = cx * u[0] + cy * u[1] + cz * u[2];
                const T cF = cx * F[0] + cy * F[1] + cz * F[2];
                const T e = T(1) + T(3) * cu + T(4.5) * cu * cu - T(1.5) * usq;
                const T lw = lam[v] * w;
                S0 += lw * e;
                Ua[0] += lw * (T(3) * cx + T(9) * cu * cx - T(3) * u[0]);
                Ua[1] += lw * (T(3) * cy + T(9) * cu * cy - T(3) * u[1]);
                Ua[2] += lw * (T(3) * cz + T(9) * cu * cz - T(3) * u[2]);
                T dom = w * rho * e - g[v];
                if (forced) {
                    Ga[0] += lw * (T(-3) * F[0] + T(9) * cx * cF);
                    Ga[1] += lw * (T(-3) * F[1] + T(9) * cy * cF);
                    Ga[2] += lw * (T(-3) * F[2] + T(9) * cz * cF);
                    Fb[0] += lw * (T(3) * (cx - u[0]) + T(9) * cu * cx);
                    Fb[1] += lw * (T(3) * (cy - u[1]) + T(9) * cu * cy);
                    Fb[2] += lw * (T(3) * (cz - u[2]) + T(9) * cu * cz);
                    dom -= T(0.5) * w * (T(3) * (cF - uF) + T(9) * cu * cF);
                }
                omb += lam[v] * dom;
            }
            T ub[3], jb[3];
            for (int d = 0; d < 3; ++d) {
                ub[d] = om * rho * Ua[d] + guo * Ga[d];
                jb[d] = ub[d] / rho;
            }
            if constexpr (L::D == 2) jb[2] = T(0);
            const T rb = om * S0 - (ub[0] * u[0] + ub[1] * u[1] + ub[2] * u[2]) / rho;
#pragma unroll
            for (int v = 0; v < Q; ++v)
                b.gbar[(size_t)v * a.ncell + cell] = (T(1) - om) * lam[v] + rb + T(L::c(v, 0)) * jb[0] +
                                                    T(L::c(v, 1)) * jb[1] + T(L::c(v, 2)) * jb[2];
            if (forced) {
                double Fbar[3] = {double(guo * Fb[0] + T(0.5) * jb[0]), double(guo * Fb[1] + T(0.5) * jb[1]),
                                  double(guo * Fb[2])};
                if constexpr (L::D == 3) Fbar[2] += double(T(0.5) * (a.fz_compat ? jb[0] : jb[2]));
                for (int d = 0; d < L::D; ++d) b.fbar[(size_t)d * a.ncell + cell] = Fbar[d];
            }
            om_site = double(omb);
        }
    }
    const double s = block_sum(om_site);
    if (threadIdx.x == 0) b.part[block_linear()] = s;
}

// Regularised velocity face sites, one thread per listed site (cells[i]): gbar from
// adj_regularised_face at the site's streamed populations; the omega and inlet-scale
// derivatives per site in fpart[i] and fpart[n + i].
template <int Q, class P, bool ODD>
__global__ void k_adj_faces(StepArgs a, AdjArgs<typename P::T> b, const int *cells, int n, double *fpart) {
    using L = Lattice<Q>;
    using T = typename P::T;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t cell = (size_t)cells[i];
    const int nx = a.nx, ny = a.ny;
    const int x = (int)(cell % nx), y = (int)((cell / nx) % ny), z = (int)(cell / ((size_t)nx * ny));
    Nbr nb;
    nb.init(x, y, z, nx, ny, a.nz);
    T fin[Q];
    gather_in<Q, P, ODD>(a, x, y, z, cell, nb, kVelocity, fin);
    double g[Q], lam[Q], gb[Q];
    for (int v = 0; v < Q; ++v) {
        g[v] = double(fin[v]) + (P::kShift ? L::w(v) * a.rho0 : 0.0);
        lam[v] = double(b.lam[(size_t)v * a.ncell + cell]);
    }
    double om = 0.0, in = 0.0;
    adj_regularised_face<Q>(a, a.flags[cell], y, g, lam, gb, om, in);
    for (int v = 0; v < Q; ++v) b.gbar[(size_t)v * a.ncell + cell] = T(gb[v]);
    fpart[i] = om;
    fpart[n + i] = in;
}

// Sites that extrapolate from their two inward neighbours (k_edge_sites), one thread per site:
// edge and corner velocity sites (regularised_edge with the density 2 rho(x+n) - rho(x+2n) from
// the neighbours' sweep-1 values) and pressure faces (regularised_pressure with the tangential
// velocity 2 u(x+n) - u(x+2n)), then BGK.  The site's own gbar comes from forward-mode passes
// over its populations; the extrapolated quantities' adjoints (dbar[i], or tbar[3 i + d] for
// the velocity components) reach the neighbours' streamed populations through
// d rho / d g_v = 1 and d u_d / d g_v = (c_vd - u_d) / rho (k_adj_edge_nbrs); omega and
// inlet-scale derivatives per site in epart[i] and epart[n + i].
template <int Q, class P, bool ODD>
__global__ void k_adj_edges(StepArgs a, EdgeArgs e, AdjArgs<typename P::T> b, double *dbar, double *tbar,
                            double *epart) {
    using L = Lattice<Q>;
    using T = typename P::T;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= e.n) return;
    const size_t cell = (size_t)e.cell[i];
    const int nx = a.nx, ny = a.ny;
    const int x = (int)(cell % nx), y = (int)((cell / nx) % ny), z = (int)(cell / ((size_t)nx * ny));
    const int pk = e.nrm[i];
    const int n[3] = {(pk & 3) - 1, ((pk >> 2) & 3) - 1, ((pk >> 4) & 3) - 1};
    Nbr nb;
    nb.init(x, y, z, nx, ny, a.nz);
    T fin[Q];
    gather_in<Q, P, ODD>(a, x, y, z, cell, nb, kVelocity, fin);
    double g[Q], lam[Q];
    for (int v = 0; v < Q; ++v) {
        g[v] = double(fin[v]) + (P::kShift ? L::w(v) * a.rho0 : 0.0);
        lam[v] = double(b.lam[(size_t)v * a.ncell + cell]);
    }
    const double *er1 = e.erho + 8 * (size_t)i, *er2 = er1 + 4;
    const bool pressure = (a.flags[cell] & kTypeMask) == kPressure;
    const int ncode = (a.flags[cell] >> kNormalShift) & 7;
    const double dens0 = 2.0 * er1[0] - er2[0];
    const double vt0[3] = {2.0 * er1[1] - er2[1], 2.0 * er1[2] - er2[2], (L::D == 3) ? 2.0 * er1[3] - er2[3] : 0.0};
    const double r = a.ramp_reg;
    const double vin0[3] = {a.uin[y] * r, a.uin[ny + y] * r, (L::D == 3) ? a.uin[2 * ny + y] * r : 0.0};
    // seeds: populations 0..Q-1, then omega, then the extrapolated density and the inlet scale
    // (velocity edges) or the three extrapolated velocity components (pressure faces)
    double omb = 0.0, inb = 0.0, db = 0.0, tb[3] = {0.0, 0.0, 0.0};
    const int ns = Q + (pressure ? 4 : 3);
    for (int s = 0; s < ns; ++s) {
        Dual fa[Q];
        for (int v = 0; v < Q; ++v) fa[v] = Dual(g[v], v == s ? 1.0 : 0.0);
        Dual rr, uu[3], fout[Q];
        if (pressure) {
            Dual vt[3];
            for (int d = 0; d < 3; ++d) vt[d] = Dual(vt0[d], s == Q + 1 + d ? 1.0 : 0.0);
            regularised_pressure<Q, Dual>(fa, ncode, vt, Dual(a.rho_p), Dual(a.cs2), Dual(a.den2), rr, uu);
        } else {
            Dual vin[3];
            for (int d = 0; d < 3; ++d) vin[d] = Dual(vin0[d], s == Q + 2 ? vin0[d] : 0.0);
            regularised_edge<Q, Dual>(fa, n, vin, Dual(dens0, s == Q + 1 ? 1.0 : 0.0), Dual(a.cs2), Dual(a.den2), rr,
                                      uu);
        }
        collide_exact<Q, Dual>(fa, rr, uu, static_cast<const Dual *>(nullptr), Dual(a.omega, s == Q ? 1.0 : 0.0),
                               Dual(a.cs2), Dual(a.den2), fout);
        double acc = 0.0;
        for (int w = 0; w < Q; ++w) acc += lam[w] * fout[w].d;
        if (s < Q) b.gbar[(size_t)s * a.ncell + cell] = T(acc);
        else if (s == Q) omb = acc;
        else if (pressure) tb[s - Q - 1] = acc;
        else if (s == Q + 1) db = acc;
        else inb = acc;
    }
    dbar[i] = db;
    for (int d = 0; d < 3; ++d) tbar[3 * (size_t)i + d] = tb[d];
    epart[i] = omb;
    epart[e.n + i] = inb;
}

// The extrapolation's reach into the neighbours: gbar_v(x) += sum over the sites that
// extrapolate from x of coef * (dbar + sum_d tbar_d (c_vd - u_d(x)) / rho(x)) (coef 2 for the
// neighbour x + n, -1 for x + 2n; rho, u the neighbour's sweep-1 values in erho); one thread per
// neighbour cell, its entries in a fixed order (no atomics).
template <int Q, class T>
__global__ void k_adj_edge_nbrs(int nn, const int *ncell, const int *ptr, const int *ent, const double *dbar,
                                const double *tbar, const double *erho, size_t ncell_total, T *gbar) {
    using L = Lattice<Q>;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nn) return;
    double s[Q];
    for (int v = 0; v < Q; ++v) s[v] = 0.0;
    for (int j = ptr[k]; j < ptr[k + 1]; ++j) {
        const int code = ent[j];   // 2 * site + (0: x + n, 1: x + 2n)
        const int i = code >> 1;
        const double coef = (code & 1) ? -1.0 : 2.0;
        const double *er = erho + 4 * (size_t)code;   // this neighbour's rho, u
        for (int v = 0; v < Q; ++v) {
            double t = dbar[i];
            for (int d = 0; d < L::D; ++d) t += tbar[3 * (size_t)i + d] * (L::c(v, d) - er[1 + d]) / er[0];
            s[v] += coef * t;
        }
    }
    const size_t cell = (size_t)ncell[k];
    for (int v = 0; v < Q; ++v) gbar[(size_t)v * ncell_total + cell] += T(s[v]);
}

// Reverse streaming: lam_w(x) gathers the adjoint of every streamed population that read
// f_w(x) -- the pull by x + c_w (a fluid or regularised boundary site), the bounce-back at
// x itself when x + c_w is solid, and the pull through an ExtrapolateRight source two cells to
// the right.  Also the derivative with
// respect to a scale s of the inlet profile (u_in -> s u_in) of the forced-equilibrium
// velocity sources, d feq_v(rho_in, s u_in) / ds at s = 1.
template <int Q, class P>
__global__ __launch_bounds__(256) void k_adj_stream(StepArgs a, AdjArgs<typename P::T> b) {
    using L = Lattice<Q>;
    using T = typename P::T;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y, z = blockIdx.z;
    double in_site = 0.0;
    if (x < a.nx) {
        const int nx = a.nx, ny = a.ny;
        const size_t cell = cell_of(x, y, z, nx, ny);
        // the site's flag, the plain-site bit and the flag two sites downstream in one round trip.
        // A plain site (every neighbour fluid) is pulled by its neighbours alone and pulls from no
        // velocity source: no neighbour flag to read
        const bool pl = (a.plain[cell >> 5] >> (cell & 31)) & 1u;   // never null in the adjoint
        const unsigned fx = a.flags[x + 2 < nx ? cell + 2 : cell];
        const bool xr = (x + 2 < nx) && ((fx & kTypeMask) == kExtrapRight);
        const unsigned fl = a.flags[cell];
        const int ty = fl & kTypeMask;
        if (pl || ty == kFluid || regularised_boundary_site(a, fl)) {   // a plain site is a fluid site
            Nbr nb;
            nb.init(x, y, z, nx, ny, a.nz);
            Nbr ns;
            if (xr) ns.init(x + 2, y, z, nx, ny, a.nz);
            // every load before the first store: a store to lam would otherwise order the
            // following loads (of gbar, and of flags, which may alias anything) behind it
            T acc[Q];
            if (pl) {
                static_for<Q>([&](auto wc) {
                    constexpr int w = decltype(wc)::value;
                    acc[w] = b.gbar[(size_t)w * a.ncell + nb.at(L::c(w, 0), L::c(w, 1), L::c(w, 2), nx, ny)];
                });
            } else {
#pragma unroll
                for (int w = 0; w < Q; ++w) {
                    const size_t yn = nb.at(L::c(w, 0), L::c(w, 1), L::c(w, 2), nx, ny);
                    const unsigned fn = a.flags[yn];
                    const int tn = fn & kTypeMask;
                    acc[w] = T(0);
                    if (tn == kFluid || regularised_boundary_site(a, fn)) acc[w] = b.gbar[(size_t)w * a.ncell + yn];
                    else if (tn == kSolid) acc[w] = b.gbar[(size_t)L::opp(w) * a.ncell + cell];
                }
            }
            if (xr) {
#pragma unroll
                for (int w = 0; w < Q; ++w) {
                    const size_t yp = ns.at(L::c(w, 0), L::c(w, 1), L::c(w, 2), nx, ny);
                    if ((a.flags[yp] & kTypeMask) == kFluid) acc[w] += b.gbar[(size_t)w * a.ncell + yp];
                }
            }
            if (!a.regularised && !pl) {
                for (int v = 0; v < Q; ++v) {
                    const size_t src = nb.at(-L::c(v, 0), -L::c(v, 1), -L::c(v, 2), nx, ny);
                    if ((a.flags[src] & kTypeMask) != kVelocity) continue;
                    // the source's velocity as the step's gather_in takes it (with a velocity ramp,
                    // the ramped profile of the pulling site's row)
                    const int js = a.ramp_on ? y : nb.ys[1 - L::c(v, 1)];
                    const double rf = a.ramp_feq;
                    const double us[3] = {a.uin[js] * rf, a.uin[ny + js] * rf, (L::D == 3) ? a.uin[2 * ny + js] * rf : 0.0};
                    const double cu = L::c(v, 0) * us[0] + L::c(v, 1) * us[1] + L::c(v, 2) * us[2];
                    // d feq_v / du . u_in = w rho_in (3 c.u + 9 (c.u)^2 - 3 u.u)
                    const double usq = us[0] * us[0] + us[1] * us[1] + us[2] * us[2];
                    const double d = L::w(v) * a.rho_in * (3.0 * cu + 9.0 * cu * cu - 3.0 * usq);
                    in_site += double(b.gbar[(size_t)v * a.ncell + cell]) * d;
                }
            }
#pragma unroll
            for (int w = 0; w < Q; ++w) b.lam[(size_t)w * a.ncell + cell] = acc[w];
        } else {
            for (int w = 0; w < Q; ++w) b.lam[(size_t)w * a.ncell + cell] = T(0);
        }
    }
    const double s = block_sum(in_site);
    if (threadIdx.x == 0) b.part[b.nblk + block_linear()] = s;
}

// IB adjoint, marker side: the spread's adjoint gathers the site force adjoint to the
// markers; the marker's force adjoint then feeds the interpolation's adjoint (mbar) and the
// derivatives with respect to the marker's epsilon, ds and position (through its delta
// weights), accumulated over the steps.
struct AdjIbm {
    const double *fbar;      // [3][ncell]
    const double *sddelta;   // [nm][S][3] d delta_ms / d x_m
    double *mbar;            // [nm][4]: adjoint of (rho u)_m, rho_m
    double *xbar, *epsbar, *dsbar;
    double dir[3];
    double invT;
};

// sum over a 64-lane wave, fixed tree, lane 0's result in every lane (so every lane goes on
// with the same value)
__device__ __forceinline__ double wave_sum_bcast(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return __shfl(v, 0, 64);
}

// One 64-lane wave per marker (blocks of 256: four markers): the lanes take the marker's
// supports, and the two sums over them are wave reductions.
static __global__ __launch_bounds__(256) void k_adj_ibm_marker(IbmArgs ib, AdjIbm aj) {
    const int m = blockIdx.x * 4 + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (m >= ib.nm) return;   // the whole wave
    const int D = ib.D;
    const double eps = ib.eps[m], ds = ib.ds[m];
    const double vd = (D == 3) ? ds : 1.0;
    const double wgt = eps * ds * vd;
    double Fs[3] = {0.0, 0.0, 0.0};
    for (int k = lane; k < ib.ns[m]; k += 64) {
        const size_t cell = (size_t)ib.scell[m * ib.S + k];
        const double dl = ib.sdelta[m * ib.S + k];
        for (int d = 0; d < D; ++d) Fs[d] += dl * aj.fbar[(size_t)d * ib.ncell + cell];
    }
    for (int d = 0; d < 3; ++d) Fs[d] = wave_sum_bcast(Fs[d]);
    double Fb[3], mom[3], FdotFs = 0.0, Fdir = 0.0, rb = 0.0;
    for (int d = 0; d < 3; ++d) {
        const double Fm = (d < D) ? ib.force[3 * m + d] : 0.0;
        Fb[d] = (d < D) ? (-wgt * Fs[d] + aj.dir[d] * wgt * aj.invT) : 0.0;
        mom[d] = 2.0 * Fb[d];
        rb += -2.0 * Fb[d] * ((d < D) ? ib.vel[3 * m + d] : 0.0);
        FdotFs += Fm * Fs[d];
        Fdir += Fm * aj.dir[d];
    }
    if (lane == 0) {
        for (int d = 0; d < 3; ++d) aj.mbar[4 * m + d] = mom[d];
        aj.mbar[4 * m + 3] = rb;
        aj.epsbar[m] += -ds * vd * FdotFs + Fdir * ds * vd * aj.invT;
        aj.dsbar[m] += ((D == 3) ? 2.0 * eps * ds : eps) * (-FdotFs + Fdir * aj.invT);
    }
    double xb[3] = {0.0, 0.0, 0.0};
    for (int k = lane; k < ib.ns[m]; k += 64) {
        const size_t cell = (size_t)ib.scell[m * ib.S + k];
        const double r = ib.rho[cell];
        double term = rb * r;
        for (int d = 0; d < D; ++d) {
            term += mom[d] * r * ib.u[(size_t)d * ib.ncell + cell];
            term -= wgt * ib.force[3 * m + d] * aj.fbar[(size_t)d * ib.ncell + cell];
        }
        for (int d = 0; d < D; ++d) xb[d] += aj.sddelta[((size_t)m * ib.S + k) * 3 + d] * term;
    }
    for (int d = 0; d < 3; ++d) xb[d] = wave_sum_bcast(xb[d]);
    if (lane == 0)
        for (int d = 0; d < D; ++d) aj.xbar[3 * m + d] += xb[d];
}

// IB adjoint, site side: the interpolation's adjoint spread through the per-site gather lists
// (no atomics): gbar_v(s) += rho_bar_s + c_v . j_bar_s at every fluid support site.
template <int Q, class T>
__global__ void k_adj_ibm_site(IbmArgs ib, const double *mbar, const uint8_t *flags, T *gbar) {
    using L = Lattice<Q>;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ib.nsites) return;
    const size_t cell = (size_t)ib.sites[s];
    if ((flags[cell] & kTypeMask) != kFluid) return;
    double jb[3] = {0.0, 0.0, 0.0}, rb = 0.0;
    for (int e = ib.sptr[s]; e < ib.sptr[s + 1]; ++e) {
        const int code = ib.sent[e];
        const int m = code / ib.S;
        const double dl = ib.sdelta[code];
        for (int d = 0; d < 3; ++d) jb[d] += dl * mbar[4 * m + d];
        rb += dl * mbar[4 * m + 3];
    }
    for (int v = 0; v < Q; ++v)
        gbar[(size_t)v * ib.ncell + cell] +=
            T(rb + L::c(v, 0) * jb[0] + L::c(v, 1) * jb[1] + L::c(v, 2) * jb[2]);
}

// (1/T) dir . sum_m F_m eps_m ds_m vd_m added to acc[0]; one 64-lane wave, fixed order
static __global__ void k_adj_objective(IbmArgs ib, double dx, double dy, double dz, double invT, double *acc) {
    const int lane = threadIdx.x;
    const double dir[3] = {dx, dy, dz};
    double s = 0.0;
    for (int m = lane; m < ib.nm; m += 64) {
        const double w = ib.eps[m] * ((ib.D == 3) ? ib.ds[m] : 1.0) * ib.ds[m];
        double fd = 0.0;
        for (int d = 0; d < ib.D; ++d) fd += ib.force[3 * m + d] * dir[d];
        s += fd * w;
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) acc[0] += s * invT;
}

// The initial populations' adjoint in LUMA's layout f[v + Q*id], id = k + j*K + i*K*M, from
// the [v][cell] array: one thread per cell, reads coalesced, each thread's Q values contiguous
template <int Q, class T>
__global__ void k_adj_to_luma(const T *lam, size_t nc, int N, int M, int K, double *out) {
    const size_t cell = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= nc) return;
    const size_t i = cell % N, j = (cell / N) % M, k = cell / ((size_t)N * M);
    const size_t id = k + j * K + i * (size_t)K * M;
    for (int v = 0; v < Q; ++v) out[v + Q * id] = double(lam[(size_t)v * nc + cell]);
}

// acc[k] += sum of part[k*n .. k*n + n), k = 0..nk-1; one block of 1024, each thread with
// four independent running sums (loads in flight), then a fixed tree: the order is fixed
static __global__ __launch_bounds__(1024) void k_adj_sum_partials(const double *part, int n, int nk, double w0,
                                                                   double *acc) {
    __shared__ double sh[1024];
    for (int k = 0; k < nk; ++k) {
        const double *p = part + (size_t)k * n;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        int i = threadIdx.x;
        for (; i + 3 * 1024 < n; i += 4 * 1024) {
            s0 += p[i];
            s1 += p[i + 1024];
            s2 += p[i + 2 * 1024];
            s3 += p[i + 3 * 1024];
        }
        for (; i < n; i += 1024) s0 += p[i];
        sh[threadIdx.x] = (s0 + s1) + (s2 + s3);
        __syncthreads();
        for (int st = 512; st > 0; st >>= 1) {
            if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
            __syncthreads();
        }
        // the first partial (omega) scaled by this step's d omega_t / d omega (1 without a ramp)
        if (threadIdx.x == 0) acc[k] += k == 0 ? w0 * sh[0] : sh[0];
        __syncthreads();
    }
}

}  // namespace lumagpu
