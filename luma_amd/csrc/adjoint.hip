// adjoint.hip -- discrete adjoint of the time step with checkpointing (the differentiation
// form): gradients of the time-averaged IB body force with respect to the relaxation
// frequency, the inlet velocity scale, the marker positions and the initial populations.
//
// Forward: T steps of the ordinary step (lumagpu_step semantics), the population array copied
// to a checkpoint every K steps and the objective accumulated on the device.  Backward: for
// each segment, newest first, the segment's K states are recomputed from its checkpoint and
// kept in HBM; then every step of the segment is reversed from its stored input state (the
// outlet side buffer and the IB forces are recomputed from that state, then the adjoint
// kernels of adjoint_kernels.hpp run).  The marker-position gradient adds, on the host, the
// dependence of epsilon (the RKPM system A eps = 1, ObjectManager_ops_ibm.cpp:773-920) and of
// ds (nearest-marker spacing, :927-981) on the positions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "adjoint_kernels.hpp"
#include "ibm_system.hpp"
#include "launch.hpp"

namespace lumagpu {
namespace {

struct DevBuf {
    void *p = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t bytes) {
        if (hipMalloc(&p, bytes) != hipSuccess) {
            p = nullptr;
            return fail(LUMAGPU_ENOMEM, "adjoint buffer allocation failed (" + std::to_string(bytes) + " bytes)");
        }
        return 0;
    }
};

template <class P>
hipError_t launch_adj_collide(const lumagpu_grid *g, const StepArgs &a, const AdjArgs<typename P::T> &b) {
    const dim3 bl = cell_block(a.nx), gr((a.nx + bl.x - 1) / bl.x, a.ny, a.nz);
    const bool odd = g->parity == 0;
    if (g->Q == 9) {
        if (odd) k_adj_collide<9, P, true><<<gr, bl, 0, g->stream>>>(a, b);
        else k_adj_collide<9, P, false><<<gr, bl, 0, g->stream>>>(a, b);
    } else {
        if (odd) k_adj_collide<19, P, true><<<gr, bl, 0, g->stream>>>(a, b);
        else k_adj_collide<19, P, false><<<gr, bl, 0, g->stream>>>(a, b);
    }
    return hipGetLastError();
}

template <class P>
hipError_t launch_adj_stream(const lumagpu_grid *g, const StepArgs &a, const AdjArgs<typename P::T> &b) {
    const dim3 bl = cell_block(a.nx), gr((a.nx + bl.x - 1) / bl.x, a.ny, a.nz);
    if (g->Q == 9) k_adj_stream<9, P><<<gr, bl, 0, g->stream>>>(a, b);
    else k_adj_stream<19, P><<<gr, bl, 0, g->stream>>>(a, b);
    return hipGetLastError();
}

int check_adjoint_grid(const lumagpu_grid *g) {
    if (g->d.precision != LUMAGPU_FP64 && g->d.precision != LUMAGPU_FP32S)
        return fail(LUMAGPU_EUNSUPPORTED, "the adjoint runs in the FP64 and FP32-shifted forms");
    if (g->parent || !g->kids.empty()) return fail(LUMAGPU_EUNSUPPORTED, "the adjoint runs on a single grid level");
    if (g->zslab) return fail(LUMAGPU_EUNSUPPORTED, "the adjoint does not cover z-slabs");
    if (g->gravity_on) return fail(LUMAGPU_EUNSUPPORTED, "the adjoint does not cover the gravity reset");
    for (size_t c = 0; c < g->hflags.size(); ++c) {
        const int ty = g->hflags[c] & kTypeMask;
        if (ty == kFluid || ty == kSolid || ty == kExtrapRight) continue;
        // forced-equilibrium velocity sources, regularised velocity and pressure sites
        if (ty == kVelocity || ty == kPressure) continue;
        return fail(LUMAGPU_EUNSUPPORTED,
                    "the adjoint covers fluid, solid, velocity, pressure and ExtrapolateRight sites only");
    }
    if (g->ibm) {
        for (const auto &b : g->ibm->bodies)
            if (b.flexible) return fail(LUMAGPU_EUNSUPPORTED, "the adjoint covers rigid IB bodies only");
        // the interpolation's adjoint spreads to fluid support sites
        for (const auto &b : g->ibm->bodies)
            for (const auto &m : b.m)
                for (int k = 0; k < m.ns; ++k) {
                    const size_t c = (size_t)m.sijk[k][0] +
                                     (size_t)g->d.N * ((size_t)m.sijk[k][1] + (size_t)g->d.M * m.sijk[k][2]);
                    const int t = c < g->hflags.size() ? (g->hflags[c] & kTypeMask) : kFluid;
                    if (t == kVelocity || t == kPressure)
                        return fail(LUMAGPU_EUNSUPPORTED,
                                    "the adjoint does not cover IB supports on velocity or pressure sites");
                }
    }
    return 0;
}

// The epsilon chain of one system (a body's markers, or with the universal system every marker
// of the grid in body order): xg += dJ/dx through eps (A eps = 1), and dsbar += the part of
// dJ/d ds that enters through A's columns.
int epsilon_chain(const lumagpu_grid *g, const std::vector<const Marker *> &ms, const double *epsbar, double *dsbar,
                  double *xg) {
    const int n = (int)ms.size(), D = g->D;
    const double dh = g->d.dh;
    // A_IJ = ds_J sum_k delta(x_I; s^I_k) delta(x_J; s^I_k) area_I; solve A^T mu = epsbar
    std::vector<double> A((size_t)n * n, 0.0);
    int kl = 0, ku = 0;   // band of A^T's column-major view, which the solvers factor: A's own
    for (int I = 0; I < n; ++I)
        for (int J = 0; J < n; ++J) {
            if (far_apart(*ms[I], *ms[J], D, dh)) continue;
            double acc = 0.0;
            for (int k = 0; k < ms[I]->ns; ++k)
                acc += ms[I]->delta[k] * delta_at(ms[J]->pos, ms[I]->sxyz[k], dh, ms[J]->dilation, D) * ms[I]->area;
            A[(size_t)I * n + J] = acc * ms[J]->ds;
            if (A[(size_t)I * n + J] != 0.0) {
                kl = std::max(kl, I - J);
                ku = std::max(ku, J - I);
            }
        }
    std::vector<double> At((size_t)n * n), mu(epsbar, epsbar + n);
    for (int I = 0; I < n; ++I)
        for (int J = 0; J < n; ++J) At[(size_t)J * n + I] = A[(size_t)I * n + J];
    // markers ordered along a body (a sphere's lattice by height, a filament along its length)
    // couple only near neighbours in index: the banded solver skips the exact zeros outside the
    // band and otherwise does the dense solver's operations in its order (the same solution)
    const int r = (kl + ku < n - 1) ? lu_solve_band(n, kl, ku, At, mu, "singular IB epsilon system")
                                    : lu_solve(n, At, mu);
    if (r) return r;
    for (int I = 0; I < n; ++I) {
        const Marker &mI = *ms[I];
        for (int J = 0; J < n; ++J) {
            const Marker &mJ = *ms[J];
            const double Abar = -mu[I] * mJ.eps;
            if (Abar == 0.0 || far_apart(mI, mJ, D, dh)) continue;
            dsbar[J] += Abar * A[(size_t)I * n + J] / mJ.ds;
            for (int k = 0; k < mI.ns; ++k) {
                double gI[3], gJ[3];
                delta_grad(mI.pos, mI.sxyz[k], dh, mI.dilation, D, gI);
                delta_grad(mJ.pos, mI.sxyz[k], dh, mJ.dilation, D, gJ);
                const double dJk = delta_at(mJ.pos, mI.sxyz[k], dh, mJ.dilation, D);
                for (int a = 0; a < D; ++a) {
                    xg[3 * I + a] += Abar * mJ.ds * gI[a] * dJk * mI.area;
                    xg[3 * J + a] += Abar * mJ.ds * mI.delta[k] * gJ[a] * mI.area;
                }
            }
        }
    }
    return 0;
}

// The ds chain of one body (ibm_computeDs is per body): ds_m = min(10, min_n |x_m - x_n| / dh)
void ds_chain(const lumagpu_grid *g, const Body &b, const double *dsbar, double *xg) {
    const int n = (int)b.m.size(), D = g->D;
    const double dh = g->d.dh;
    for (int m = 0; m < n; ++m) {
        double best = 10.0;
        int arg = -1;
        for (int q = 0; q < n; ++q) {
            if (q == m) continue;
            double acc = 0.0;
            for (int a = 0; a < 3; ++a) acc += std::pow(b.m[m].pos[a] - b.m[q].pos[a], 2);
            const double dist = std::sqrt(acc) / dh;
            if (dist < best) {
                best = dist;
                arg = q;
            }
        }
        if (arg < 0 || best <= 0.0) continue;
        for (int a = 0; a < D; ++a) {
            const double gdir = (b.m[m].pos[a] - b.m[arg].pos[a]) / (best * dh * dh);
            xg[3 * m + a] += dsbar[m] * gdir;
            xg[3 * arg + a] -= dsbar[m] * gdir;
        }
    }
}

// dJ/dx of every marker through epsilon and ds, per body or through the universal system
int position_chains(const lumagpu_grid *g, const double *epsbar, const double *dsbar_dev, double *xg) {
    const IbmSystem *s = g->ibm;
    std::vector<double> dsbar(dsbar_dev, dsbar_dev + s->nm);
    if (s->universal) {
        std::vector<const Marker *> ms;
        for (const auto &b : s->bodies)
            for (const auto &m : b.m) ms.push_back(&m);
        if (int r = epsilon_chain(g, ms, epsbar, dsbar.data(), xg)) return r;
    } else {
        for (const auto &b : s->bodies) {
            std::vector<const Marker *> ms;
            for (const auto &m : b.m) ms.push_back(&m);
            if (int r = epsilon_chain(g, ms, epsbar + b.first, dsbar.data() + b.first, xg + 3 * (size_t)b.first))
                return r;
        }
    }
    for (const auto &b : s->bodies) ds_chain(g, b, dsbar.data() + b.first, xg + 3 * (size_t)b.first);
    return 0;
}

template <class P>
int adj_run(lumagpu_grid *g, int nsteps, int every, const double dir[3], int backward, lumagpu_adj_result *res,
            double *dpos, double *lambda0) {
    g->xr_ready = 0;   // the populations change outside the step kernels
    using T = typename P::T;
    IbmSystem *s = g->ibm;
    const bool ib = s && s->nm > 0;
    const size_t fbytes = (size_t)g->elem * g->Q * g->ps;
    const double invT = 1.0 / nsteps;
    const int nck = (nsteps + every - 1) / every;
    const int t_start = g->t;   // step t of the run runs at grid time t_start + t (velocity ramp)

    std::vector<DevBuf> ck(nck);
    std::vector<int> ckpar(nck);
    for (auto &b : ck)
        if (int r = b.alloc(fbytes)) return r;
    DevBuf fin, accJ;
    if (int r = fin.alloc(fbytes)) return r;
    // J; omega (fluid sites) and inlet scale (sources); omega and inlet scale (edge, corner and
    // pressure sites); omega and inlet scale (regularised faces)
    if (int r = accJ.alloc(7 * sizeof(double))) return r;
    HIP_TRY(hipMemsetAsync(accJ.p, 0, 7 * sizeof(double), g->stream));

    // forward with checkpoints
    for (int t = 0; t < nsteps; ++t) {
        if (t % every == 0) {
            HIP_TRY(hipMemcpyAsync(ck[t / every].p, g->f, fbytes, hipMemcpyDeviceToDevice, g->stream));
            ckpar[t / every] = g->parity;
        }
        if (int r = step_grid(g, 0)) return r;
        if (ib) {
            k_adj_objective<<<1, 64, 0, g->stream>>>(ibm_args(g), dir[0], dir[1], dir[2], invT,
                                                     static_cast<double *>(accJ.p));
            HIP_TRY(hipGetLastError());
        }
    }
    HIP_TRY(hipMemcpyAsync(fin.p, g->f, fbytes, hipMemcpyDeviceToDevice, g->stream));
    const int fin_par = g->parity, fin_t = g->t;
    const double fin_omega = g->d.omega;
    double hacc[7] = {0, 0, 0, 0, 0, 0, 0};
    if (!backward) {
        HIP_TRY(hipMemcpyAsync(hacc, accJ.p, sizeof(hacc), hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
        if (res) *res = lumagpu_adj_result{hacc[0], 0.0, 0.0};
        return 0;
    }

    // backward buffers
    const size_t nc = g->ncell;
    // blocks of the collide and stream launches (cell_block rows)
    const int nblk = (int)((g->d.N + cell_block(g->d.N).x - 1) / cell_block(g->d.N).x) * g->d.M * g->d.K;
    std::vector<DevBuf> tape(every);
    std::vector<int> tpar(every);
    for (auto &b : tape)
        if (int r = b.alloc(fbytes)) return r;
    DevBuf lam, gbar, fbar, part;
    if (int r = lam.alloc((size_t)g->Q * nc * sizeof(T))) return r;
    if (int r = gbar.alloc((size_t)g->Q * nc * sizeof(T))) return r;
    if (int r = fbar.alloc(3 * nc * sizeof(double))) return r;
    // omega (collision), inlet scale (streaming from forced-equilibrium sources)
    if (int r = part.alloc(2 * (size_t)nblk * sizeof(double))) return r;
    HIP_TRY(hipMemsetAsync(lam.p, 0, (size_t)g->Q * nc * sizeof(T), g->stream));
    HIP_TRY(hipMemsetAsync(fbar.p, 0, 3 * nc * sizeof(double), g->stream));
    AdjArgs<T> ab{static_cast<T *>(lam.p), static_cast<T *>(gbar.p), static_cast<double *>(fbar.p),
                  static_cast<double *>(part.p), nblk};
    DevBuf mbar, xbar, epsbar, dsbar, sdd;
    AdjIbm aj{};
    if (ib) {
        const size_t nm = (size_t)s->nm;
        if (int r = mbar.alloc(4 * nm * sizeof(double))) return r;
        if (int r = xbar.alloc(3 * nm * sizeof(double))) return r;
        if (int r = epsbar.alloc(nm * sizeof(double))) return r;
        if (int r = dsbar.alloc(nm * sizeof(double))) return r;
        if (int r = sdd.alloc(nm * s->S * 3 * sizeof(double))) return r;
        HIP_TRY(hipMemsetAsync(xbar.p, 0, 3 * nm * sizeof(double), g->stream));
        HIP_TRY(hipMemsetAsync(epsbar.p, 0, nm * sizeof(double), g->stream));
        HIP_TRY(hipMemsetAsync(dsbar.p, 0, nm * sizeof(double), g->stream));
        std::vector<double> h(nm * s->S * 3, 0.0);
        for (const auto &b : s->bodies)
            for (size_t i = 0; i < b.m.size(); ++i) {
                const Marker &m = b.m[i];
                for (int k = 0; k < m.ns; ++k)
                    delta_grad(m.pos, m.sxyz[k], g->d.dh, m.dilation, g->D,
                               &h[((size_t)(b.first + i) * s->S + k) * 3]);
            }
        HIP_TRY(hipMemcpyAsync(sdd.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
        aj.fbar = static_cast<double *>(fbar.p);
        aj.sddelta = static_cast<double *>(sdd.p);
        aj.mbar = static_cast<double *>(mbar.p);
        aj.xbar = static_cast<double *>(xbar.p);
        aj.epsbar = static_cast<double *>(epsbar.p);
        aj.dsbar = static_cast<double *>(dsbar.p);
        for (int d = 0; d < 3; ++d) aj.dir[d] = dir[d];
        aj.invT = invT;
    }
    double *acc = static_cast<double *>(accJ.p);
    // edge and corner velocity sites: per-site density adjoints and parameter derivatives, and
    // the neighbour cells they extrapolate from with their entries (2 * site + 0 for x + n,
    // + 1 for x + 2n), in ascending order
    EdgeArgs ea{};
    ea.n = g->nedge;
    ea.cell = g->edge;
    ea.nrm = g->edge + g->nedge;
    ea.n1 = g->edge + 2 * g->nedge;
    ea.n2 = g->edge + 3 * g->nedge;
    ea.erho = g->erho;
    DevBuf dbar, tbar, epart, enb_cell, enb_ptr, enb_ent;
    int n_enb = 0;
    if (g->nedge > 0) {
        const int ne = g->nedge;
        std::vector<int> he(4 * (size_t)ne);
        HIP_TRY(hipMemcpyAsync(he.data(), g->edge, he.size() * sizeof(int), hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
        std::vector<std::pair<int, int>> pairs;   // (cell, code)
        for (int i = 0; i < ne; ++i) {
            pairs.emplace_back(he[2 * (size_t)ne + i], 2 * i);
            pairs.emplace_back(he[3 * (size_t)ne + i], 2 * i + 1);
        }
        std::sort(pairs.begin(), pairs.end());
        std::vector<int> cells, ptr{0}, ent;
        for (size_t k = 0; k < pairs.size(); ++k) {
            if (k == 0 || pairs[k].first != pairs[k - 1].first) {
                if (k) ptr.push_back((int)ent.size());
                cells.push_back(pairs[k].first);
            }
            ent.push_back(pairs[k].second);
        }
        ptr.push_back((int)ent.size());
        n_enb = (int)cells.size();
        if (int r = dbar.alloc(ne * sizeof(double))) return r;
        if (int r = tbar.alloc(3 * (size_t)ne * sizeof(double))) return r;
        if (int r = epart.alloc(2 * (size_t)ne * sizeof(double))) return r;
        if (int r = enb_cell.alloc(cells.size() * sizeof(int))) return r;
        if (int r = enb_ptr.alloc(ptr.size() * sizeof(int))) return r;
        if (int r = enb_ent.alloc(ent.size() * sizeof(int))) return r;
        HIP_TRY(hipMemcpy(enb_cell.p, cells.data(), cells.size() * sizeof(int), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(enb_ptr.p, ptr.data(), ptr.size() * sizeof(int), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(enb_ent.p, ent.data(), ent.size() * sizeof(int), hipMemcpyHostToDevice));
    }

    // the adjoint kernels read the plain-site bitmap without a null check: a grid whose flags
    // the IB set-up made (no bitmap) gets one with no plain site
    DevBuf nobits;
    if (!g->plain) {
        if (int r = nobits.alloc((nc + 31) / 32 * sizeof(uint32_t))) return r;
        HIP_TRY(hipMemsetAsync(nobits.p, 0, (nc + 31) / 32 * sizeof(uint32_t), g->stream));
    }

    // regularised velocity faces (not edges or corners): their own kernel, over a list
    DevBuf face_cells, fpart;
    int nface = 0;
    if (g->d.regularised) {
        std::vector<int> fc;
        for (size_t c = 0; c < g->hflags.size(); ++c) {
            const unsigned fl = g->hflags[c];
            if ((fl & kTypeMask) == kVelocity && ((fl >> kNormalShift) & 7) != kNormalEdge) fc.push_back((int)c);
        }
        nface = (int)fc.size();
        if (nface > 0) {
            if (int r = face_cells.alloc(fc.size() * sizeof(int))) return r;
            if (int r = fpart.alloc(2 * fc.size() * sizeof(double))) return r;
            HIP_TRY(hipMemcpy(face_cells.p, fc.data(), fc.size() * sizeof(int), hipMemcpyHostToDevice));
        }
    }

    void *const f_own = g->f;
    int rc = 0;
    for (int seg = nck - 1; seg >= 0 && !rc; --seg) {
        const int t0 = seg * every, t1 = std::min(nsteps, t0 + every);
        HIP_TRY(hipMemcpyAsync(g->f, ck[seg].p, fbytes, hipMemcpyDeviceToDevice, g->stream));
        g->parity = ckpar[seg];
        g->t = t_start + t0;
        g->xr_ready = 0;
        for (int t = t0; t < t1 && !rc; ++t) {
            HIP_TRY(hipMemcpyAsync(tape[t - t0].p, g->f, fbytes, hipMemcpyDeviceToDevice, g->stream));
            tpar[t - t0] = g->parity;
            if (t + 1 < t1) rc = step_grid(g, 0);
        }
        for (int t = t1 - 1; t >= t0 && !rc; --t) {
            // the step's input state stands in for the grid's populations while it is reversed
            g->f = tape[t - t0].p;
            g->t = t_start + t;
            g->parity = tpar[t - t0];
            if (g->has_outlet) {
                hipError_t e = outlet_prepass(g);
                if (e != hipSuccess) rc = fail(LUMAGPU_EHIP, std::string("adjoint outlet: ") + hipGetErrorString(e));
            }
            if (!rc && ib) rc = apply_core(g);
            // a Reynolds ramp: this step's omega, and its omega partial weighted onto the omega
            // after the ramp
            double w_om = 1.0;
            if (g->re_ramp > 0.0) {
                update_reynolds(g);
                w_om = reynolds_omega_weight(g);
            }
            StepArgs a = make_args(g);
            if (!a.plain) a.plain = static_cast<const uint32_t *>(nobits.p);
            if (!rc) {
                // the edge sites' extrapolation densities of this step
                hipError_t e = g->nedge ? launch_edges(g, a, 0) : hipSuccess;
                if (e == hipSuccess) e = launch_adj_collide<P>(g, a, ab);
                if (e == hipSuccess && nface) {
                    const dim3 fb(64), fg((unsigned)((nface + 63) / 64));
                    const int *fcp = static_cast<const int *>(face_cells.p);
                    double *fp = static_cast<double *>(fpart.p);
                    const bool odd = g->parity == 0;
                    if (g->Q == 9) {
                        if (odd) k_adj_faces<9, P, true><<<fg, fb, 0, g->stream>>>(a, ab, fcp, nface, fp);
                        else k_adj_faces<9, P, false><<<fg, fb, 0, g->stream>>>(a, ab, fcp, nface, fp);
                    } else {
                        if (odd) k_adj_faces<19, P, true><<<fg, fb, 0, g->stream>>>(a, ab, fcp, nface, fp);
                        else k_adj_faces<19, P, false><<<fg, fb, 0, g->stream>>>(a, ab, fcp, nface, fp);
                    }
                    e = hipGetLastError();
                    if (e == hipSuccess) {
                        k_adj_sum_partials<<<1, 1024, 0, g->stream>>>(fp, nface, 2, w_om, acc + 5);
                        e = hipGetLastError();
                    }
                }
                if (e == hipSuccess && g->nedge) {
                    const dim3 eb(64), eg((unsigned)((g->nedge + 63) / 64));
                    double *db = static_cast<double *>(dbar.p), *tb = static_cast<double *>(tbar.p);
                    double *ep = static_cast<double *>(epart.p);
                    const int *nc_p = static_cast<const int *>(enb_cell.p), *np_p = static_cast<const int *>(enb_ptr.p);
                    const int *ne_p = static_cast<const int *>(enb_ent.p);
                    const dim3 nbb(128), nbg((unsigned)((n_enb + 127) / 128));
                    if (g->Q == 9) {
                        if (g->parity == 0) k_adj_edges<9, P, true><<<eg, eb, 0, g->stream>>>(a, ea, ab, db, tb, ep);
                        else k_adj_edges<9, P, false><<<eg, eb, 0, g->stream>>>(a, ea, ab, db, tb, ep);
                        e = hipGetLastError();
                        if (e == hipSuccess)
                            k_adj_edge_nbrs<9, T><<<nbg, nbb, 0, g->stream>>>(n_enb, nc_p, np_p, ne_p, db, tb, g->erho, nc,
                                                                             ab.gbar);
                    } else {
                        if (g->parity == 0) k_adj_edges<19, P, true><<<eg, eb, 0, g->stream>>>(a, ea, ab, db, tb, ep);
                        else k_adj_edges<19, P, false><<<eg, eb, 0, g->stream>>>(a, ea, ab, db, tb, ep);
                        e = hipGetLastError();
                        if (e == hipSuccess)
                            k_adj_edge_nbrs<19, T><<<nbg, nbb, 0, g->stream>>>(n_enb, nc_p, np_p, ne_p, db, tb, g->erho,
                                                                              nc, ab.gbar);
                    }
                    if (e == hipSuccess) e = hipGetLastError();
                    if (e == hipSuccess) {
                        k_adj_sum_partials<<<1, 1024, 0, g->stream>>>(ep, g->nedge, 2, w_om, acc + 3);
                        e = hipGetLastError();
                    }
                }
                if (e == hipSuccess && ib) {
                    IbmArgs ia = ibm_args(g);
                    k_adj_ibm_marker<<<(s->nm + 3) / 4, 256, 0, g->stream>>>(ia, aj);
                    e = hipGetLastError();
                    if (e == hipSuccess) {
                        if (g->Q == 9)
                            k_adj_ibm_site<9, T><<<(s->nsites + 127) / 128, 128, 0, g->stream>>>(
                                ia, aj.mbar, g->flags, ab.gbar);
                        else
                            k_adj_ibm_site<19, T><<<(s->nsites + 127) / 128, 128, 0, g->stream>>>(
                                ia, aj.mbar, g->flags, ab.gbar);
                        e = hipGetLastError();
                    }
                }
                if (e == hipSuccess) e = launch_adj_stream<P>(g, a, ab);
                if (e == hipSuccess) {
                    k_adj_sum_partials<<<1, 1024, 0, g->stream>>>(ab.part, nblk, 2, w_om, acc + 1);
                    e = hipGetLastError();
                }
                if (e != hipSuccess) rc = fail(LUMAGPU_EHIP, std::string("adjoint launch: ") + hipGetErrorString(e));
            }
            g->f = f_own;
        }
    }
    // the grid returns to the end of the forward run
    HIP_TRY(hipMemcpyAsync(g->f, fin.p, fbytes, hipMemcpyDeviceToDevice, g->stream));
    g->parity = fin_par;
    g->t = fin_t;
    g->d.omega = fin_omega;
    g->xr_ready = 0;
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(hacc, accJ.p, sizeof(hacc), hipMemcpyDeviceToHost, g->stream));
    if (lambda0) {
        // LUMA's layout on the device, then one copy
        DevBuf lout;
        if (int r = lout.alloc((size_t)g->Q * nc * sizeof(double))) return r;
        const unsigned nb = (unsigned)((nc + 255) / 256);
        if (g->Q == 9)
            k_adj_to_luma<9, T><<<nb, 256, 0, g->stream>>>(static_cast<const T *>(lam.p), nc, g->d.N, g->d.M, g->d.K,
                                                         static_cast<double *>(lout.p));
        else
            k_adj_to_luma<19, T><<<nb, 256, 0, g->stream>>>(static_cast<const T *>(lam.p), nc, g->d.N, g->d.M,
                                                          g->d.K, static_cast<double *>(lout.p));
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(lambda0, lout.p, (size_t)g->Q * nc * sizeof(double), hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
    }
    std::vector<double> hx, he, hd;
    if (ib && dpos) {
        hx.resize(3 * (size_t)s->nm);
        he.resize(s->nm);
        hd.resize(s->nm);
        HIP_TRY(hipMemcpyAsync(hx.data(), xbar.p, hx.size() * sizeof(double), hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipMemcpyAsync(he.data(), epsbar.p, he.size() * sizeof(double), hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipMemcpyAsync(hd.data(), dsbar.p, hd.size() * sizeof(double), hipMemcpyDeviceToHost, g->stream));
    }
    HIP_TRY(hipStreamSynchronize(g->stream));
    if (res) *res = lumagpu_adj_result{hacc[0], hacc[1] + hacc[3] + hacc[5], hacc[2] + hacc[4] + hacc[6]};
    if (ib && dpos) {
        for (size_t i = 0; i < hx.size(); ++i) dpos[i] = hx[i];
        if (int r = position_chains(g, he.data(), hd.data(), dpos)) return r;
    }
    return 0;
}

}  // namespace
}  // namespace lumagpu

using namespace lumagpu;

extern "C" int lumagpu_adj_run(lumagpu_grid *g, int nsteps, int checkpoint_every, const double *dir, int backward,
                               lumagpu_adj_result *res, double *dJ_dpos, double *lambda0) {
    if (!g) return fail(LUMAGPU_EINVAL, "null grid handle");
    if (!g->has_state) return fail(LUMAGPU_ESTATE, "adjoint before the populations were initialised");
    if (nsteps <= 0 || checkpoint_every <= 0) return fail(LUMAGPU_EINVAL, "step count and checkpoint interval must be positive");
    if (!dir) return fail(LUMAGPU_EINVAL, "null objective direction");
    if (int r = check_adjoint_grid(g)) return r;
    if (g->ibm && !g->ibm->initialised && g->ibm->nm > 0)
        return fail(LUMAGPU_ESTATE, "IB bodies added but lumagpu_ibm_initialise not called");
    if (!g->flags) {
        if (int r = ensure_flags(g)) return r;
        if (int r = upload_flags(g)) return r;
    }
    if (g->d.precision == LUMAGPU_FP64)
        return adj_run<P64>(g, nsteps, checkpoint_every, dir, backward, res, dJ_dpos, lambda0);
    return adj_run<P32S>(g, nsteps, checkpoint_every, dir, backward, res, dJ_dpos, lambda0);
}
