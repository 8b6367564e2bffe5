// ibm.hip -- immersed-boundary bodies of a grid: host set-up (support stencils, ds, epsilon)
// in FP64 and the device step (support-site macro, interpolate + force, spread).
//
// Host set-up restates ObjectManager::ibm_findSupport, ibm_initialiseSupport, ibm_computeDs,
// ibm_findEpsilon and ibm_deltaKernel (LUMA/src/ObjectManager_ops_ibm.cpp:255-981) for a
// serial run; the dense epsilon system is solved here by LU with partial pivoting in the order
// of LAPACK's unblocked dgetf2/dgetrs('T'), which the reference calls through
// GridUtils::solveLinearSystem (GridUtils.cpp:356-395).  Built with -ffp-contract=off.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <cmath>
#include <string>
#include <vector>

#include <memory>

#include "ibm_system.hpp"

namespace lumagpu {

// ibm_deltaKernel, :255-272: Roma et al. three-point kernel
double delta_kernel(double radius, double dilation) {
    const double mag_r = std::fabs(radius) / dilation;
    if (mag_r > 1.5) return 0.0;
    if (mag_r > 0.5) return (5.0 - (3.0 * mag_r) - std::sqrt(-3.0 * std::pow(1.0 - mag_r, 2) + 1.0)) / 6.0;
    return (1.0 + std::sqrt(1.0 - 3.0 * std::pow(mag_r, 2))) / 3.0;
}

// d delta_kernel / d radius (zero outside the support and at the kinks' one-sided limits)
double delta_kernel_deriv(double radius, double dilation) {
    const double mag_r = std::fabs(radius) / dilation;
    const double sgn = radius < 0.0 ? -1.0 : 1.0;
    if (mag_r > 1.5) return 0.0;
    double d;
    if (mag_r > 0.5) {
        const double q = 1.0 - mag_r;
        d = (-3.0 - 3.0 * q / std::sqrt(1.0 - 3.0 * q * q)) / 6.0;
    } else {
        d = -mag_r / std::sqrt(1.0 - 3.0 * mag_r * mag_r);
    }
    return sgn * d / dilation;
}

double delta_at(const double *x, const double *s, double dh, double dil, int D) {
    double v = 1.0;
    for (int a = 0; a < D; ++a) v *= delta_kernel((x[a] - s[a]) / dh, dil);
    return v;
}

void delta_grad(const double *x, const double *s, double dh, double dil, int D, double out[3]) {
    double ph[3], dph[3];
    for (int a = 0; a < 3; ++a) {
        const double r = (a < D) ? (x[a] - s[a]) / dh : 0.0;
        ph[a] = (a < D) ? delta_kernel(r, dil) : 1.0;
        dph[a] = (a < D) ? delta_kernel_deriv(r, dil) / dh : 0.0;
    }
    for (int a = 0; a < 3; ++a) {
        out[a] = 0.0;
        if (a >= D) continue;
        double v = dph[a];
        for (int b = 0; b < D; ++b)
            if (b != a) v *= ph[b];
        out[a] = v;
    }
}

IbmArgs ibm_args(lumagpu_grid *g) {
    IbmSystem *s = g->ibm;
    IbmArgs b;
    std::memset(&b, 0, sizeof b);
    b.nm = s->nm;
    b.S = s->S;
    b.D = g->D;
    b.scell = s->d_scell;
    b.sdelta = s->d_sdelta;
    b.ns = s->d_ns;
    b.vel = s->d_vel;
    b.eps = s->d_eps;
    b.ds = s->d_ds;
    b.interp = s->d_interp;
    b.force = s->d_force;
    b.rho = g->rho;
    b.u = g->u;
    b.fxyz = g->force;
    b.flags = g->flags;
    b.ncell = g->ncell;
    b.nsites = s->nsites;
    b.sites = s->d_sites;
    b.sptr = s->d_sptr;
    b.sent = s->d_sent;
    b.accumulate = g->all_forced;
    return b;
}

int build_site_map(lumagpu_grid *g);

void ibm_graph_key(const lumagpu_grid *g, std::vector<uint8_t> &key) {
    const IbmSystem *s = g->ibm;
    auto put = [&](const void *p, size_t n) {
        const uint8_t *b = static_cast<const uint8_t *>(p);
        key.insert(key.end(), b, b + n);
    };
    const int present = s && s->nm > 0;
    put(&present, sizeof present);
    if (!present) return;
    const IbmArgs b = ibm_args(const_cast<lumagpu_grid *>(g));
    put(&b, sizeof b);
    const int state[4] = {s->sites_dirty, (int)s->flagged.size(), s->initialised, s->site_map_ok};
    put(state, sizeof state);
    put(&s->d_site_of, sizeof s->d_site_of);
}

namespace {

// GridUtils::getEnclosingVoxel (serial), GridUtils.cpp:1630-1719
int enclosing(const lumagpu_grid *g, double x, int axis) {
    return (int)std::floor((x - g->d.origin[axis]) / g->d.dh);
}

// GridUtils::isWithinDomain, GridUtils.cpp:1311-1330 (level-0 edges from 0)
bool within_domain(const lumagpu_grid *g, const double p[3]) {
    if (p[0] < 0.0 || p[0] > g->d.xmax || p[1] < 0.0 || p[1] > g->d.ymax) return false;
    if (g->D == 3 && (p[2] < 0.0 || p[2] > g->d.zmax)) return false;
    return true;
}

// ibm_initialiseSupport, :462-491
double support_delta(const lumagpu_grid *g, const Marker &m, const double est[3]) {
    const double dh = g->d.dh;
    const double dx = (est[0] - m.pos[0]) / dh, dy = (est[1] - m.pos[1]) / dh;
    if (g->D == 3) {
        const double dz = (est[2] - m.pos[2]) / dh;
        return delta_kernel(dx, m.dilation) * delta_kernel(dy, m.dilation) * delta_kernel(dz, m.dilation);
    }
    return delta_kernel(dx, m.dilation) * delta_kernel(dy, m.dilation);
}

// ibm_findSupport, :279-453: the enclosing voxel first, then every site of the +-5 cage
// within 1.5 dilations on each axis and inside the domain.
int find_support(const lumagpu_grid *g, Body &b, int ib) {
    const int D = g->D, N = g->d.N, M = g->d.M, K = g->d.K;
    const double dh = g->d.dh;
    for (size_t mi = 0; mi < b.m.size(); ++mi) {
        Marker &m = b.m[mi];
        m.ns = 0;
        const int in = enclosing(g, m.pos[0], 0), jn = enclosing(g, m.pos[1], 1);
        const int kn = (D == 3) ? enclosing(g, m.pos[2], 2) : 0;
        if (in < 0 || jn < 0 || kn < 0 || in >= N || jn >= M || kn >= K)
            return fail(LUMAGPU_EPHYS, "Body " + std::to_string(ib) +
                                           " is no-longer inside the domain! Simulation has likely crashed.");
        const double nearpos[3] = {g->xpos[in], g->ypos[jn], (D == 3) ? g->zpos[kn] : 0.0};
        auto add = [&](int i, int j, int k, const double p[3]) {
            m.sijk[m.ns][0] = i;
            m.sijk[m.ns][1] = j;
            m.sijk[m.ns][2] = k;
            for (int a = 0; a < 3; ++a) m.sxyz[m.ns][a] = p[a];
            m.delta[m.ns] = support_delta(g, m, p);
            m.ns++;
        };
        add(in, jn, kn, nearpos);
        // the cage test per axis, once per offset (each axis' test depends on that offset only)
        double ex[11], ey[11], ez[11];
        bool cx[11], cy[11], cz[11];
        for (int o = 0; o < 11; ++o) {
            ex[o] = nearpos[0] + (o - 5) * dh;
            ey[o] = nearpos[1] + (o - 5) * dh;
            ez[o] = (D == 3) ? nearpos[2] + (o - 5) * dh : 0.0;
            cx[o] = std::fabs(m.pos[0] - ex[o]) / dh < 1.5 * m.dilation;
            cy[o] = std::fabs(m.pos[1] - ey[o]) / dh < 1.5 * m.dilation;
            cz[o] = D != 3 || std::fabs(m.pos[2] - ez[o]) / dh < 1.5 * m.dilation;
        }
        for (int i = in - 5; i <= in + 5; i++)
            for (int j = jn - 5; j <= jn + 5; j++)
                for (int k = (D == 3 ? kn - 5 : 0); k <= (D == 3 ? kn + 5 : 0); k++) {
                    const int oi = i - in + 5, oj = j - jn + 5, ok = (D == 3) ? k - kn + 5 : 5;
                    if (!(cx[oi] && cy[oj] && cz[ok])) continue;
                    double est[3] = {ex[oi], ey[oj], (D == 3) ? ez[ok] : 0.0};
                    if (!within_domain(g, est)) continue;
                    if (i == in && j == jn && (D != 3 || k == kn)) continue;
                    if (m.ns >= kSmax) return fail(LUMAGPU_EINVAL, "IB support stencil overflow");
                    if (i < 0 || j < 0 || k < 0 || i >= N || j >= M || k >= K)
                        return fail(LUMAGPU_EUNSUPPORTED, "IB support stencil leaves the grid of body " +
                                                              std::to_string(ib));
                    add(i, j, k, est);
                }
    }
    return 0;
}

// ibm_computeDs, :927-981; on a tangent-form grid also d ds along the marker tangents
// (through the nearest marker: ds = |x_m - x_n| / dh)
// Markers in the order of their coordinate on the axis of the largest extent, for sweeps that
// visit only the markers within a distance of one marker on that axis.
struct AxisOrder {
    int axis = 0;
    std::vector<int> idx, rank;
};
AxisOrder axis_order(const std::vector<double> &p, size_t stride, size_t n, int D) {
    AxisOrder o;
    double ext = -1.0;
    for (int a = 0; a < D; ++a) {
        double lo = HUGE_VAL, hi = -HUGE_VAL;
        for (size_t i = 0; i < n; ++i) {
            lo = std::min(lo, p[stride * i + a]);
            hi = std::max(hi, p[stride * i + a]);
        }
        if (n && hi - lo > ext) {
            ext = hi - lo;
            o.axis = a;
        }
    }
    o.idx.resize(n);
    o.rank.resize(n);
    for (size_t i = 0; i < n; ++i) o.idx[i] = (int)i;
    std::stable_sort(o.idx.begin(), o.idx.end(),
                     [&](int x, int y) { return p[stride * x + o.axis] < p[stride * y + o.axis]; });
    for (size_t k = 0; k < n; ++k) o.rank[o.idx[k]] = (int)k;
    return o;
}

void compute_ds(const lumagpu_grid *g, Body &b) {
    const double dh = g->d.dh;
    const size_t nm = b.m.size();
    // The result of the all-pairs scan (the smallest distance, at the lowest marker index among
    // equal distances) from a sweep outwards along the sorted axis: the squared offset on that
    // axis is a lower bound of the rounded squared distance and grows along the sweep, so the
    // sweep stops once it exceeds the best squared distance (with a margin far above the
    // rounding that could make two different sums give one distance), and the root and
    // division are taken only for sums within that bound.
    std::vector<double> p(3 * nm);
    for (size_t i = 0; i < nm; ++i)
        for (int a = 0; a < 3; ++a) p[3 * i + a] = b.m[i].pos[a];
    const AxisOrder ord = axis_order(p, 3, nm, g->D);
    const int ax = ord.axis;
    for (size_t mi = 0; mi < nm; ++mi) {
        double ds = 10.0, best = HUGE_VAL;
        size_t arg = mi;
        const double *pm = &p[3 * mi];
        auto visit = [&](size_t n) {
            const double da = pm[ax] - p[3 * n + ax];
            const double bound = best * (1.0 + 1e-12);
            if (std::pow(da, 2) > bound) return false;
            double acc = 0.0;
            for (int a = 0; a < 3; ++a) acc += std::pow(pm[a] - p[3 * n + a], 2);
            if (acc > bound) return true;
            const double dist = std::sqrt(acc) / dh;
            if (dist < ds) {
                ds = dist;
                arg = n;
                best = acc;
            } else if (dist == ds && arg != mi && n < arg) {
                arg = n;
                best = std::min(best, acc);
            }
            return true;
        };
        const int r = ord.rank[mi];
        for (int k = r + 1; k < (int)nm && visit((size_t)ord.idx[k]); ++k) {
        }
        for (int k = r - 1; k >= 0 && visit((size_t)ord.idx[k]); --k) {
        }
        Marker &m = b.m[mi];
        m.ds = ds;
        m.dds = 0.0;
        if (tangent_form(g) && arg != mi && ds > 0.0) {
            const Marker &q = b.m[arg];
            double dd = 0.0;
            for (int a = 0; a < 3; ++a) dd += (m.pos[a] - q.pos[a]) * (m.dpos[a] - q.dpos[a]);
            m.dds = dd / (ds * dh * dh);
        }
    }
}

// derivatives of the support weights along the marker tangents (tangent-form grids)
void support_tangents(const lumagpu_grid *g, Body &b) {
    for (Marker &m : b.m) {
        for (int k = 0; k < m.ns; ++k) {
            double gr[3];
            delta_grad(m.pos, m.sxyz[k], g->d.dh, m.dilation, g->D, gr);
            m.ddelta[k] = gr[0] * m.dpos[0] + gr[1] * m.dpos[1] + gr[2] * m.dpos[2];
        }
    }
}

}  // namespace

// A x = rhs for row-major A: factor the column-major view (A^T) with row pivoting, then solve
// through its transpose; the operation order of dgetf2 followed by dgetrs('T').  Pivots are
// chosen on values; on Duals every operation also carries the derivative.
template <class Sc>
int lu_solve_t(int n, std::vector<Sc> &a, std::vector<Sc> &x) {
    auto M = [&](int i, int j) -> Sc & { return a[(size_t)j * n + i]; };
    std::vector<int> piv(n), rows;
    rows.reserve(n);
    for (int j = 0; j < n; ++j) {
        int p = j;
        double best = std::fabs(value_of(M(j, j)));
        for (int i = j + 1; i < n; ++i)
            if (std::fabs(value_of(M(i, j))) > best) {
                best = std::fabs(value_of(M(i, j)));
                p = i;
            }
        piv[j] = p;
        if (value_of(M(p, j)) == 0.0) return fail(LUMAGPU_EPHYS, "singular IB epsilon system");
        if (p != j)
            for (int k = 0; k < n; ++k) std::swap(M(j, k), M(p, k));
        const Sc r = Sc(1.0) / M(j, j);
        rows.clear();
        for (int i = j + 1; i < n; ++i) {
            M(i, j) *= r;
            if (!is_zero(M(i, j))) rows.push_back(i);
        }
        // rank-1 update; a zero multiplier or pivot-row entry changes nothing but the sign of a
        // zero, so only the nonzero pairs are visited (the IB systems are sparse and banded)
        for (int k = j + 1; k < n; ++k) {
            const Sc t = M(j, k);
            if (!is_zero(t))
                for (int i : rows) M(i, k) -= M(i, j) * t;
        }
    }
    for (int j = 0; j < n; ++j) {
        Sc t = x[j];
        for (int i = 0; i < j; ++i) t -= M(i, j) * x[i];
        x[j] = t / M(j, j);
    }
    for (int j = n - 1; j >= 0; --j) {
        Sc t = x[j];
        for (int i = j + 1; i < n; ++i) t -= M(i, j) * x[i];
        x[j] = t;
    }
    for (int j = n - 1; j >= 0; --j)
        if (piv[j] != j) std::swap(x[j], x[piv[j]]);
    return 0;
}

int lu_solve(int n, std::vector<double> &a, std::vector<double> &x) { return lu_solve_t(n, a, x); }
int lu_solve(int n, std::vector<Dual> &a, std::vector<Dual> &x) { return lu_solve_t(n, a, x); }

// lu_solve_t for a matrix A whose entries more than kl below or ku above the diagonal are zero
// (a beam's effective stiffness: kl = ku = 5).  The loops skip only positions that hold exact
// zeros -- pivot candidates past row j + kl, multipliers past it, update columns past
// j + kl + ku (the reach of U under row pivoting), the same reach in the first substitution --
// and every operation on the other entries is lu_solve_t's, in its order.  A skipped term
// changes no value: a zero never wins the pivot search, the rank-1 update already passes over
// zero pivot-row entries, and subtracting a zero product leaves a nonzero sum as it is.  (A
// zero solution component may come out as -0 where lu_solve_t gives +0 or the reverse; the
// beam adds the solution to displacements that are never -0 and squares it for the residual,
// so no result sees the difference.)  The L multipliers of earlier columns move down with later
// row swaps, so after any swap the second substitution runs in full.
template <class Sc>
int lu_solve_band_t(int n, int kl, int ku, std::vector<Sc> &a, std::vector<Sc> &x,
                    const char *singular = "singular beam system") {
    auto M = [&](int i, int j) -> Sc & { return a[(size_t)j * n + i]; };
    std::vector<int> piv(n), rows;
    rows.reserve(kl);
    bool swapped = false;
    for (int j = 0; j < n; ++j) {
        const int ilast = std::min(n - 1, j + kl);
        int p = j;
        double best = std::fabs(value_of(M(j, j)));
        for (int i = j + 1; i <= ilast; ++i)
            if (std::fabs(value_of(M(i, j))) > best) {
                best = std::fabs(value_of(M(i, j)));
                p = i;
            }
        piv[j] = p;
        if (value_of(M(p, j)) == 0.0) return fail(LUMAGPU_EPHYS, singular);
        if (p != j) {
            swapped = true;
            for (int k = 0; k < n; ++k) std::swap(M(j, k), M(p, k));
        }
        const Sc r = Sc(1.0) / M(j, j);
        rows.clear();
        for (int i = j + 1; i <= ilast; ++i) {
            M(i, j) *= r;
            if (!is_zero(M(i, j))) rows.push_back(i);
        }
        const int klast = std::min(n - 1, j + kl + ku);
        for (int k = j + 1; k <= klast; ++k) {
            const Sc t = M(j, k);
            if (!is_zero(t))
                for (int i : rows) M(i, k) -= M(i, j) * t;
        }
    }
    for (int j = 0; j < n; ++j) {
        Sc t = x[j];
        for (int i = std::max(0, j - (kl + ku)); i < j; ++i) t -= M(i, j) * x[i];
        x[j] = t / M(j, j);
    }
    for (int j = n - 1; j >= 0; --j) {
        Sc t = x[j];
        const int ilast = swapped ? n - 1 : std::min(n - 1, j + kl);
        for (int i = j + 1; i <= ilast; ++i) t -= M(i, j) * x[i];
        x[j] = t;
    }
    for (int j = n - 1; j >= 0; --j)
        if (piv[j] != j) std::swap(x[j], x[piv[j]]);
    return 0;
}

int lu_solve_band(int n, int kl, int ku, std::vector<double> &a, std::vector<double> &x, const char *singular) {
    return lu_solve_band_t(n, kl, ku, a, x, singular);
}
int lu_solve_band(int n, int kl, int ku, std::vector<Dual> &a, std::vector<Dual> &x) {
    return lu_solve_band_t(n, kl, ku, a, x);
}

namespace {

// ibm_findEpsilon, :773-920: A_IJ = ds_J * sum_s delta_I(s) delta(x_J - x_s) A_I, A eps = 1.
// On a tangent-form grid also d eps from A d(eps) = -dA eps, with dA from the marker, delta
// and ds tangents.
int find_epsilon(const lumagpu_grid *g, Body &b) {
    const int n = (int)b.m.size(), D = g->D;
    const double dh = g->d.dh;
    const bool tan = tangent_form(g);
    std::vector<double> A((size_t)n * n, 0.0), rhs(n, 1.0), dA;
    if (tan) dA.assign((size_t)n * n, 0.0);
    // positions and dilations side by side for the all-pairs test (a marker record spans ~1.6 KB)
    std::vector<double> pd(4 * (size_t)n);
    for (int I = 0; I < n; ++I) {
        for (int a = 0; a < 3; ++a) pd[4 * (size_t)I + a] = b.m[I].pos[a];
        pd[4 * (size_t)I + 3] = b.m[I].dilation;
    }
    // the markers J near I come from a sweep along the sorted axis that stops past the largest
    // reach any J can have on it; every entry is computed on its own, so the visiting order
    // leaves A as the all-pairs loop makes it
    const AxisOrder ord = axis_order(pd, 4, (size_t)n, D);
    double dmax = 0.0;
    for (int J = 0; J < n; ++J) dmax = std::max(dmax, pd[4 * (size_t)J + 3]);
    std::vector<int> near;
    int kl = 0, ku = 0;
    for (int I = 0; I < n; ++I) {
        const Marker &mI = b.m[I];
        const double *qI = &pd[4 * (size_t)I];
        const double reach = (1.5 * qI[3] + 1.5 * dmax + 0.01) * dh;
        near.clear();
        for (int k = ord.rank[I]; k < n && std::fabs(pd[4 * (size_t)ord.idx[k] + ord.axis] - qI[ord.axis]) <= reach; ++k)
            near.push_back(ord.idx[k]);
        for (int k = ord.rank[I] - 1; k >= 0 && std::fabs(pd[4 * (size_t)ord.idx[k] + ord.axis] - qI[ord.axis]) <= reach; --k)
            near.push_back(ord.idx[k]);
        for (int J : near) {
            // every support of I lies within 1.5 dilations of I per axis, so a marker J further
            // than 1.5 + 1.5 dilations (+ margin) away on some axis has a zero delta at all of
            // them; skipping those exact zero terms leaves the sum unchanged (far_apart's test)
            const double *qJ = &pd[4 * (size_t)J];
            const double lim = (1.5 * qI[3] + 1.5 * qJ[3] + 0.01) * dh;
            bool far = false;
            for (int a = 0; a < D; ++a) far = far || std::fabs(qJ[a] - qI[a]) > lim;
            if (far) continue;
            const Marker &mJ = b.m[J];
            double acc = 0.0;
            for (int k = 0; k < mI.ns; ++k) {
                double dJ;
                if (D == 3)
                    dJ = delta_kernel((mJ.pos[0] - mI.sxyz[k][0]) / dh, mJ.dilation) *
                         delta_kernel((mJ.pos[1] - mI.sxyz[k][1]) / dh, mJ.dilation) *
                         delta_kernel((mJ.pos[2] - mI.sxyz[k][2]) / dh, mJ.dilation);
                else
                    dJ = delta_kernel((mJ.pos[0] - mI.sxyz[k][0]) / dh, mJ.dilation) *
                         delta_kernel((mJ.pos[1] - mI.sxyz[k][1]) / dh, mJ.dilation);
                acc += mI.delta[k] * dJ * mI.area;
            }
            A[(size_t)I * n + J] = acc * mJ.ds;
            if (A[(size_t)I * n + J] != 0.0) {
                // the band of the transpose, which lu_solve_t factors: row J, column I
                kl = std::max(kl, J - I);
                ku = std::max(ku, I - J);
            }
            if (tan) {
                double dacc = 0.0;
                for (int k = 0; k < mI.ns; ++k) {
                    double gJ[3];
                    delta_grad(mJ.pos, mI.sxyz[k], dh, mJ.dilation, D, gJ);
                    const double dJ = delta_at(mJ.pos, mI.sxyz[k], dh, mJ.dilation, D);
                    const double ddJ = gJ[0] * mJ.dpos[0] + gJ[1] * mJ.dpos[1] + gJ[2] * mJ.dpos[2];
                    dacc += (mI.ddelta[k] * dJ + mI.delta[k] * ddJ) * mI.area;
                }
                dA[(size_t)I * n + J] = acc * mJ.dds + dacc * mJ.ds;
            }
        }
    }
    std::vector<double> A0;
    if (tan) A0 = A;
    // a chain of markers (a filament) gives a matrix banded in marker order: the banded solver,
    // which skips only the exact zeros outside the band, otherwise the dense one
    const bool band = 4 * (kl + ku) < n;
    if (int r = band ? lu_solve_band_t(n, kl, ku, A, rhs, "singular IB epsilon system") : lu_solve(n, A, rhs)) return r;
    for (int i = 0; i < n; ++i) b.m[i].eps = rhs[i];
    if (tan) {
        std::vector<double> drhs(n, 0.0);
        for (int I = 0; I < n; ++I) {
            double t = 0.0;
            for (int J = 0; J < n; ++J) t += dA[(size_t)I * n + J] * rhs[J];
            drhs[I] = -t;
        }
        if (int r = lu_solve(n, A0, drhs)) return r;
        for (int i = 0; i < n; ++i) b.m[i].deps = drhs[i];
    }
    return 0;
}

// ibm_findEpsilon over the grid's bodies: per body, or with L_UNIVERSAL_EPSILON_CALC one system
// over every marker of the level in body order (ibm_universalEpsilonGather, :1067-1107), its
// epsilon (and on a tangent-form grid its derivative) scattered back in the same order
// (ibm_universalEpsilonScatter, :1113-1141)
int find_epsilon_all(const lumagpu_grid *g) {
    IbmSystem *s = g->ibm;
    if (!s->universal) {
        for (auto &b : s->bodies)
            if (int r = find_epsilon(g, b)) return r;
        return 0;
    }
    Body all;
    for (const auto &b : s->bodies) all.m.insert(all.m.end(), b.m.begin(), b.m.end());
    if (all.m.empty()) return 0;
    if (int r = find_epsilon(g, all)) return r;
    size_t k = 0;
    for (auto &b : s->bodies)
        for (auto &m : b.m) {
            m.eps = all.m[k].eps;
            m.deps = all.m[k].deps;
            ++k;
        }
    return 0;
}

// Pack markers and supports for the device; build the per-site gather lists in LUMA's spread
// order; clear sites that are no longer supports.
int upload(lumagpu_grid *g) {
    IbmSystem *s = g->ibm;
    const int S = s->S, N = g->d.N, M = g->d.M;
    int nm = 0;
    for (auto &b : s->bodies) {
        b.first = nm;
        nm += (int)b.m.size();
    }
    s->nm = nm;
    // tangent form: every FP64 marker quantity is a (value, derivative) pair
    const size_t w = tangent_form(g) ? 2 : 1;
    auto put = [w](std::vector<double> &v, size_t i, double x, double dx) {
        v[w * i] = x;
        if (w == 2) v[w * i + 1] = dx;
    };
    std::vector<int> scell((size_t)nm * S, 0), ns(nm);
    std::vector<double> sdelta((size_t)nm * S * w, 0.0), vel((size_t)nm * 3 * w), eps(nm * w), ds(nm * w);
    std::vector<int> &site_of = s->site_of, order;
    if (site_of.size() != g->ncell) site_of.assign(g->ncell, -1);
    std::vector<std::vector<int>> entries;
    // back to all -1 on every exit
    struct Reset {
        std::vector<int> &so;
        const std::vector<int> &cells;
        ~Reset() {
            for (int c : cells) so[c] = -1;
        }
    } reset{site_of, order};
    for (const auto &b : s->bodies)
        for (size_t mi = 0; mi < b.m.size(); ++mi) {
            const Marker &m = b.m[mi];
            const int gm = b.first + (int)mi;
            ns[gm] = m.ns;
            for (int a = 0; a < 3; ++a) put(vel, 3 * (size_t)gm + a, m.vel[a], m.dvel[a]);
            put(eps, gm, m.eps, m.deps);
            put(ds, gm, m.ds, m.dds);
            for (int k = 0; k < m.ns; ++k) {
                const int cell = m.sijk[k][0] + N * (m.sijk[k][1] + M * m.sijk[k][2]);
                scell[(size_t)gm * S + k] = cell;
                put(sdelta, (size_t)gm * S + k, m.delta[k], m.ddelta[k]);
                if (site_of[cell] < 0) {
                    const unsigned fl = g->hflags.empty() ? 0u : g->hflags[cell];
                    if (((fl & kTypeMask) == kVelocity && ((fl >> kNormalShift) & 7) == kNormalEdge) ||
                        (fl & kTypeMask) == kPressure)
                        return fail(LUMAGPU_EUNSUPPORTED,
                                    "IB support on an edge/corner velocity site or a pressure site is not supported");
                    site_of[cell] = (int)order.size();
                    order.push_back(cell);
                    entries.emplace_back();
                }
                entries[site_of[cell]].push_back(gm * S + k);
            }
        }
    std::vector<int> sptr(order.size() + 1, 0), sent;
    for (size_t i = 0; i < order.size(); ++i) {
        sptr[i + 1] = sptr[i] + (int)entries[i].size();
        sent.insert(sent.end(), entries[i].begin(), entries[i].end());
    }
    // Interpolated values and marker forces persist across re-uploads (a marker keeps them
    // when its body moves); they start at zero when the marker set changes size.
    struct Part {
        void **dev;
        const void *src;
        size_t bytes, off;
    };
    Part parts[] = {{(void **)&s->d_scell, scell.data(), scell.size() * sizeof(int), 0},
                    {(void **)&s->d_ns, ns.data(), ns.size() * sizeof(int), 0},
                    {(void **)&s->d_sites, order.data(), order.size() * sizeof(int), 0},
                    {(void **)&s->d_sptr, sptr.data(), sptr.size() * sizeof(int), 0},
                    {(void **)&s->d_sent, sent.data(), sent.size() * sizeof(int), 0},
                    {(void **)&s->d_sdelta, sdelta.data(), sdelta.size() * sizeof(double), 0},
                    {(void **)&s->d_vel, vel.data(), vel.size() * sizeof(double), 0},
                    {(void **)&s->d_eps, eps.data(), eps.size() * sizeof(double), 0},
                    {(void **)&s->d_ds, ds.data(), ds.size() * sizeof(double), 0}};
    size_t total = 0;
    for (Part &q : parts) {
        q.off = total;
        total += (q.bytes + 255) & ~size_t(255);
    }
    if (!s->pack_ev && hipEventCreateWithFlags(&s->pack_ev, hipEventDisableTiming) != hipSuccess)
        return fail(LUMAGPU_EHIP, "IB staging event creation failed");
    if (total > s->pack_cap) {
        // kernels still queued may read the old arena
        HIP_TRY(hipStreamSynchronize(g->stream));
        if (s->d_pack) (void)hipFree(s->d_pack);
        if (s->h_pack) (void)hipHostFree(s->h_pack);
        s->d_pack = s->h_pack = nullptr;
        s->pack_cap = 0;
        const size_t cap = 2 * total;
        if (hipMalloc(&s->d_pack, cap) != hipSuccess || hipHostMalloc(&s->h_pack, cap, hipHostMallocDefault) != hipSuccess)
            return fail(LUMAGPU_ENOMEM, "IB buffer allocation failed");
        s->pack_cap = cap;
    } else {
        // the previous copy out of the staging buffer has finished
        HIP_TRY(hipEventSynchronize(s->pack_ev));
    }
    for (const Part &q : parts) {
        if (q.bytes) std::memcpy(static_cast<char *>(s->h_pack) + q.off, q.src, q.bytes);
        *q.dev = static_cast<char *>(s->d_pack) + q.off;
    }
    if (total) HIP_TRY(hipMemcpyAsync(s->d_pack, s->h_pack, total, hipMemcpyHostToDevice, g->stream));
    HIP_TRY(hipEventRecord(s->pack_ev, g->stream));
    int r = 0;
    size_t *c = s->cap;
    if ((r = dev_resize(&s->d_interp, c + 9, (size_t)nm * 4 * w, g->stream)) ||
        (r = dev_resize(&s->d_force, c + 10, (size_t)nm * 3 * w, g->stream)))
        return r;
    if (!s->d_out && hipMalloc(&s->d_out, 6 * sizeof(double)) != hipSuccess)
        return fail(LUMAGPU_ENOMEM, "IB buffer allocation failed");
    if (order != s->sites) {
        s->sites_dirty = 1;
        s->site_map_ok = 0;
    }
    if (!s->d_site_of && !ibm_has_flexible(g) && hipMalloc(&s->d_site_of, g->ncell * sizeof(int)) != hipSuccess)
        return fail(LUMAGPU_ENOMEM, "IB site map allocation failed");
    s->sites = order;
    s->nsites = (int)order.size();
    return 0;
}

}  // namespace

hipError_t launch_sites(const lumagpu_grid *g, const StepArgs &a, bool odd, int n, const int *cells, int ib_force) {
    switch (g->d.precision) {
        case LUMAGPU_FP64: return launch_sites_f64(g->Q, a, odd, n, cells, ib_force, g->stream);
        case LUMAGPU_FP32S: return launch_sites_f32s(g->Q, a, odd, n, cells, ib_force, g->stream);
        case LUMAGPU_FP64T: return launch_sites_dual(g->Q, a, odd, n, cells, ib_force, g->stream);
        case LUMAGPU_FP32T: return launch_sites_dual32(g->Q, a, odd, n, cells, ib_force, g->stream);
        default: return launch_sites_f16s(g->Q, a, odd, n, cells, ib_force, g->stream);
    }
}

namespace {

IbmSystem *system_of(lumagpu_grid *g) {
    if (!g->ibm) {
        g->ibm = new IbmSystem();
        g->ibm->S = (g->D == 3) ? 27 : 9;
    }
    return g->ibm;
}

int check_body(lumagpu_grid *g, int body) {
    if (!g) return fail(LUMAGPU_EINVAL, "null grid handle");
    if (!g->ibm || body < 0 || body >= (int)g->ibm->bodies.size()) return fail(LUMAGPU_EINVAL, "no such IB body");
    return 0;
}

}  // namespace

int ensure_flags(lumagpu_grid *g) {
    if (g->hflags.size() != g->ncell) g->hflags.assign(g->ncell, (uint8_t)kFluid);
    if (!g->flags && hipMalloc(&g->flags, g->ncell) != hipSuccess)
        return fail(LUMAGPU_ENOMEM, "flag array allocation failed");
    return 0;
}

void ibm_forget_flags(lumagpu_grid *g) {
    if (!g->ibm) return;
    g->ibm->flagged.clear();
    g->ibm->sites_dirty = 1;
}

int ibm_marker_count(const lumagpu_grid *g) {
    if (!g->ibm) return 0;
    int n = 0;
    for (const auto &b : g->ibm->bodies) n += (int)b.m.size();
    return n;
}

bool ibm_has_flexible(const lumagpu_grid *g) {
    if (!g->ibm) return false;
    for (const auto &b : g->ibm->bodies)
        if (b.flexible) return true;
    return false;
}

void ibm_destroy(lumagpu_grid *g) {
    delete g->ibm;
    g->ibm = nullptr;
}

// Sites flagged by the previous application and no longer supports are reset
// (_LBM_resetForces before every application, ObjectManager_ops_ibm.cpp:120-128 and
// GridObj_ops_lbm_optimised.cpp:55-58); current supports are overwritten by the spread, which
// starts from zero.
int clear_gone(lumagpu_grid *g) {
    IbmSystem *s = g->ibm;
    // a gravity reset has rewritten every site's force: nothing to clear
    if (s->sites_dirty && !s->flagged.empty() && !g->all_forced) {
        // site_of is all -1 between uses: mark the current sites, collect, unmark
        std::vector<int> &mark = s->site_of;
        if (mark.size() != g->ncell) mark.assign(g->ncell, -1);
        for (int c : s->sites) mark[c] = 0;
        std::vector<int> gone;
        for (int c : s->flagged)
            if (mark[c] < 0) gone.push_back(c);
        for (int c : s->sites) mark[c] = -1;
        if (!gone.empty()) {
            if (gone.size() > s->gone_cap) {
                HIP_TRY(hipStreamSynchronize(g->stream));
                if (s->d_gone) (void)hipFree(s->d_gone);
                if (s->h_gone) (void)hipHostFree(s->h_gone);
                s->d_gone = s->h_gone = nullptr;
                s->gone_cap = 0;
                const size_t cap = 2 * gone.size();
                if (hipMalloc(&s->d_gone, cap * sizeof(int)) != hipSuccess ||
                    hipHostMalloc(&s->h_gone, cap * sizeof(int), hipHostMallocDefault) != hipSuccess)
                    return fail(LUMAGPU_ENOMEM, "IB buffer allocation failed");
                s->gone_cap = cap;
            }
            std::memcpy(s->h_gone, gone.data(), gone.size() * sizeof(int));
            HIP_TRY(hipMemcpyAsync(s->d_gone, s->h_gone, gone.size() * sizeof(int), hipMemcpyHostToDevice, g->stream));
            IbmArgs b = ibm_args(g);
            const int nb = ((int)gone.size() + 127) / 128;
            if (tangent_form(g)) k_ibm_clear<Dual><<<nb, 128, 0, g->stream>>>(b, (int)gone.size(), s->d_gone);
            else k_ibm_clear<double><<<nb, 128, 0, g->stream>>>(b, (int)gone.size(), s->d_gone);
            HIP_TRY(hipGetLastError());
            // h_gone is free again once the copy has run; the spread that follows needs the clear
            // ordered before it, which the stream gives
            HIP_TRY(hipStreamSynchronize(g->stream));
            for (int c : gone) g->hflags[c] &= (uint8_t)~kForcedBit;
        }
    }
    return 0;
}

hipError_t launch_ibm_interp(const lumagpu_grid *g, const StepArgs &a, const IbmArgs &b, int mode) {
    const bool odd = g->parity == 0;
    switch (g->d.precision) {
        case LUMAGPU_FP64: return launch_ibm_interp_f64(g->Q, a, b, odd, mode, g->stream);
        case LUMAGPU_FP32S: return launch_ibm_interp_f32s(g->Q, a, b, odd, mode, g->stream);
        case LUMAGPU_FP64T: return launch_ibm_interp_dual(g->Q, a, b, odd, mode, g->stream);
        case LUMAGPU_FP32T: return launch_ibm_interp_dual32(g->Q, a, b, odd, mode, g->stream);
        default: return launch_ibm_interp_f16s(g->Q, a, b, odd, mode, g->stream);
    }
}

// ibm_interpolate (:498-591) from the sweep-1 rho, u of the supports, recomputed here from the
// populations (LUMA's u = u_n reset before each sub-iteration) and stored at the support sites;
// mode 0 also computes the marker forces (ibm_computeForce, :598-614) in the same pass
int interpolate(lumagpu_grid *g, int mode) {
    const StepArgs a = make_args(g);
    const IbmArgs b = ibm_args(g);
    const hipError_t e = launch_ibm_interp(g, a, b, mode);
    if (e != hipSuccess) return fail(LUMAGPU_EHIP, std::string("IB interpolation launch: ") + hipGetErrorString(e));
    return 0;
}

// ibm_computeForce (:598-614) from the values of the last interpolation
int compute_force(lumagpu_grid *g) {
    IbmSystem *s = g->ibm;
    IbmArgs b = ibm_args(g);
    if (tangent_form(g)) k_ibm_interp_force<Dual><<<(s->nm + 127) / 128, 128, 0, g->stream>>>(b, 2);
    else k_ibm_interp_force<double><<<(s->nm + 127) / 128, 128, 0, g->stream>>>(b, 2);
    HIP_TRY(hipGetLastError());
    return 0;
}

// ibm_spread (:621-689): gone supports cleared, then every support site gathers its entries
int spread(lumagpu_grid *g) {
    IbmSystem *s = g->ibm;
    if (int r = clear_gone(g)) return r;
    IbmArgs b = ibm_args(g);
    if (tangent_form(g)) k_ibm_spread<Dual><<<(s->nsites + 127) / 128, 128, 0, g->stream>>>(b);
    else k_ibm_spread<double><<<(s->nsites + 127) / 128, 128, 0, g->stream>>>(b);
    HIP_TRY(hipGetLastError());
    if (s->sites_dirty) {
        for (int c : s->sites) g->hflags[c] |= kForcedBit;
        s->flagged = s->sites;
        s->sites_dirty = 0;
        if (!ibm_has_flexible(g))
            if (int r = build_site_map(g)) return r;
    }
    return 0;
}

// Interpolate with forces, then spread, at the current supports.
int apply_core(lumagpu_grid *g) {
    if (int r = interpolate(g, 0)) return r;
    return spread(g);
}

namespace {
__global__ void k_site_map_fill(int *map, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) map[i] = -1;
}
__global__ void k_site_map_set(int *map, const int *sites, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) map[sites[i]] = i;
}
}  // namespace

// the device cell -> support index map of a rigid system (built once its supports settle;
// upload allocated it)
int build_site_map(lumagpu_grid *g) {
    IbmSystem *s = g->ibm;
    if (!s->d_site_of) return 0;
    k_site_map_fill<<<(unsigned)((g->ncell + 255) / 256), 256, 0, g->stream>>>(s->d_site_of, g->ncell);
    if (s->nsites > 0) k_site_map_set<<<(s->nsites + 255) / 256, 256, 0, g->stream>>>(s->d_site_of, s->d_sites, s->nsites);
    HIP_TRY(hipGetLastError());
    s->site_map_ok = 1;
    return 0;
}

void ibm_fold_args(const lumagpu_grid *g, StepArgs &a) {
    const IbmSystem *s = g->ibm;
    a.ib_site_of = s->d_site_of;
    a.ib_sptr = s->d_sptr;
    a.ib_sent = s->d_sent;
    a.ib_sdelta = s->d_sdelta;
    a.ib_force = s->d_force;
    a.ib_eps = s->d_eps;
    a.ib_ds = s->d_ds;
    a.ib_S = s->S;
    a.ib_accumulate = g->all_forced;
}

namespace {

// ibm_moveBodies (:61-104): structural step of every flexible body from the marker forces,
// new supports for flexible bodies, then ds and epsilon of all bodies on the grid.
// The tangent form runs the same beam on Duals: marker forces, epsilon, positions and
// velocities carry their derivatives through the Newton solve.
int move_bodies_tangent(lumagpu_grid *g) {
    IbmSystem *s = g->ibm;
    std::vector<Dual> force((size_t)s->nm * 3);
    HIP_TRY(hipMemcpyAsync(force.data(), s->d_force, force.size() * sizeof(Dual), hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    const FemBeamT<Dual>::Solver solver = [](int n, std::vector<Dual> &a, std::vector<Dual> &x) {
        return lu_solve_band(n, FemBeamT<Dual>::kBand, FemBeamT<Dual>::kBand, a, x);
    };
    for (auto &b : s->bodies) {
        if (!b.flexible || !b.femd) continue;
        const size_t n = b.m.size();
        std::vector<Dual> eps(n), pos(3 * n), pos0(3 * n), vel(3 * n), vkm1(3 * n);
        for (size_t i = 0; i < n; ++i) {
            const Marker &m = b.m[i];
            eps[i] = Dual(m.eps, m.deps);
            for (int a = 0; a < 3; ++a) {
                pos[3 * i + a] = Dual(m.pos[a], m.dpos[a]);
                pos0[3 * i + a] = Dual(m.pos0[a]);
                vel[3 * i + a] = Dual(m.vel[a], m.dvel[a]);
                vkm1[3 * i + a] = Dual(m.vel_km1[a], m.dvel_km1[a]);
            }
        }
        if (int r = b.femd->dynamic(force.data() + 3 * (size_t)b.first, eps.data(), pos.data(), pos0.data(),
                                    vel.data(), vkm1.data(), solver))
            return r;
        for (size_t i = 0; i < n; ++i) {
            Marker &m = b.m[i];
            for (int a = 0; a < 3; ++a) {
                m.pos[a] = pos[3 * i + a].v;
                m.dpos[a] = pos[3 * i + a].d;
                m.vel[a] = vel[3 * i + a].v;
                m.dvel[a] = vel[3 * i + a].d;
                m.vel_km1[a] = vkm1[3 * i + a].v;
                m.dvel_km1[a] = vkm1[3 * i + a].d;
            }
        }
    }
    for (size_t i = 0; i < s->bodies.size(); ++i)
        if (s->bodies[i].flexible) {
            if (int r = find_support(g, s->bodies[i], (int)i)) return r;
            support_tangents(g, s->bodies[i]);
        }
    for (auto &b : s->bodies) compute_ds(g, b);
    if (int r = find_epsilon_all(g)) return r;
    return upload(g);
}

int move_bodies(lumagpu_grid *g) {
    if (tangent_form(g)) return move_bodies_tangent(g);
    IbmSystem *s = g->ibm;
    const size_t nf = (size_t)s->nm * 3;
    if (nf > s->h_force_cap) {
        if (s->h_force) (void)hipHostFree(s->h_force);
        s->h_force = nullptr;
        s->h_force_cap = 0;
        if (hipHostMalloc(&s->h_force, nf * sizeof(double), hipHostMallocDefault) != hipSuccess)
            return fail(LUMAGPU_ENOMEM, "IB force buffer allocation failed");
        s->h_force_cap = nf;
    }
    HIP_TRY(hipMemcpyAsync(s->h_force, s->d_force, nf * sizeof(double), hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    const double *force = s->h_force;
    const FemBeam::Solver solver = [](int n, std::vector<double> &a, std::vector<double> &x) {
        return lu_solve_band(n, FemBeam::kBand, FemBeam::kBand, a, x);
    };
    for (auto &b : s->bodies) {
        if (!b.flexible || !b.fem) continue;
        const size_t n = b.m.size();
        std::vector<double> eps(n), pos(3 * n), pos0(3 * n), vel(3 * n), vkm1(3 * n);
        for (size_t i = 0; i < n; ++i) {
            eps[i] = b.m[i].eps;
            for (int a = 0; a < 3; ++a) {
                pos[3 * i + a] = b.m[i].pos[a];
                pos0[3 * i + a] = b.m[i].pos0[a];
                vel[3 * i + a] = b.m[i].vel[a];
                vkm1[3 * i + a] = b.m[i].vel_km1[a];
            }
        }
        if (int r = b.fem->dynamic(force + 3 * (size_t)b.first, eps.data(), pos.data(), pos0.data(), vel.data(),
                                   vkm1.data(), solver))
            return r;
        for (size_t i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) {
                b.m[i].pos[a] = pos[3 * i + a];
                b.m[i].vel[a] = vel[3 * i + a];
                b.m[i].vel_km1[a] = vkm1[3 * i + a];
            }
    }
    for (size_t i = 0; i < s->bodies.size(); ++i)
        if (s->bodies[i].flexible)
            if (int r = find_support(g, s->bodies[i], (int)i)) return r;
    for (auto &b : s->bodies) compute_ds(g, b);
    if (int r = find_epsilon_all(g)) return r;
    return upload(g);
}

// ibm_checkVelDiff (:988-1032), normalised by L_UX0 dt/dh with L_UX0 = 1
double vel_diff(const lumagpu_grid *g) {
    double res = 0.0;
    for (const auto &b : g->ibm->bodies) {
        if (!b.flexible) continue;
        for (const auto &m : b.m) {
            double acc = 0.0;
            for (int a = 0; a < 3; ++a) acc += std::pow(m.vel[a] - m.vel_km1[a], 2);
            const double v = std::sqrt(acc) / (1.0 * g->d.dt / g->d.dh);
            if (std::fabs(v) > res) res = std::fabs(v);
        }
    }
    return res;
}

}  // namespace

// ObjectManager::ibm_apply (:33-54) with ibm_subIterate (:111-185) for flexible bodies.  The
// macroscopic update at the support sites (ibm_updateMacroscopic, :696-766) happens in the
// step kernel, which recomputes u with +F/2 wherever the force bit is set; every application
// recomputes the sweep-1 rho, u at its supports from the populations, which is LUMA's
// u = u_n reset.
int ibm_apply(lumagpu_grid *g, bool fold_ok) {
    IbmSystem *s = g->ibm;
    if (!s || s->nm == 0) return 0;
    if (!s->initialised) return fail(LUMAGPU_ESTATE, "IB bodies added but lumagpu_ibm_initialise not called");
    if (g->all_forced && ibm_has_flexible(g))
        return fail(LUMAGPU_EUNSUPPORTED, "gravity with flexible bodies is not built (each FSI sub-iteration "
                                          "re-resets gravity from the updated density)");
    // Rigid bodies whose supports are where the last spread left them: the step kernel that
    // follows sums the spread force at each support itself (one launch fewer per step).  Not with
    // gravity, whose reset value the edge-site pre-pass reads from force_xyz before the step.
    if (fold_ok && !ibm_has_flexible(g) && !s->sites_dirty && !g->all_forced && s->site_map_ok &&
        s->flagged == s->sites) {
        if (int r = interpolate(g, 0)) return r;
        g->ib_fold = 1;
        return 0;
    }
    if (int r = apply_core(g)) return r;
    bool flexible = false;
    for (const auto &b : s->bodies) flexible = flexible || b.flexible;
    if (!flexible) return 0;
    if (int r = move_bodies(g)) return r;
    // a pinned count replaces the residual test for this step (lumagpu_ibm_pin_subiterations)
    const int pin = s->pinned_pos < s->pinned.size() ? s->pinned[s->pinned_pos++] : 0;
    int it = 0;
    double res = 0.0;
    do {
        if (int r = apply_core(g)) return r;
        if (int r = move_bodies(g)) return r;
        res = vel_diff(g);
        it++;
    } while (pin > 0 ? it < pin : (res > 1e-4 && it < 10));
    s->sub_it = it;
    s->sub_res = res;
    s->sub_log.push_back(it);
    for (auto &b : s->bodies) {
        if (b.flexible && b.fem) b.fem->commit();
        if (b.flexible && b.femd) b.femd->commit();
    }
    return 0;
}

namespace {

// 16 doubles per marker: position, markerVel, force_xyz, interpMom, interpRho, epsilon, ds and
// the support count -- or, with tan, the derivatives of the same (the count's is 0)
int marker_record(lumagpu_grid *g, int body, double *rec, bool tan) {
    if (int r = check_body(g, body)) return r;
    if (!rec) return fail(LUMAGPU_EINVAL, "null output");
    IbmSystem *s = g->ibm;
    const Body &b = s->bodies[body];
    const size_t n = b.m.size(), w = tangent_form(g) ? 2 : 1, part = tan ? 1 : 0;
    std::vector<double> interp(4 * n * w, 0.0), force(3 * n * w, 0.0);
    if (s->initialised) {
        HIP_TRY(hipMemcpyAsync(interp.data(), s->d_interp + 4 * w * b.first, interp.size() * sizeof(double),
                               hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipMemcpyAsync(force.data(), s->d_force + 3 * w * b.first, force.size() * sizeof(double),
                               hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
    }
    for (size_t i = 0; i < n; ++i) {
        double *o = rec + 16 * i;
        const Marker &m = b.m[i];
        for (int a = 0; a < 3; ++a) {
            o[a] = tan ? m.dpos[a] : m.pos[a];
            o[3 + a] = tan ? m.dvel[a] : m.vel[a];
            o[6 + a] = force[w * (3 * i + a) + part];
            o[9 + a] = interp[w * (4 * i + 1 + a) + part];
        }
        o[12] = interp[w * (4 * i) + part];
        o[13] = tan ? m.deps : m.eps;
        o[14] = tan ? m.dds : m.ds;
        o[15] = tan ? 0.0 : m.ns;
    }
    return 0;
}

int body_force(lumagpu_grid *g, int body, double *out, double *dout) {
    if (int r = check_body(g, body)) return r;
    if (!out) return fail(LUMAGPU_EINVAL, "null output");
    IbmSystem *s = g->ibm;
    if (!s->initialised) return fail(LUMAGPU_ESTATE, "call lumagpu_ibm_initialise first");
    const Body &b = s->bodies[body];
    IbmArgs a = ibm_args(g);
    if (tangent_form(g)) k_ibm_body_force<Dual><<<1, 64, 0, g->stream>>>(a, b.first, b.first + (int)b.m.size(), s->d_out);
    else k_ibm_body_force<double><<<1, 64, 0, g->stream>>>(a, b.first, b.first + (int)b.m.size(), s->d_out);
    HIP_TRY(hipGetLastError());
    double h[6] = {0, 0, 0, 0, 0, 0};
    HIP_TRY(hipMemcpyAsync(h, s->d_out, (tangent_form(g) ? 6 : 3) * sizeof(double), hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    for (int d = 0; d < 3; ++d) {
        out[d] = h[d];
        if (dout) dout[d] = h[3 + d];
    }
    return 0;
}

}  // namespace

}  // namespace lumagpu

using namespace lumagpu;

namespace {
// The IB stages run between the halves of a split step, on bodies that ibm_initialise has set
// up; 1 = no markers on the grid (nothing to do), 0 = ready, < 0 = error
int stage_ready(lumagpu_grid *g, const char *what) {
    if (!g) return fail(LUMAGPU_EINVAL, "null grid handle");
    if (!g->half)
        return fail(LUMAGPU_ESTATE, std::string(what) + " runs between lumagpu_stream_bc_macro and "
                                                        "lumagpu_force_collide_swap");
    if (!g->ibm || g->ibm->nm == 0) return 1;
    if (!g->ibm->initialised) return fail(LUMAGPU_ESTATE, "IB bodies added but lumagpu_ibm_initialise not called");
    return 0;
}
}  // namespace

extern "C" {

int lumagpu_ibm_apply(lumagpu_grid *g) {
    const int r = stage_ready(g, "lumagpu_ibm_apply");
    return r == 1 ? 0 : (r ? r : ibm_apply(g));
}

int lumagpu_ibm_interpolate(lumagpu_grid *g) {
    const int r = stage_ready(g, "lumagpu_ibm_interpolate");
    return r == 1 ? 0 : (r ? r : interpolate(g, 1));
}

int lumagpu_ibm_compute_force(lumagpu_grid *g) {
    const int r = stage_ready(g, "lumagpu_ibm_compute_force");
    return r == 1 ? 0 : (r ? r : compute_force(g));
}

int lumagpu_ibm_spread(lumagpu_grid *g) {
    const int r = stage_ready(g, "lumagpu_ibm_spread");
    return r == 1 ? 0 : (r ? r : spread(g));
}

int lumagpu_ibm_update_macro(lumagpu_grid *g) {
    const int r = stage_ready(g, "lumagpu_ibm_update_macro");
    if (r) return r == 1 ? 0 : r;
    IbmSystem *s = g->ibm;
    const StepArgs a = make_args(g);
    HIP_TRY(launch_sites(g, a, g->parity == 0, s->nsites, s->d_sites, 1));
    return 0;
}

int lumagpu_set_positions(lumagpu_grid *g, const double *x, const double *y, const double *z) {
    if (!g) return fail(LUMAGPU_EINVAL, "null grid handle");
    if (!x || !y || !z) return fail(LUMAGPU_EINVAL, "null position array");
    g->xpos.assign(x, x + g->d.N);
    g->ypos.assign(y, y + g->d.M);
    g->zpos.assign(z, z + g->d.K);
    return 0;
}

int lumagpu_ibm_set_universal_epsilon(lumagpu_grid *g, int on) {
    if (!g) return fail(LUMAGPU_EINVAL, "null grid handle");
    IbmSystem *s = system_of(g);
    if (s->initialised) return fail(LUMAGPU_ESTATE, "set the epsilon mode before lumagpu_ibm_initialise");
    s->universal = on ? 1 : 0;
    return 0;
}

int lumagpu_ibm_add_body(lumagpu_grid *g, int nmarkers, const double *pos, int flexible, int *body) {
    if (!g) return fail(LUMAGPU_EINVAL, "null grid handle");
    if (!g->d.forced) return fail(LUMAGPU_ESTATE, "IB bodies need a grid created with desc.forced = 1");
    if (nmarkers < 2 || !pos) return fail(LUMAGPU_EINVAL, "an IB body needs at least two markers");
    if (flexible) return fail(LUMAGPU_EUNSUPPORTED, "flexible bodies are built with lumagpu_ibm_add_filament");
    IbmSystem *s = system_of(g);
    Body b;
    b.flexible = flexible;
    b.m.resize(nmarkers);
    for (int i = 0; i < nmarkers; ++i)
        for (int a = 0; a < 3; ++a) b.m[i].pos[a] = b.m[i].pos0[a] = pos[3 * i + a];
    s->bodies.push_back(std::move(b));
    s->initialised = 0;
    if (body) *body = (int)s->bodies.size() - 1;
    return 0;
}

int lumagpu_ibm_add_filament(lumagpu_grid *g, const double *start, double length, double height, double depth,
                             double angle_deg, int n_elements, int clamped, double density, double youngs,
                             int *body) {
    if (!g) return fail(LUMAGPU_EINVAL, "null grid handle");
    if (!g->d.forced) return fail(LUMAGPU_ESTATE, "IB bodies need a grid created with desc.forced = 1");
    if (g->D != 2) return fail(LUMAGPU_EUNSUPPORTED, "flexible filaments are two-dimensional");
    if (!start || !(length > 0.0) || n_elements < 1 || !(g->d.dh > 0.0) || !(g->d.dt > 0.0))
        return fail(LUMAGPU_EINVAL, "filament needs start, length > 0, elements >= 1 and desc.dh, desc.dt > 0");
    (void)depth;   // 2D: one lattice spacing deep (FEMBody.cpp:103-105)
    const double pi = 3.14159265358979323846, dh = g->d.dh;
    // Body filament constructor (Body.h:144-187): floor(L/dh) + 1 markers
    const int n = (int)std::floor(length / dh) + 1;
    const double spacing = length / (n - 1);
    const double spacing_h = spacing * std::cos(angle_deg * pi / 180);
    std::vector<double> pos(3 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        pos[3 * i] = start[0] + i * spacing_h * std::cos(0.0 * pi / 180.0);
        pos[3 * i + 1] = start[1] + i * spacing * std::sin(angle_deg * pi / 180.0);
        pos[3 * i + 2] = start[2] + i * spacing_h * std::sin(0.0 * pi / 180.0);
    }
    IbmSystem *s = system_of(g);
    Body b;
    b.flexible = 1;
    b.m.resize(n);
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) b.m[i].pos[a] = b.m[i].pos0[a] = pos[3 * i + a];
    if (tangent_form(g)) {
        b.femd.reset(new FemBeamT<Dual>(start, length, height, angle_deg, n_elements, clamped != 0, Dual(density),
                                        Dual(youngs), dh, g->d.dt, g->d.dm));
        if (b.femd->ib_nodes() != n) return fail(LUMAGPU_EINVAL, "filament IB-FEM node mapping mismatch");
    } else {
        b.fem.reset(new FemBeam(start, length, height, angle_deg, n_elements, clamped != 0, density, youngs, dh,
                                g->d.dt, g->d.dm));
        if (b.fem->ib_nodes() != n) return fail(LUMAGPU_EINVAL, "filament IB-FEM node mapping mismatch");
    }
    s->bodies.push_back(std::move(b));
    s->initialised = 0;
    if (body) *body = (int)s->bodies.size() - 1;
    return 0;
}

int lumagpu_ibm_subiterations(const lumagpu_grid *g, int *iterations, double *residual) {
    if (!g) return fail(LUMAGPU_EINVAL, "null grid handle");
    if (!g->ibm) return fail(LUMAGPU_ESTATE, "no IB bodies on this grid");
    if (iterations) *iterations = g->ibm->sub_it;
    if (residual) *residual = g->ibm->sub_res;
    return 0;
}

int lumagpu_ibm_subiteration_log(const lumagpu_grid *g, int *counts, int max, int *n) {
    if (!g) return fail(LUMAGPU_EINVAL, "null grid handle");
    if (!g->ibm) return fail(LUMAGPU_ESTATE, "no IB bodies on this grid");
    const std::vector<int> &l = g->ibm->sub_log;
    if (n) *n = (int)l.size();
    if (counts)
        for (int i = 0; i < max && i < (int)l.size(); ++i) counts[i] = l[i];
    return 0;
}

int lumagpu_ibm_pin_subiterations(lumagpu_grid *g, const int *counts, int n) {
    if (!g) return fail(LUMAGPU_EINVAL, "null grid handle");
    if (!g->ibm) return fail(LUMAGPU_ESTATE, "no IB bodies on this grid");
    if (n < 0 || (n > 0 && !counts)) return fail(LUMAGPU_EINVAL, "bad sub-iteration schedule");
    std::vector<int> v(counts, counts + n);
    for (int c : v)
        if (c < 1 || c > 10) return fail(LUMAGPU_EINVAL, "pinned sub-iteration counts must lie in 1..10");
    g->ibm->pinned = std::move(v);
    g->ibm->pinned_pos = 0;
    return 0;
}

int lumagpu_ibm_initialise(lumagpu_grid *g) {
    if (!g) return fail(LUMAGPU_EINVAL, "null grid handle");
    if (!g->ibm || g->ibm->bodies.empty()) return fail(LUMAGPU_ESTATE, "no IB bodies on this grid");
    if (!(g->d.dh > 0.0)) return fail(LUMAGPU_EINVAL, "desc.dh must be positive for IB bodies");
    if ((int)g->xpos.size() != g->d.N || (int)g->ypos.size() != g->d.M || (int)g->zpos.size() != g->d.K)
        return fail(LUMAGPU_ESTATE, "grid positions are not set (lumagpu_set_labels or lumagpu_set_positions)");
    if (int r = ensure_flags(g)) return r;
    IbmSystem *s = g->ibm;
    for (size_t i = 0; i < s->bodies.size(); ++i) {
        if (int r = find_support(g, s->bodies[i], (int)i)) return r;
        if (tangent_form(g)) support_tangents(g, s->bodies[i]);
    }
    for (auto &b : s->bodies) compute_ds(g, b);
    if (int r = find_epsilon_all(g)) return r;
    if (int r = upload(g)) return r;
    if (int r = upload_flags(g)) return r;
    HIP_TRY(hipStreamSynchronize(g->stream));
    g->masked = 1;
    s->initialised = 1;
    s->sub_log.clear();
    return 0;
}

int lumagpu_ibm_set_velocity(lumagpu_grid *g, int body, const double *vel) {
    if (int r = check_body(g, body)) return r;
    if (!vel) return fail(LUMAGPU_EINVAL, "null velocity array");
    Body &b = g->ibm->bodies[body];
    for (size_t i = 0; i < b.m.size(); ++i)
        for (int a = 0; a < 3; ++a) {
            b.m[i].vel[a] = vel[3 * i + a];
            b.m[i].dvel[a] = 0.0;
        }
    if (g->ibm->initialised) {
        const size_t w = tangent_form(g) ? 2 : 1;
        std::vector<double> v(3 * b.m.size() * w, 0.0);
        for (size_t i = 0; i < b.m.size(); ++i)
            for (int a = 0; a < 3; ++a) v[w * (3 * i + a)] = b.m[i].vel[a];
        HIP_TRY(hipMemcpyAsync(g->ibm->d_vel + 3 * w * b.first, v.data(), v.size() * sizeof(double),
                               hipMemcpyHostToDevice, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
    }
    return 0;
}

int lumagpu_ibm_set_epsilon(lumagpu_grid *g, int body, const double *eps) {
    if (int r = check_body(g, body)) return r;
    if (!eps) return fail(LUMAGPU_EINVAL, "null epsilon array");
    if (!g->ibm->initialised) return fail(LUMAGPU_ESTATE, "call lumagpu_ibm_initialise first");
    Body &b = g->ibm->bodies[body];
    const size_t w = tangent_form(g) ? 2 : 1;
    std::vector<double> h(b.m.size() * w, 0.0);
    for (size_t i = 0; i < b.m.size(); ++i) {
        b.m[i].eps = eps[i];
        b.m[i].deps = 0.0;
        h[w * i] = eps[i];
    }
    HIP_TRY(hipMemcpyAsync(g->ibm->d_eps + w * b.first, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                           g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    return 0;
}

int lumagpu_ibm_get_markers(lumagpu_grid *g, int body, double *rec) { return marker_record(g, body, rec, false); }

int lumagpu_ibm_get_marker_tangents(lumagpu_grid *g, int body, double *rec) {
    if (int r = check_body(g, body)) return r;
    if (!tangent_form(g)) return fail(LUMAGPU_ESTATE, "marker derivatives exist on tangent-form grids only");
    return marker_record(g, body, rec, true);
}

int lumagpu_ibm_seed_youngs(lumagpu_grid *g, int body, double dE) {
    if (int r = check_body(g, body)) return r;
    if (!tangent_form(g)) return fail(LUMAGPU_ESTATE, "derivative seeds need a tangent-form grid (LUMAGPU_FP64T or LUMAGPU_FP32T)");
    Body &b = g->ibm->bodies[body];
    if (!b.flexible || !b.femd) return fail(LUMAGPU_EINVAL, "body " + std::to_string(body) + " is not a flexible filament");
    b.femd->seed_youngs(dE);
    return 0;
}

int lumagpu_ibm_count(const lumagpu_grid *g, int *nbodies, int body, int *nmarkers) {
    if (!g) return fail(LUMAGPU_EINVAL, "null grid handle");
    const int nb = g->ibm ? (int)g->ibm->bodies.size() : 0;
    if (nbodies) *nbodies = nb;
    if (nmarkers) {
        if (body < 0 || body >= nb) return fail(LUMAGPU_EINVAL, "no such IB body");
        *nmarkers = (int)g->ibm->bodies[body].m.size();
    }
    return 0;
}

int lumagpu_ibm_body_force(lumagpu_grid *g, int body, double out[3]) { return body_force(g, body, out, nullptr); }

int lumagpu_ibm_body_force_tangent(lumagpu_grid *g, int body, double out[3], double dout[3]) {
    if (int r = check_body(g, body)) return r;
    if (!tangent_form(g)) return fail(LUMAGPU_ESTATE, "derivatives exist on tangent-form grids only");
    if (!dout) return fail(LUMAGPU_EINVAL, "null output");
    return body_force(g, body, out, dout);
}

}  // extern "C"
