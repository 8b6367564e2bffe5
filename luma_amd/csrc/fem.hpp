// fem.hpp -- host FP64 structural solver for flexible IB filaments (2D corotational
// Euler-Bernoulli beams, 3 DOF per node), coupled to the IB markers of one body.
//
// Restates LUMA's FEMBody (LUMA/src/FEMBody.cpp:62-424), FEMElement (FEMElement.cpp:61-354)
// and FEMNode (FEMNode.cpp:49-68): consistent mass and linear + geometric stiffness in the
// element frame, Newton-Raphson on the Newmark-beta effective system (alpha 1/4, delta 1/2,
// tolerance 1e-10, at most 20 iterations), clamped or supported first node, and the IB-FEM
// node mapping with under-relaxed marker velocities (L_RELAX = 0.5).  Small matrix products
// sum rows, then columns, then the inner index from 0.0, as GridUtils::matrix_multiply does
// (GridUtils.cpp:257-308); the dense solve is the caller's (the same LU as epsilon).
//
// The scalar type Sc is double, or Dual for the tangent form (the same operations on value
// and derivative; math functions are called unqualified so a Dual finds its own).
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <functional>
#include <type_traits>
#include <vector>

#include "dual.hpp"

namespace lumagpu {

template <class Sc>
class FemBeamT {
  public:
    using Mat6 = std::array<std::array<Sc, 6>, 6>;
    using Vec6 = std::array<Sc, 6>;
    using Solver = std::function<int(int, std::vector<Sc> &, std::vector<Sc> &)>;
    // half-bandwidth of the assembled matrices: an element couples the six DOFs of its two
    // nodes, so no entry lies more than five places from the diagonal
    static constexpr int kBand = 5;

    struct Element {
        Sc length0 = 0, length = 0, area = 0, E = 0, density = 0, angle = 0, I = 0;
        Mat6 T{};
        Vec6 Fint{};
        int dof0 = 0;
        struct Child {
            int node;
            double z1, z2;
        };
        std::vector<Child> children;   // IB nodes integrated over this element
    };

    FemBeamT(const double *start, double length, double height, double angle_deg, int n_el, bool clamped,
             Sc density, Sc E, double dh, double dt, double dm)
        : dh_(dh), dt_(dt), dm_(dm) {
        using std::cos;
        using std::sin;
        const double pi = 3.14159265358979323846;
        ndof_ = (n_el + 1) * 3;
        bc_ = clamped ? 3 : 2;
        const double spacing = length / n_el;
        const double spacing_h = spacing * std::cos(angle_deg * pi / 180);
        nodes_.resize(n_el + 1);
        for (int i = 0; i <= n_el; ++i) {
            Node &nd = nodes_[i];
            nd.p0[0] = start[0] + i * spacing_h * cos(0.0 * pi / 180.0);
            nd.p0[1] = start[1] + i * spacing * sin(angle_deg * pi / 180.0);
            nd.p0[2] = start[2] + i * spacing_h * sin(0.0 * pi / 180.0);
            nd.p = nd.p0;
            nd.a0 = angle_deg * pi / 180.0;
            nd.a = nd.a0;
        }
        const double depth = dh;   // 2D: one lattice spacing deep (FEMBody.cpp:103-105)
        els_.resize(n_el);
        for (int i = 0; i < n_el; ++i) {
            Element &e = els_[i];
            e.length0 = e.length = spacing;
            e.area = height * depth;
            e.angle = angle_deg * pi / 180.0;
            e.I = depth * (height * height * height) / 12.0;
            e.E = E;
            e.density = density;
            e.dof0 = 3 * i;
            set_rotation(e);
        }
        map_ib_nodes((int)std::floor(length / dh) + 1);
        for (auto *v : {&U_, &Un_, &dU_, &V_, &Vn_, &A_, &An_, &R_, &F_}) v->assign(ndof_, 0.0);
        M_.assign((size_t)ndof_ * ndof_, 0.0);
        K_.assign((size_t)ndof_ * ndof_, 0.0);
    }

    int ib_nodes() const { return (int)parent_el_.size(); }
    int ndof() const { return ndof_; }
    int iterations() const { return it_; }
    double residual() const { return value_of(res_); }
    // derivative of every element's Young's modulus (the tangent form's seed; Sc = Dual)
    void seed_youngs(double dE) {
        if constexpr (!std::is_same_v<Sc, double>)
            for (auto &e : els_) e.E.d = dE;
    }

    // FEMBody::dynamicFEM (FEMBody.cpp:136-179) followed by updateIBMarkers (:329-370).
    // force/eps: marker force_xyz and epsilon; pos/pos0/vel/vel_km1: marker state, 3 per marker.
    int dynamic(const Sc *force, const Sc *eps, Sc *pos, const Sc *pos0, Sc *vel, Sc *vel_km1,
                const Solver &solve) {
        U_ = Un_;
        V_ = Vn_;
        A_ = An_;
        update_geometry();
        std::fill(R_.begin(), R_.end(), 0.0);
        for (auto &e : els_) load_vector(e, force, eps);
        it_ = 0;
        do {
            if (int r = newton_iteration(solve)) return r;
            res_ = residual_norm();
            it_++;
        } while (value_of(res_) > 1e-10 && it_ < 20);
        finish_newmark();
        move_markers(pos, pos0, vel, vel_km1);
        return 0;
    }

    // start-of-step values for the next step (ObjectManager_ops_ibm.cpp:150-158)
    void commit() {
        Un_ = U_;
        Vn_ = V_;
        An_ = A_;
    }

  private:
    struct Node {
        std::array<Sc, 3> p0{}, p{};
        Sc a0 = 0, a = 0;
    };

    static constexpr double kAlpha = 0.25, kDelta = 0.5, kRelax = 0.5;

    static void set_rotation(Element &e) {
        using std::cos;
        using std::sin;
        e.T = Mat6{};
        e.T[0][0] = e.T[1][1] = e.T[3][3] = e.T[4][4] = cos(e.angle);
        e.T[0][1] = e.T[3][4] = sin(e.angle);
        e.T[1][0] = e.T[4][3] = -sin(e.angle);
        e.T[2][2] = e.T[5][5] = 1.0;
    }

    // Positions set_rotation can make nonzero: the two 3 x 3 blocks, each an in-plane 2 x 2
    // rotation and a 1.  The others hold +0 from Mat6{}, and the products below skip them: a +0
    // factor gives a zero term, and adding a zero to a sum started at +0 leaves it as it is, so
    // the results equal the full 6 x 6 sums bit for bit.
    static constexpr bool tnz(int i, int j) {
        return (i < 3) == (j < 3) && ((i % 3 < 2 && j % 3 < 2) || (i % 3 == 2 && j % 3 == 2));
    }
    // T^T v
    static Vec6 rotate_back(const Mat6 &T, const Vec6 &v) {
        Vec6 out{};
        for (int r = 0; r < 6; ++r) {
            out[r] = 0.0;
            for (int c = 0; c < 6; ++c)
                if (tnz(c, r)) out[r] += T[c][r] * v[c];
        }
        return out;
    }
    // T^T L T, evaluated as (T^T L) T
    static Mat6 to_global(const Mat6 &T, const Mat6 &L) {
        Mat6 TL{}, G{};
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) {
                TL[r][c] = 0.0;
                for (int k = 0; k < 6; ++k)
                    if (tnz(k, r)) TL[r][c] += T[k][r] * L[k][c];
            }
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) {
                G[r][c] = 0.0;
                for (int k = 0; k < 6; ++k)
                    if (tnz(k, c)) G[r][c] += TL[r][k] * T[k][c];
            }
        return G;
    }
    void add_matrix(std::vector<Sc> &Gm, const Element &e, const Mat6 &G) {
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) Gm[(size_t)(e.dof0 + i) * ndof_ + (e.dof0 + j)] += G[i][j];
    }

    // FEMBody::computeNodeMapping (FEMBody.cpp:378-424)
    void map_ib_nodes(int nIB) {
        const double r = (nodes_.size() - 1.0) / (nIB - 1.0);
        parent_el_.assign(nIB, 0);
        parent_zeta_.assign(nIB, 0.0);
        parent_zeta_[0] = -1.0;
        for (int i = 1; i < nIB; ++i) {
            if (std::fmod(i * r, 1.0) == 0.0) {
                parent_el_[i] = (int)std::floor(i * r) - 1;
                parent_zeta_[i] = 1.0;
            } else {
                parent_el_[i] = (int)std::floor(i * r);
                parent_zeta_[i] = std::fmod(i * r, 1.0) * 2.0 - 1.0;
            }
        }
        for (size_t el = 0; el < els_.size(); ++el)
            for (int node = 0; node < nIB; ++node) {
                double n1 = -1 + 2.0 * ((node - 0.5) * r - (double)el) / 1.0;
                double n2 = -1 + 2.0 * ((node + 0.5) * r - (double)el) / 1.0;
                if ((n1 > -1.0 && n1 < 1.0) || (n2 > -1.0 && n2 < 1.0)) {
                    if (n1 < -1.0) n1 = -1.0;
                    if (n2 > 1.0) n2 = 1.0;
                    els_[el].children.push_back({node, n1, n2});
                }
            }
    }

    // FEMBody::updateFEMValues (:297-325)
    void update_geometry() {
        using std::atan2;
        using std::cos;
        using std::pow;
        using std::sin;
        using std::sqrt;
        for (size_t n = 0; n < nodes_.size(); ++n) {
            for (int d = 0; d < 2; ++d) nodes_[n].p[d] = nodes_[n].p0[d] + U_[n * 3 + d];
            nodes_[n].a = nodes_[n].a0 + U_[n * 3 + 2];
        }
        for (size_t el = 0; el < els_.size(); ++el) {
            Sc v[3], sq = 0.0;
            for (int a = 0; a < 3; ++a) v[a] = nodes_[el + 1].p[a] - nodes_[el].p[a];
            Element &e = els_[el];
            e.angle = atan2(v[1], v[0]);
            for (int a = 0; a < 3; ++a) sq += pow(v[a], 2);
            e.length = sqrt(sq);
            e.T[0][0] = e.T[1][1] = e.T[3][3] = e.T[4][4] = cos(e.angle);
            e.T[0][1] = e.T[3][4] = sin(e.angle);
            e.T[1][0] = e.T[4][3] = -sin(e.angle);
            e.T[2][2] = e.T[5][5] = 1.0;
        }
    }

    // FEMElement::loadVector (FEMElement.cpp:133-173): marker forces integrated against the
    // element shape functions over each IB node's share of the element
    void load_vector(const Element &e, const Sc *force, const Sc *eps) {
        const double forceScale = dm_ / (dt_ * dt_);
        for (const auto &ch : e.children) {
            const double a = ch.z1, b = ch.z2;
            const Sc L = e.length;
            const Sc s = eps[ch.node] * 1.0 * forceScale;
            const Sc fx = force[3 * ch.node] * s, fy = force[3 * ch.node + 1] * s;
            Sc F0 = 0.0, F1 = 0.0;
            F0 += e.T[0][0] * fx;
            F0 += e.T[0][1] * fy;
            F1 += e.T[1][0] * fx;
            F1 += e.T[1][1] * fy;
            const double a2 = a * a, b2 = b * b, a3 = (a * a) * a, b3 = (b * b) * b;
            Vec6 Rl{};
            Rl[0] = F0 * 0.5 * L * (0.5 * b - 0.5 * a + 0.25 * a2 - 0.25 * b2);
            Rl[1] = F1 * 0.5 * L * (0.5 * b - 0.5 * a - a2 * a2 / 16.0 + b2 * b2 / 16.0 + 3.0 * a2 / 8.0 - 3.0 * b2 / 8.0);
            Rl[2] = F1 * 0.5 * L * (L * (-a2 * a2 + b2 * b2) / 32.0 - L * (-a3 + b3) / 24.0 - L * (-a2 + b2) / 16.0 + L * (b - a) / 8.0);
            Rl[3] = F0 * 0.5 * L * (-0.25 * a2 + 0.25 * b2 + 0.5 * b - 0.5 * a);
            Rl[4] = F1 * 0.5 * L * (0.5 * b - 0.5 * a + a2 * a2 / 16.0 - b2 * b2 / 16.0 - 3.0 * a2 / 8.0 + 3.0 * b2 / 8.0);
            Rl[5] = F1 * 0.5 * L * (L * (-a2 * a2 + b2 * b2) / 32.0 + L * (-a3 + b3) / 24.0 - L * (-a2 + b2) / 16.0 - L * (b - a) / 8.0);
            const auto Rg = rotate_back(e.T, Rl);
            for (int k = 0; k < 6; ++k) R_[e.dof0 + k] += Rg[k];
        }
    }

    // FEMElement::forceVector (:273-302): internal forces from the corotated deformation
    void internal_force(Element &e, int id) {
        using std::atan2;
        using std::cos;
        using std::sin;
        const Sc t1 = nodes_[id].a, t2 = nodes_[id + 1].a, ca = cos(e.angle), sa = sin(e.angle);
        const Sc u = (e.length * e.length - e.length0 * e.length0) / (e.length + e.length0);
        const Sc th1 = atan2(ca * sin(t1) - sa * cos(t1), ca * cos(t1) + sa * sin(t1));
        const Sc th2 = atan2(ca * sin(t2) - sa * cos(t2), ca * cos(t2) + sa * sin(t2));
        const Sc F0 = (e.E * e.area / e.length0) * u;
        const Sc M1 = (Sc(2) * e.E * e.I / e.length0) * (Sc(2.0) * th1 + th2);
        const Sc M2 = (Sc(2) * e.E * e.I / e.length0) * (th1 + Sc(2.0) * th2);
        e.Fint = {-F0, (Sc(1.0) / e.length0) * (M1 + M2), M1, F0, -(Sc(1.0) / e.length0) * (M1 + M2), M2};
        const auto Fg = rotate_back(e.T, e.Fint);
        for (int k = 0; k < 6; ++k) F_[e.dof0 + k] += Fg[k];
    }

    // FEMElement::massMatrix (:178-213): consistent beam mass
    void mass(const Element &e) {
        Mat6 Ml{};
        const Sc C1 = e.density * e.area * e.length0 / 420.0, l = e.length0;
        Ml[0][0] = C1 * 140.0;
        Ml[0][3] = C1 * 70.0;
        Ml[1][1] = C1 * 156.0;
        Ml[1][2] = C1 * 22.0 * l;
        Ml[1][4] = C1 * 54;
        Ml[1][5] = C1 * (-13.0 * l);
        Ml[2][2] = C1 * 4.0 * (l * l);
        Ml[2][4] = C1 * 13.0 * l;
        Ml[2][5] = C1 * (-3.0 * (l * l));
        Ml[3][3] = C1 * 140.0;
        Ml[4][4] = C1 * 156.0;
        Ml[4][5] = C1 * (-22.0 * l);
        Ml[5][5] = C1 * 4.0 * (l * l);
        for (int i = 1; i < 6; i++)
            for (int j = 0; j < i; j++) Ml[i][j] = Ml[j][i];
        add_matrix(M_, e, to_global(e.T, Ml));
    }

    // FEMElement::stiffMatrix (:218-268): linear + geometric (axial and shear force) parts
    void stiffness(const Element &e) {
        Mat6 Kl{};
        const Sc E = e.E, A = e.area, I = e.I, l = e.length0, l2 = l * l, l3 = (l * l) * l;
        Kl[0][0] = E * A / l;
        Kl[0][3] = -E * A / l;
        Kl[1][1] = 12.0 * E * I / l3;
        Kl[1][2] = 6.0 * E * I / l2;
        Kl[1][4] = -12.0 * E * I / l3;
        Kl[1][5] = 6.0 * E * I / l2;
        Kl[2][2] = 4.0 * E * I / l;
        Kl[2][4] = -6.0 * E * I / l2;
        Kl[2][5] = 2.0 * E * I / l;
        Kl[3][3] = E * A / l;
        Kl[4][4] = 12.0 * E * I / l3;
        Kl[4][5] = -6.0 * E * I / l2;
        Kl[5][5] = 4.0 * E * I / l;
        for (int i = 1; i < 6; i++)
            for (int j = 0; j < i; j++) Kl[i][j] = Kl[j][i];
        const Sc F0 = -e.Fint[0], V0 = e.Fint[4];
        Kl[0][1] += -V0 / l;
        Kl[0][4] += V0 / l;
        Kl[1][0] += -V0 / l;
        Kl[1][1] += F0 / l;
        Kl[1][3] += V0 / l;
        Kl[1][4] += -F0 / l;
        Kl[3][1] += V0 / l;
        Kl[3][4] += -V0 / l;
        Kl[4][0] += V0 / l;
        Kl[4][1] += -F0 / l;
        Kl[4][3] += -V0 / l;
        Kl[4][4] += F0 / l;
        add_matrix(K_, e, to_global(e.T, Kl));
    }

    // FEMBody::newtonRaphsonIterator (:183-218) with setNewmark (:264-293)
    // The matrices are banded (kBand): entries off the band are written once, as +0, and never
    // again, so the loops below visit the band only.  A skipped term is a product with a +0
    // entry added to a sum started at +0, which leaves the sum as it is: the results equal the
    // full loops' bit for bit.
    int newton_iteration(const Solver &solve) {
        std::fill(F_.begin(), F_.end(), 0.0);
        for (const Element &e : els_)
            for (int i = 0; i < 6; ++i)
                for (int j = 0; j < 6; ++j) {
                    M_[(size_t)(e.dof0 + i) * ndof_ + (e.dof0 + j)] = 0.0;
                    K_[(size_t)(e.dof0 + i) * ndof_ + (e.dof0 + j)] = 0.0;
                }
        for (size_t el = 0; el < els_.size(); ++el) {
            internal_force(els_[el], (int)el);
            mass(els_[el]);
            stiffness(els_[el]);
        }
        const double a0 = 1.0 / (kAlpha * (dt_ * dt_)), a2 = 1.0 / (kAlpha * dt_), a3 = 1.0 / (2.0 * kAlpha) - 1.0;
        std::vector<Sc> meff(ndof_), mf(ndof_, Sc(0.0));
        for (int i = 0; i < ndof_; i++) meff[i] = a0 * (Un_[i] - U_[i]) + a2 * V_[i] + a3 * A_[i];
        auto lo = [&](int r) { return std::max(0, r - kBand); };
        auto hi = [&](int r) { return std::min(ndof_ - 1, r + kBand); };
        for (int r = 0; r < ndof_; ++r)
            for (int c = lo(r); c <= hi(r); ++c) mf[r] += M_[(size_t)r * ndof_ + c] * meff[c];
        for (int i = 0; i < ndof_; i++) {
            F_[i] = R_[i] - F_[i] + mf[i];
            for (int j = lo(i); j <= hi(i); j++) K_[(size_t)i * ndof_ + j] += a0 * M_[(size_t)i * ndof_ + j];
        }
        // GridUtils::solveLinearSystem(K, F, BC): the block past the fixed DOFs
        const int m = ndof_ - bc_;
        std::vector<Sc> a((size_t)m * m, Sc(0.0)), x(m);
        for (int i = 0; i < m; ++i) {
            for (int j = std::max(0, i - kBand); j <= std::min(m - 1, i + kBand); ++j)
                a[(size_t)i * m + j] = K_[(size_t)(bc_ + i) * ndof_ + (bc_ + j)];
            x[i] = F_[bc_ + i];
        }
        if (int r = solve(m, a, x)) return r;
        for (int i = 0; i < bc_; ++i) dU_[i] = 0.0;
        for (int i = 0; i < m; ++i) dU_[bc_ + i] = x[i];
        for (int i = 0; i < ndof_; i++) U_[i] = U_[i] + dU_[i];
        update_geometry();
        return 0;
    }

    // FEMBody::checkNRConvergence (:224-231)
    Sc residual_norm() const {
        using std::sqrt;
        const Sc L = els_[0].length0 * (double)els_.size();
        Sc dp = 0.0;
        for (const Sc &d : dU_) dp += d * d;
        return sqrt(dp / (double)ndof_) / L;
    }

    // FEMBody::finishNewmark (:236-255)
    void finish_newmark() {
        const double a6 = 1.0 / (kAlpha * (dt_ * dt_)), a7 = -1.0 / (kAlpha * dt_), a8 = -(1.0 / (2.0 * kAlpha) - 1.0);
        const double a9 = dt_ * (1.0 - kDelta), a10 = kDelta * dt_;
        for (int i = 0; i < ndof_; i++) {
            const Sc st = a6 * (U_[i] - Un_[i]) + a7 * V_[i] + a8 * A_[i];
            V_[i] = V_[i] + a9 * A_[i] + a10 * st;
            A_[i] = st;
        }
    }

    // FEMElement::shapeFuns (FEMElement.cpp:109-128)
    static void shape(const Element &e, const Vec6 &v, double zeta, Sc out[2]) {
        const double h = (zeta + 1.0) / 2.0, h2 = h * h, h3 = (h * h) * h;
        const double N0 = 1.0 - h, N1 = 1.0 - 3.0 * h2 + 2.0 * h3;
        const Sc N2 = (h - 2.0 * h2 + h3) * e.length;
        const double N3 = h, N4 = 3.0 * h2 - 2.0 * h3;
        const Sc N5 = (-h2 + h3) * e.length;
        out[0] = v[0] * N0 + v[3] * N3;
        out[1] = v[1] * N1 + v[2] * N2 + v[4] * N4 + v[5] * N5;
    }

    // FEMBody::updateIBMarkers (:329-370)
    void move_markers(Sc *pos, const Sc *pos0, Sc *vel, Sc *vel_km1) const {
        for (size_t node = 0; node < parent_el_.size(); ++node) {
            const Element &e = els_[parent_el_[node]];
            const double zeta = parent_zeta_[node];
            Vec6 dU{}, dV{}, lU{}, lV{};
            for (int k = 0; k < 6; ++k) {
                dU[k] = U_[e.dof0 + k];
                dV[k] = V_[e.dof0 + k];
            }
            for (int r = 0; r < 6; ++r) {
                lU[r] = 0.0;
                lV[r] = 0.0;
                for (int c = 0; c < 6; ++c) lU[r] += e.T[r][c] * dU[c];
                for (int c = 0; c < 6; ++c) lV[r] += e.T[r][c] * dV[c];
            }
            Sc sU[2], sV[2], gU[2], gV[2];
            shape(e, lU, zeta, sU);
            shape(e, lV, zeta, sV);
            for (int r = 0; r < 2; ++r) {
                gU[r] = 0.0;
                gV[r] = 0.0;
                for (int c = 0; c < 2; ++c) gU[r] += e.T[c][r] * sU[c];
                for (int c = 0; c < 2; ++c) gV[r] += e.T[c][r] * sV[c];
            }
            const double scale = dt_ / dh_;
            for (int d = 0; d < 2; ++d) {
                const Sc v = gV[d] * scale;
                pos[3 * node + d] = pos0[3 * node + d] + gU[d];
                vel_km1[3 * node + d] = vel[3 * node + d];
                vel[3 * node + d] = kRelax * v + (1.0 - kRelax) * vel_km1[3 * node + d];
            }
        }
    }

    double dh_, dt_, dm_;
    int ndof_ = 0, bc_ = 3, it_ = 0;
    Sc res_ = 0.0;
    std::vector<Node> nodes_;
    std::vector<Element> els_;
    std::vector<int> parent_el_;
    std::vector<double> parent_zeta_;
    std::vector<Sc> U_, Un_, dU_, V_, Vn_, A_, An_, R_, F_, M_, K_;
};

using FemBeam = FemBeamT<double>;

}  // namespace lumagpu
