// geometry.hip -- ObjectManager::io_readInGeomConfig (LUMA/src/ObjectManager_ops_io.cpp:322-800):
// the bodies of input/geometry.config built on the GPU grids.  Covered keywords, TYPE IBM:
// FROM_FILE (point clouds, io_readInCloud :812-1049 with Body::buildFromCloud's voxel filter,
// Body.h:633-785), FILAMENT_ARRAY, CIRCLE_SPHERE, SQUARE_CUBE and PLATE; marker generators
// restate the Body<IBMarker> constructors (LUMA/inc/Body.h:144-500).  Host code.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "internal.hpp"
#include "lumagpu.h"

namespace lumagpu {
namespace {

constexpr double kPi = 3.14159265358979323846;   // L_PI

using Pts = std::vector<double>;   // 3 per marker

void add(Pts &p, double x, double y, double z) {
    p.push_back(x);
    p.push_back(y);
    p.push_back(z);
}

// GridUtils::linspace (GridUtils.h:289-309)
std::vector<double> linspace(double a, double b, int n) {
    if (n < 2) n = 2;
    std::vector<double> v(n);
    const double spacing = (b - a) / static_cast<double>(n - 1);
    for (int i = 0; i < n; ++i) v[i] = a + spacing * i;
    return v;
}

// circle (markers 1..n-1, Body.h:230-243) or Fibonacci sphere (:215-228)
Pts circle_sphere(int D, double dh, const double c[3], double r) {
    Pts p;
    if (D == 3) {
        const int n = (int)std::floor(4.0 * kPi * std::pow(r, 2.0) / std::pow(dh, 2.0));
        const double inc = kPi * (3 - std::sqrt(5));
        const double off = 2.0 / (float)n;
        for (int k = 0; k < n; ++k) {
            const double y = k * off - 1 + (off / 2);
            const double rr = std::sqrt(1 - y * y);
            const double phi = k * inc;
            add(p, c[0] + (std::cos(phi) * rr * r), y * r + c[1], c[2] + (std::sin(phi) * rr * r));
        }
    } else {
        const int n = (int)std::floor(2.0 * kPi * r / dh);
        const std::vector<double> th = linspace(0.0, 2.0 * kPi - (2.0 * kPi / (double)n), n);
        for (size_t i = 1; i < th.size(); ++i) add(p, c[0] + r * std::cos(th[i]), c[1] + r * std::sin(th[i]), c[2]);
    }
    return p;
}

// square / cuboid surface (Body.h:262-380)
Pts square_cube(int D, double dh, const double c[3], const double dims[3], const double ang[2]) {
    Pts p;
    const int nl = (int)std::floor(dims[0] / dh) + 1, nh = (int)std::floor(dims[1] / dh) + 1;
    const int nd = (int)std::floor(dims[2] / dh) + 1;
    const double sl = dims[0] / (nl - 1), sh = dims[1] / (nh - 1), sd = dims[2] / (nd - 1);
    const double a0 = ang[0] * kPi / 180, a1 = ang[1] * kPi / 180;
    for (int i = 0; i < nl; i++)
        for (int j = 0; j < nh; j++) {
            if (D == 3) {
                for (int k = 0; k < nd; k++) {
                    const double x = -dims[0] / 2.0 + i * sl, y = -dims[1] / 2.0 + j * sh, z = -dims[2] / 2.0 + k * sd;
                    if ((i == 0 || i == nl - 1) || (j == 0 || j == nh - 1) || (k == 0 || k == nd - 1)) {
                        const double xd = (x * std::cos(a0) - y * std::sin(a0)) * std::cos(a1) - z * std::sin(a1);
                        const double yd = (y * std::cos(a0) + x * std::sin(a0));
                        const double zd = (x * std::cos(a0) - y * std::sin(a0)) * std::sin(a1) + z * std::cos(a1);
                        add(p, xd + c[0], yd + c[1], zd + c[2]);
                    }
                }
            } else {
                const double x = -dims[0] / 2.0 + i * sl, y = -dims[1] / 2.0 + j * sh, z = 0.0;
                if ((i == 0 || i == nl - 1) || (j == 0 || j == nh - 1)) {
                    const double xd = x * std::cos(a0) - y * std::sin(a0);
                    const double yd = x * std::sin(a0) + y * std::cos(a0);
                    add(p, xd + c[0], yd + c[1], z + c[2]);
                }
            }
        }
    return p;
}

// straight filament (Body.h:144-187); the horizontal angle is always 0 (the reference tests
// L_DIM, which is never defined)
Pts filament(double dh, const double s[3], double length, double angle_v) {
    Pts p;
    const int n = (int)std::floor(length / dh) + 1;
    const double spacing = length / (n - 1);
    const double spacing_h = spacing * std::cos(angle_v * kPi / 180);
    const double ah = 0.0;
    for (int i = 0; i < n; i++)
        add(p, s[0] + i * spacing_h * std::cos(ah * kPi / 180.0), s[1] + i * spacing * std::sin(angle_v * kPi / 180.0),
            s[2] + i * spacing_h * std::sin(ah * kPi / 180.0));
    return p;
}

// plate of markers, 3D (Body.h:410-490): floor(L/dh)+1 by floor(W/dh)+1 markers in the x-z plane
// about the centre, rotated by Tz then Tx (the product R = Tz Ty Tx is built there but unused)
Pts plate(double dh, const double c[3], double length, double width, const double ang[3]) {
    Pts p;
    const int nl = (int)std::floor(length / dh) + 1, nw = (int)std::floor(width / dh) + 1;
    const double sl = length / (nl - 1), sw = width / (nw - 1);
    const double tx = ang[0] * kPi / 180.0, tz = ang[2] * kPi / 180.0;
    const double Tx[3][3] = {{1.0, 0.0, 0.0}, {0.0, std::cos(tx), -std::sin(tx)}, {0.0, std::sin(tx), std::cos(tx)}};
    const double Tz[3][3] = {{std::cos(tz), -std::sin(tz), 0.0}, {std::sin(tz), std::cos(tz), 0.0}, {0.0, 0.0, 1.0}};
    // GridUtils::matrix_multiply (GridUtils.cpp:257-276): each row summed from 0.0 over the columns
    auto mul = [](const double A[3][3], const double x[3], double out[3]) {
        for (int r = 0; r < 3; ++r) {
            out[r] = 0.0;
            for (int k = 0; k < 3; ++k) out[r] += A[r][k] * x[k];
        }
    };
    for (int i = 0; i < nl; i++)
        for (int j = 0; j < nw; j++) {
            const double q0[3] = {-length / 2.0 + i * sl, 0.0, -width / 2.0 + j * sw};
            double q1[3], q2[3];
            mul(Tz, q0, q1);
            mul(Tx, q1, q2);
            add(p, q2[0] + c[0], q2[1] + c[1], q2[2] + c[2]);
        }
    return p;
}

// io_readInCloud (ObjectManager_ops_io.cpp:812-1049) and Body::buildFromCloud (Body.h:759-785)
// for one IBM body on grid g.  refc[a]: reference is a CENTRE (else START); ref[a], length and
// dir: the reference point, the body's extent and the direction it is measured along.
int cloud_body(const lumagpu_grid *g, const std::string &path, const bool refc[3], const double ref[3], double length,
               int dir, Pts &out) {
    if (g->xpos.size() != (size_t)g->d.N || g->ypos.size() != (size_t)g->d.M || g->zpos.size() != (size_t)g->d.K)
        return fail(LUMAGPU_ESTATE, "grid positions are not set (lumagpu_set_labels / lumagpu_set_positions)");
    std::ifstream file(path);
    if (!file) return fail(LUMAGPU_EINVAL, "Error opening cloud input file: " + path + ". Exiting.");
    const int D = g->D;
    std::vector<double> P[3];
    // LUMA's loop (:871-898): a line that holds no number leaves the previous point's values,
    // so a file ending in a newline repeats its last point
    double tx = 0.0, ty = 0.0, tz = 0.0;
    while (!file.eof()) {
        std::string line;
        std::getline(file, line, '\n');
        std::istringstream iss(line);
        iss >> tx;
        iss >> ty;
        iss >> tz;
        P[0].push_back(tx);
        P[1].push_back(ty);
        P[2].push_back(D == 3 ? tz : 0.0);
    }
    if (P[0].empty()) return fail(LUMAGPU_EINVAL, "Failed to read object data from cloud input file.");
    const double dc = g->d.dh, small = 1e-8;   // L_SMALL_NUMBER (stdafx.h:117)
    double r[3];
    for (int a = 0; a < 3; ++a) r[a] = std::round(ref[a] / dc) * dc;
    const double blen = std::round(length / dc) * dc;
    double lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = *std::min_element(P[a].begin(), P[a].end());
        hi[a] = *std::max_element(P[a].begin(), P[a].end());
    }
    const double scale = (blen - 2 * small * dc) / std::fabs(hi[dir] - lo[dir]);
    double shift[3];
    for (int a = 0; a < 3; ++a) {
        if (refc[a]) {
            const double rc = r[a] + dc / 2.0;
            double dist = scale * std::fabs(hi[a] - lo[a]);
            dist = std::round(dist / dc) * dc;
            shift[a] = (rc - (dist / 2.0)) - scale * lo[a];
        } else {
            shift[a] = (r[a] + small * dc) - scale * lo[a];
        }
    }
    // scale and shift (z only in 3D), then keep the points on the grid (GridUtils::isOnThisRank,
    // serial branch: inside the level-0 edges, and a voxel of this grid)
    struct Mk {
        double pos[3];
        int vox[3], count;
    };
    std::vector<Mk> mk;
    auto voxel = [&](const double q[3], int v[3]) {
        v[0] = (int)std::floor((q[0] - g->d.origin[0]) / dc);
        v[1] = (int)std::floor((q[1] - g->d.origin[1]) / dc);
        v[2] = (D == 3) ? (int)std::floor((q[2] - g->d.origin[2]) / dc) : 0;
        return v[0] >= 0 && v[0] < g->d.N && v[1] >= 0 && v[1] < g->d.M && v[2] >= 0 && v[2] < g->d.K;
    };
    auto on_grid = [&](const double q[3]) {
        if (!(0.0 <= q[0] && q[0] < g->d.xmax && 0.0 <= q[1] && q[1] < g->d.ymax)) return false;
        if (D == 3 && !(0.0 <= q[2] && q[2] < g->d.zmax)) return false;
        int v[3];
        return voxel(q, v);
    };
    // Body::isInVoxel (Body.h:696-727): within half a spacing of the voxel centre on every axis
    auto in_voxel = [&](const double q[3], const Mk &m) {
        const double h = dc / 2.0;
        const double vx = g->xpos[m.vox[0]], vy = g->ypos[m.vox[1]], vz = g->zpos[m.vox[2]];
        return (q[0] >= vx - h && q[0] < vx + h) && (q[1] >= vy - h && q[1] < vy + h) && (q[2] >= vz - h && q[2] < vz + h);
    };
    // running mean of the points of a voxel (passToVoxelFilter, Body.h:633-688)
    auto absorb = [](Mk &m, const double q[3]) {
        m.count++;
        for (int a = 0; a < 3; ++a) m.pos[a] = ((m.pos[a] * (m.count - 1)) + q[a]) / m.count;
    };
    int cur = 0;
    for (size_t i = 0; i < P[0].size(); ++i) {
        double q[3] = {P[0][i] * scale + shift[0], P[1][i] * scale + shift[1], P[2][i]};
        if (D == 3) q[2] = P[2][i] * scale + shift[2];
        if (!on_grid(q)) continue;
        if (mk.empty()) {
            Mk m{{q[0], q[1], q[2]}, {0, 0, 0}, 1};
            voxel(q, m.vox);
            mk.push_back(m);
            cur = 0;
            continue;
        }
        if (in_voxel(q, mk[cur])) {
            absorb(mk[cur], q);
            continue;
        }
        int v[3];
        voxel(q, v);
        int found = -1;
        for (size_t k = 0; k < mk.size(); ++k)
            if (mk[k].vox[0] == v[0] && mk[k].vox[1] == v[1] && mk[k].vox[2] == v[2]) {
                found = (int)k;
                break;
            }
        if (found >= 0) {
            cur = found;
            absorb(mk[cur], q);
        } else {
            Mk m{{q[0], q[1], q[2]}, {v[0], v[1], v[2]}, 1};
            mk.push_back(m);
            cur = (int)mk.size() - 1;
        }
    }
    out.clear();
    for (const Mk &m : mk) add(out, m.pos[0], m.pos[1], m.pos[2]);
    return 0;
}

lumagpu_grid *grid_at(lumagpu_grid *const *grids, int n, int lev, int reg) {
    if (reg != 0) return nullptr;
    for (int i = 0; i < n; ++i)
        if (grids[i] && grids[i]->d.level == lev) return grids[i];
    return nullptr;
}

}  // namespace
}  // namespace lumagpu

using namespace lumagpu;

extern "C" int lumagpu_ibm_read_geometry(lumagpu_grid *const *grids, int ngrids, const char *path, const double *shift,
                                         int *nbodies) {
    if (!grids || ngrids <= 0 || !path || !shift) return fail(LUMAGPU_EINVAL, "null argument");
    std::ifstream in(path);
    if (!in) return fail(LUMAGPU_EINVAL, "Error opening geometry configuration file. Exiting.");
    std::stringstream all;
    std::string line;
    while (std::getline(in, line)) {
        const size_t p = line.find_first_not_of(" \t\r");
        if (p == std::string::npos || line[p] == '#') continue;
        all << line << "\n";
    }
    // cloud files sit next to the geometry file (LUMA reads both from ./input/)
    std::string dir = path;
    const size_t slash = dir.find_last_of('/');
    dir = (slash == std::string::npos) ? std::string(".") : dir.substr(0, slash);
    int count = 0;
    std::string key;
    while (all >> key) {
        std::string type, cloud;
        int lev, reg;
        all >> type;
        if (key == "FROM_FILE") all >> cloud;
        all >> lev >> reg;
        if (!all) return fail(LUMAGPU_EINVAL, "malformed geometry line: " + key);
        if (type != "IBM")
            return fail(LUMAGPU_EUNSUPPORTED, key + " " + type + " bodies are not built by this build (IBM bodies only: "
                                                                "bounce-back and BFL walls are not on the path)");
        lumagpu_grid *g = grid_at(grids, ngrids, lev, reg);
        if (!g) return fail(LUMAGPU_EINVAL, "no grid at level " + std::to_string(lev) + " region " + std::to_string(reg));
        const int D = g->D;
        const double dh = g->d.dh;
        int body = -1;
        if (key == "FROM_FILE") {
            std::string rt[3], direction, flex, bc;
            double ref[3], length;
            all >> rt[0] >> ref[0] >> rt[1] >> ref[1] >> rt[2] >> ref[2] >> length >> direction >> flex >> bc;
            if (!all) return fail(LUMAGPU_EINVAL, "malformed FROM_FILE line");
            bool refc[3];
            for (int a = 0; a < 3; ++a) {
                if (rt[a] != "CENTRE" && rt[a] != "START")
                    return fail(LUMAGPU_EINVAL, "Unknown reference type in geometry file. Exiting.");
                refc[a] = rt[a] == "CENTRE";
            }
            const int sd = direction == "X" ? 0 : direction == "Y" ? 1 : direction == "Z" ? 2 : -1;
            if (sd < 0) return fail(LUMAGPU_EINVAL, "unknown scaling direction " + direction);
            if (flex == "FLEXIBLE" || flex == "MOVABLE")
                return fail(LUMAGPU_EUNSUPPORTED, "point-cloud IB bodies are built rigid here (" + flex + " requested)");
            Pts p;
            if (int rc = cloud_body(g, dir + "/" + cloud, refc, ref, length, sd, p)) return rc;
            if (int rc = lumagpu_ibm_add_body(g, (int)(p.size() / 3), p.data(), 0, &body)) return rc;
        } else if (key == "PLATE") {
            double c[3], length, width, ang[3];
            std::string flex;
            all >> c[0] >> c[1] >> c[2] >> length >> width >> ang[0] >> ang[1] >> ang[2] >> flex;
            if (!all) return fail(LUMAGPU_EINVAL, "malformed PLATE line");
            if (D != 3) return fail(LUMAGPU_EINVAL, "Can't build plate in 2D. Exiting.");
            if (flex == "FLEXIBLE") return fail(LUMAGPU_EINVAL, "Plate cannot be flexible. Exiting.");
            for (int a = 0; a < 3; ++a) c[a] += shift[a];
            const Pts p = plate(dh, c, length, width, ang);
            if (int rc = lumagpu_ibm_add_body(g, (int)(p.size() / 3), p.data(), 0, &body)) return rc;
        } else if (key == "CIRCLE_SPHERE") {
            double c[3], r;
            std::string flex;
            all >> c[0] >> c[1] >> c[2] >> r >> flex;
            if (!all) return fail(LUMAGPU_EINVAL, "malformed CIRCLE_SPHERE line");
            if (flex == "FLEXIBLE") return fail(LUMAGPU_EINVAL, "Circle/sphere cannot be flexible. Exiting.");
            for (int a = 0; a < 3; ++a) c[a] += shift[a];
            const Pts p = circle_sphere(D, dh, c, r);
            if (int rc = lumagpu_ibm_add_body(g, (int)(p.size() / 3), p.data(), 0, &body)) return rc;
        } else if (key == "SQUARE_CUBE") {
            double c[3], dims[3], ang[2];
            std::string flex;
            all >> c[0] >> c[1] >> c[2] >> dims[0] >> dims[1] >> dims[2] >> ang[0] >> ang[1] >> flex;
            if (!all) return fail(LUMAGPU_EINVAL, "malformed SQUARE_CUBE line");
            if (flex == "FLEXIBLE") return fail(LUMAGPU_EINVAL, "Circle/sphere cannot be flexible. Exiting.");
            for (int a = 0; a < 3; ++a) c[a] += shift[a];
            const Pts p = square_cube(D, dh, c, dims, ang);
            if ((D == 3 && p.size() / 3 < 24) || (D == 2 && p.size() / 3 < 8))
                return fail(LUMAGPU_EINVAL, "Resolution not high enough to build square this small. Change its "
                                            "dimensions. Exiting.");
            if (int rc = lumagpu_ibm_add_body(g, (int)(p.size() / 3), p.data(), 0, &body)) return rc;
        } else if (key == "FILAMENT_ARRAY") {
            int nfil;
            double s[3], sp[3], length, height, depth, av, ah, density, E;
            std::string flex, nel, bc;
            all >> nfil >> s[0] >> s[1] >> s[2] >> sp[0] >> sp[1] >> sp[2] >> length >> height >> depth >> av >> ah >>
                flex >> nel >> bc >> density >> E;
            if (!all) return fail(LUMAGPU_EINVAL, "malformed FILAMENT_ARRAY line");
            const int nelem = (nel == "CONFORMING") ? (int)std::floor(length / dh) : (int)std::stod(nel);
            const bool clamped = bc == "CLAMPED";
            for (int i = 0; i < nfil; ++i) {
                // only the first D components are set (ObjectManager_ops_io.cpp:540-545)
                double pos[3] = {0.0, 0.0, 0.0};
                for (int d = 0; d < D; ++d) pos[d] = s[d] + shift[d] + i * sp[d];
                if (flex == "FLEXIBLE") {
                    if (int rc = lumagpu_ibm_add_filament(g, pos, length, height, depth, av, nelem, clamped ? 1 : 0,
                                                          density, E, &body))
                        return rc;
                } else {
                    const Pts p = filament(dh, pos, length, av);
                    if (int rc = lumagpu_ibm_add_body(g, (int)(p.size() / 3), p.data(), 0, &body)) return rc;
                }
                ++count;
            }
            continue;
        } else {
            return fail(LUMAGPU_EINVAL, "unknown geometry keyword " + key);
        }
        ++count;
    }
    if (nbodies) *nbodies = count;
    return 0;
}
