// geometry.hip -- ObjectManager::io_readInGeomConfig (LUMA/src/ObjectManager_ops_io.cpp:322-800):
// the bodies of input/geometry.config built on the GPU grids.  Covered keywords: CIRCLE_SPHERE,
// SQUARE_CUBE and FILAMENT_ARRAY with TYPE IBM; marker generators restate the Body<IBMarker>
// constructors (LUMA/inc/Body.h:144-380).  Host code.
#include <hip/hip_runtime.h>

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
    int count = 0;
    std::string key;
    while (all >> key) {
        std::string type;
        int lev, reg;
        all >> type >> lev >> reg;
        if (!all) return fail(LUMAGPU_EINVAL, "malformed geometry line: " + key);
        if (key == "FROM_FILE" || key == "PLATE" || type != "IBM")
            return fail(LUMAGPU_EUNSUPPORTED, key + " " + type + " bodies are not built by this build (IBM prefab "
                                                                "circle/sphere, square/cube and filament arrays only)");
        lumagpu_grid *g = grid_at(grids, ngrids, lev, reg);
        if (!g) return fail(LUMAGPU_EINVAL, "no grid at level " + std::to_string(lev) + " region " + std::to_string(reg));
        const int D = g->D;
        const double dh = g->d.dh;
        int body = -1;
        if (key == "CIRCLE_SPHERE") {
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
