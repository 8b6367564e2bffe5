// io.hip -- LUMA's on-disk formats for a grid's state: the restart file (GridObj::io_restart,
// LUMA/src/GridObj_ops_io.cpp:406-590), the io_lite row dump (GridObj::io_lite, :706-810) and
// the per-body lift/drag file (ObjectManager::io_writeLiftDrag, ObjectManager_ops_io.cpp:64-122).
// Host code: the state is downloaded in LUMA's layout and written line by line.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "ibm_system.hpp"
#include "internal.hpp"
#include "lumagpu.h"

namespace lumagpu {
namespace {

// _LBM_equilibrium_opt (GridObj_ops_lbm_optimised.cpp:674-705) on the host
double feq_host(const lumagpu_grid *g, int v, double rho, const double *u) {
    const double cs = 1.0 / std::sqrt(3.0);
    const double cs2 = cs * cs, den2 = 2.0 * (cs * cs) * (cs * cs);
    int c[3];
    double w;
    if (g->Q == 9) {
        c[0] = Lattice<9>::c(v, 0);
        c[1] = Lattice<9>::c(v, 1);
        c[2] = 0;
        w = Lattice<9>::w(v);
    } else {
        c[0] = Lattice<19>::c(v, 0);
        c[1] = Lattice<19>::c(v, 1);
        c[2] = Lattice<19>::c(v, 2);
        w = Lattice<19>::w(v);
    }
    double A, B;
    if (g->D == 3) {
        A = (c[0] * u[0]) + (c[1] * u[1]) + (c[2] * u[2]);
        B = (double(c[0] * c[0]) - cs2) * (u[0] * u[0]) + (double(c[1] * c[1]) - cs2) * (u[1] * u[1]) +
            (double(c[2] * c[2]) - cs2) * (u[2] * u[2]) + double(2 * c[0] * c[1]) * u[0] * u[1] +
            double(2 * c[0] * c[2]) * u[0] * u[2] + double(2 * c[1] * c[2]) * u[1] * u[2];
    } else {
        A = (c[0] * u[0]) + (c[1] * u[1]);
        B = (double(c[0] * c[0]) - cs2) * (u[0] * u[0]) + (double(c[1] * c[1]) - cs2) * (u[1] * u[1]) +
            double(2 * c[0] * c[1]) * u[0] * u[1];
    }
    return rho * w * (1.0 + (A / cs2) + (B / den2));
}

struct HostState {
    std::vector<double> f, rho, u;
};

int download_state(lumagpu_grid *g, HostState &h) {
    h.f.resize(g->ncell * g->Q);
    h.rho.resize(g->ncell);
    h.u.resize(g->ncell * g->D);
    if (int r = lumagpu_download_f(g, h.f.data())) return r;
    return lumagpu_download_macro(g, h.rho.data(), h.u.data());
}

int check_positions(const lumagpu_grid *g) {
    if (g->xpos.size() != (size_t)g->d.N || g->ypos.size() != (size_t)g->d.M || g->zpos.size() != (size_t)g->d.K)
        return fail(LUMAGPU_ESTATE, "grid positions are not set (lumagpu_set_labels / lumagpu_set_positions)");
    return 0;
}

// index of a position on one axis (the site whose position matches to within 1e-6 dh): the
// uniformly spaced guess first, a scan only if the positions are not uniform
int locate(const std::vector<double> &pos, double x, double dh) {
    const double tol = 1e-6 * dh;
    if (pos.empty()) return -1;
    const long guess = std::lround((x - pos[0]) / dh);
    if (guess >= 0 && guess < (long)pos.size() && std::fabs(pos[guess] - x) <= tol) return (int)guess;
    for (size_t i = 0; i < pos.size(); ++i)
        if (std::fabs(pos[i] - x) <= tol) return (int)i;
    return -1;
}

}  // namespace
}  // namespace lumagpu

using namespace lumagpu;

extern "C" {

int lumagpu_write_restart(lumagpu_grid *g, const char *path, int append, int region) {
    if (!g || !path) return fail(LUMAGPU_EINVAL, "null argument");
    if (int r = check_positions(g)) return r;
    HostState h;
    if (int r = download_state(g, h)) return r;
    std::ofstream out(path, append ? std::ios::out | std::ios::app : std::ios::out);
    if (!out) return fail(LUMAGPU_EINVAL, std::string("cannot open ") + path);
    const int N = g->d.N, M = g->d.M, K = g->d.K, Q = g->Q, D = g->D;
    const double dh = g->d.dh, dt = g->d.dt, om = g->d.omega;
    char buf[64];
    auto put = [&](double v) {
        std::snprintf(buf, sizeof buf, "%.17g\t", v);
        out << buf;
    };
    for (int k = 0; k < K; ++k)
        for (int j = 0; j < M; ++j)
            for (int i = 0; i < N; ++i) {
                const size_t id = (size_t)k + (size_t)j * K + (size_t)i * K * M;
                out << g->d.level << "\t" << region << "\t";
                put(g->xpos[i]);
                put(g->ypos[j]);
                put(g->zpos[k]);
                double u[3] = {0.0, 0.0, 0.0};
                for (int v = 0; v < D; ++v) {
                    u[v] = h.u[v + D * id];
                    put((u[v] * dh) / dt);   // GridUnits::ulbm2ud
                }
                put(h.rho[id]);
                for (int v = 0; v < Q; ++v) {
                    const double fe = feq_host(g, v, h.rho[id], u);
                    put(((h.f[v + Q * id] - fe) * om) / (fe * dt));
                }
                out << "\n";
            }
    if (!out) return fail(LUMAGPU_EINVAL, std::string("write failed: ") + path);
    return 0;
}

int lumagpu_read_restart(lumagpu_grid *g, const char *path, int region) {
    if (!g || !path) return fail(LUMAGPU_EINVAL, "null argument");
    if (int r = check_positions(g)) return r;
    std::ifstream in(path);
    if (!in) return fail(LUMAGPU_EINVAL, "Error opening LBM restart file. Exiting.");
    const int M = g->d.M, K = g->d.K, Q = g->Q, D = g->D;
    const double dh = g->d.dh, dt = g->d.dt, om = g->d.omega;
    HostState h;
    h.f.assign(g->ncell * Q, 0.0);
    h.rho.assign(g->ncell, 0.0);
    h.u.assign(g->ncell * D, 0.0);
    std::vector<char> seen(g->ncell, 0);
    size_t nseen = 0;
    std::string line;
    while (std::getline(in, line)) {
        if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
        std::istringstream iss(line);
        int lev, reg;
        double x, y, z;
        if (!(iss >> lev >> reg >> x >> y >> z)) return fail(LUMAGPU_EINVAL, "malformed restart line");
        if (lev != g->d.level || reg != region) continue;
        const int i = locate(g->xpos, x, dh), j = locate(g->ypos, y, dh), k = locate(g->zpos, z, dh);
        if (i < 0 || j < 0 || k < 0) continue;   // a site of another rank (GridUtils::isOnThisRank)
        const size_t id = (size_t)k + (size_t)j * K + (size_t)i * K * M;
        double u[3] = {0.0, 0.0, 0.0};
        for (int v = 0; v < D; ++v) {
            double ud;
            if (!(iss >> ud)) return fail(LUMAGPU_EINVAL, "malformed restart line");
            u[v] = (ud * dt) / dh;   // GridUnits::ud2ulbm
            h.u[v + D * id] = u[v];
        }
        if (!(iss >> h.rho[id])) return fail(LUMAGPU_EINVAL, "malformed restart line");
        for (int v = 0; v < Q; ++v) {
            double fn;
            if (!(iss >> fn)) return fail(LUMAGPU_EINVAL, "malformed restart line");
            const double fe = feq_host(g, v, h.rho[id], u);
            h.f[v + Q * id] = fe * (1 + (dt * fn) / om);
        }
        if (!seen[id]) {
            seen[id] = 1;
            ++nseen;
        }
    }
    if (nseen != g->ncell)
        return fail(LUMAGPU_EINVAL, "restart file covers " + std::to_string(nseen) + " of " +
                                        std::to_string(g->ncell) + " sites of this grid");
    if (int r = lumagpu_upload_f(g, h.f.data())) return r;
    return lumagpu_upload_macro(g, h.rho.data(), h.u.data());
}

int lumagpu_io_lite(lumagpu_grid *g, const char *path, int rank, int precision) {
    if (!g || !path) return fail(LUMAGPU_EINVAL, "null argument");
    if (int r = check_positions(g)) return r;
    if (precision < 0 || precision > 30) return fail(LUMAGPU_EINVAL, "output precision out of range");
    HostState h;
    if (int r = download_state(g, h)) return r;
    FILE *fp = std::fopen(path, "w");
    if (!fp) return fail(LUMAGPU_EINVAL, std::string("cannot open ") + path);
    const int N = g->d.N, M = g->d.M, K = g->d.K, Q = g->Q, D = g->D;
    auto put = [&](double v) { std::fprintf(fp, "%.*f\t", precision, v); };
    for (int k = 0; k < K; ++k)
        for (int j = 0; j < M; ++j)
            for (int i = 0; i < N; ++i) {
                const size_t id = (size_t)k + (size_t)j * K + (size_t)i * K * M;
                const size_t cell = (size_t)i + (size_t)N * ((size_t)j + (size_t)M * k);
                const int ty = g->hflags.size() == g->ncell ? (g->hflags[cell] & kTypeMask) : kFluid;
                std::fprintf(fp, "%d\t%d\t", rank, ty);
                put(g->xpos[i]);
                put(g->ypos[j]);
                put(g->zpos[k]);
                put(h.rho[id]);
                for (int v = 0; v < D; ++v) put(h.u[v + D * id]);
                if (D != 3) std::fprintf(fp, "%s\t", std::to_string(0.0).c_str());
                // f, then LUMA's fNew columns: this build keeps one population array, so the
                // post-collision state is written in both blocks
                for (int rep = 0; rep < 2; ++rep)
                    for (int v = 0; v < Q; ++v) put(h.f[v + Q * id]);
                std::fprintf(fp, "\n");
            }
    const bool ok = std::ferror(fp) == 0;
    std::fclose(fp);
    return ok ? 0 : fail(LUMAGPU_EINVAL, std::string("write failed: ") + path);
}

int lumagpu_ibm_write_ld(lumagpu_grid *g, const char *path, int body, int t, double time, int precision) {
    if (!g || !path) return fail(LUMAGPU_EINVAL, "null argument");
    if (!g->ibm || body < 0 || body >= (int)g->ibm->bodies.size()) return fail(LUMAGPU_EINVAL, "no such IB body");
    if (precision < 1 || precision > 30) return fail(LUMAGPU_EINVAL, "output precision out of range");
    const Body &b = g->ibm->bodies[body];
    const int n = (int)b.m.size(), D = g->D;
    std::vector<double> rec((size_t)n * 16);
    if (int r = lumagpu_ibm_get_markers(g, body, rec.data())) return r;
    FILE *fp = std::fopen(path, "a");
    if (!fp) return fail(LUMAGPU_EINVAL, std::string("cannot open ") + path);
    if (t == 0) std::fprintf(fp, "Timestep\tTime (s)\tDrag (N)\tLift (N)\n");
    const double dm = g->d.dm, dh = g->d.dh, dt = g->d.dt;
    const double scale = (D == 2) ? dm * dh / (dt * dt) * 1.0 / dh : dm * dh / (dt * dt);
    std::fprintf(fp, "%d\t%.*g", t, precision, time);
    for (int m = 0; m < n; ++m) {
        const double *r = &rec[(size_t)m * 16];
        const double eps = r[13], ds = r[14];
        const double vd = (D == 3) ? ds : 1.0;
        for (int d = 0; d < D; ++d) std::fprintf(fp, "\t%.*g", precision, r[6 + d] * eps * vd * ds * scale);
    }
    std::fprintf(fp, "\n");
    const bool ok = std::ferror(fp) == 0;
    std::fclose(fp);
    return ok ? 0 : fail(LUMAGPU_EINVAL, std::string("write failed: ") + path);
}

// ObjectManager::io_writeTipPositions (ObjectManager_ops_io.cpp:1220-1245) for one flexible
// body: appends "t<TAB>time<TAB>x<TAB>y<TAB>z" of its last marker in the stream's default
// (%g-style) format at the given precision, after the header line when t == 0.  The reference
// writes time = t * (level-0 grid's dt), whatever the body's level, so the caller passes it.
int lumagpu_ibm_write_tip(lumagpu_grid *g, const char *path, int body, int t, double time, int precision) {
    if (!g || !path) return fail(LUMAGPU_EINVAL, "null argument");
    if (!g->ibm || body < 0 || body >= (int)g->ibm->bodies.size()) return fail(LUMAGPU_EINVAL, "no such IB body");
    if (precision < 1 || precision > 30) return fail(LUMAGPU_EINVAL, "output precision out of range");
    const Body &b = g->ibm->bodies[body];
    if (!b.flexible) return fail(LUMAGPU_EINVAL, "tip positions are written for flexible bodies only");
    const Marker &m = b.m.back();
    FILE *fp = std::fopen(path, "a");
    if (!fp) return fail(LUMAGPU_EINVAL, std::string("cannot open ") + path);
    if (t == 0) std::fprintf(fp, "Timestep\tt\tTipX\tTipY\tTipZ\n");
    std::fprintf(fp, "%d\t%.*g\t%.*g\t%.*g\t%.*g\n", t, precision, time, precision, m.pos[0], precision,
                 m.pos[1], precision, m.pos[2]);
    const bool ok = std::ferror(fp) == 0;
    std::fclose(fp);
    return ok ? 0 : fail(LUMAGPU_EINVAL, std::string("write failed: ") + path);
}

// ObjectManager::io_restart (ObjectManager_ops_io.cpp:136-290) over the grids of a hierarchy,
// bodies in grid order.  The layout is LUMA's: the body count on a line, then per body
// "\t/\t" n "\t/\t" and the marker positions (and position0 for flexible bodies), written
// at full precision.
int lumagpu_ibm_write_restart(lumagpu_grid *const *grids, int ngrids, const char *path) {
    if (!grids || ngrids <= 0 || !path) return fail(LUMAGPU_EINVAL, "null argument");
    size_t nb = 0;
    for (int i = 0; i < ngrids; ++i) {
        if (!grids[i]) return fail(LUMAGPU_EINVAL, "null grid handle");
        if (grids[i]->ibm) nb += grids[i]->ibm->bodies.size();
    }
    FILE *fp = std::fopen(path, "w");
    if (!fp) return fail(LUMAGPU_EINVAL, std::string("cannot open ") + path);
    std::fprintf(fp, "%zu\n", nb);
    for (int i = 0; i < ngrids; ++i) {
        if (!grids[i]->ibm) continue;
        for (const Body &b : grids[i]->ibm->bodies) {
            std::fprintf(fp, "\t/\t%zu\t/\t", b.m.size());
            for (const Marker &m : b.m) {
                for (int a = 0; a < 3; ++a) std::fprintf(fp, "%.17g\t", m.pos[a]);
                if (b.flexible)
                    for (int a = 0; a < 3; ++a) std::fprintf(fp, "%.17g\t", m.pos0[a]);
            }
        }
    }
    const bool ok = std::ferror(fp) == 0;
    std::fclose(fp);
    return ok ? 0 : fail(LUMAGPU_EINVAL, std::string("write failed: ") + path);
}

// The read side of the same routine.  The file is split at the '/' separators: the body count,
// then per body its marker count and its marker data (LUMA's own reader takes the marker data
// from the count's segment and so reads nothing; here each segment is read whole).  Counts
// must match the bodies already built; the positions are applied only when the whole file is
// accepted, and the bodies then need lumagpu_ibm_initialise (as main_lbm.cpp:307-318 does).
// Flexible bodies are refused: their FEM state is not part of the file.
int lumagpu_ibm_read_restart(lumagpu_grid *const *grids, int ngrids, const char *path) {
    if (!grids || ngrids <= 0 || !path) return fail(LUMAGPU_EINVAL, "null argument");
    std::vector<Body *> bodies;
    for (int i = 0; i < ngrids; ++i) {
        if (!grids[i]) return fail(LUMAGPU_EINVAL, "null grid handle");
        if (grids[i]->ibm)
            for (Body &b : grids[i]->ibm->bodies) bodies.push_back(&b);
    }
    std::ifstream in(path);
    if (!in) return fail(LUMAGPU_EINVAL, "Error opening IBM restart file. Exiting.");
    std::vector<std::string> seg;
    std::string part;
    while (std::getline(in, part, '/')) seg.push_back(part);
    size_t nb = 0;
    {
        std::istringstream iss(seg.empty() ? std::string() : seg[0]);
        if (!(iss >> nb)) return fail(LUMAGPU_EINVAL, "malformed IBM restart file");
    }
    if (nb != bodies.size())
        return fail(LUMAGPU_EINVAL, "Number of IBM bodies does not match the number specified in the restart file. Exiting.");
    if (seg.size() < 1 + 2 * nb) return fail(LUMAGPU_EINVAL, "IBM restart file ends early");
    std::vector<std::vector<double>> data(nb);
    for (size_t b = 0; b < nb; ++b) {
        if (bodies[b]->flexible)
            return fail(LUMAGPU_EUNSUPPORTED, "body " + std::to_string(b) +
                                                  " is flexible: its FEM state is not part of the IBM restart file");
        std::istringstream cnt(seg[1 + 2 * b]);
        size_t nm = 0;
        if (!(cnt >> nm) || nm != bodies[b]->m.size())
            return fail(LUMAGPU_EINVAL, "Number of IBM markers does not match the number specified for body " +
                                            std::to_string(b) + " in the restart file. Exiting.");
        std::istringstream vals(seg[2 + 2 * b]);
        data[b].resize(3 * nm);
        for (double &v : data[b])
            if (!(vals >> v)) return fail(LUMAGPU_EINVAL, "IBM restart file: too few marker values for body " +
                                                              std::to_string(b));
    }
    for (size_t b = 0; b < nb; ++b)
        for (size_t m = 0; m < bodies[b]->m.size(); ++m)
            for (int a = 0; a < 3; ++a) bodies[b]->m[m].pos[a] = data[b][3 * m + a];
    for (int i = 0; i < ngrids; ++i)
        if (grids[i]->ibm && !grids[i]->ibm->bodies.empty()) grids[i]->ibm->initialised = 0;
    return 0;
}

// GridObj::LBM_initVelocity with L_INIT_VELOCITY_FROM_FILE (GridObj_init_grids.cpp:43-78): the
// first N*M*K rows of an x y z ux uy uz file (fewer is an error), each placed at the voxel that
// encloses its position (rows off the grid skipped, a later row wins a voxel) as
// u = ud2ulbm(u_file) = u_file dt / dh, into u[d + D*id] (LUMA layout).  Sites no row reaches
// keep what u held.  Read with the same stream handling as the inlet profile.
int lumagpu_read_initial_velocity(const lumagpu_grid *g, const char *path, double *u) {
    if (!g || !path || !u) return fail(LUMAGPU_EINVAL, "null argument");
    if (!(g->d.dh > 0.0) || !(g->d.dt > 0.0)) return fail(LUMAGPU_EINVAL, "desc.dh and desc.dt must be positive");
    std::ifstream datafile(path, std::ios::in);
    if (!datafile.is_open())
        return fail(LUMAGPU_EINVAL, std::string("Cannot open velocity data file named ") + path + ". Exiting.");
    std::vector<double> col[6];
    std::string line_in;
    std::istringstream iss;
    double tmp = 0.0;
    while (!datafile.eof()) {
        std::getline(datafile, line_in, '\n');
        iss.str(line_in);
        iss.seekg(0);
        for (auto &c : col) {
            iss >> tmp;
            c.push_back(tmp);
        }
    }
    const int N = g->d.N, M = g->d.M, K = g->d.K, D = g->D;
    const size_t n = (size_t)N * M * K;
    if (col[0].size() < n)
        return fail(LUMAGPU_EINVAL, "The initial velocity file has less points than the LBM grid -- change "
                                    "L_BX/L_BY/L_BZ/L_RESOLUTION or change the file initial_velocity.in. Exiting.");
    const double dh = g->d.dh, dt = g->d.dt;
    for (size_t r = 0; r < n; ++r) {
        const double x = col[0][r], y = col[1][r], z = col[2][r];
        // GridUtils::isOnThisRank, serial: inside the level-0 edges, then the enclosing voxel
        if (!(0.0 <= x && x < g->d.xmax && 0.0 <= y && y < g->d.ymax)) continue;
        if (D == 3 && !(0.0 <= z && z < g->d.zmax)) continue;
        const int i = (int)std::floor((x - g->d.origin[0]) / dh), j = (int)std::floor((y - g->d.origin[1]) / dh);
        const int k = (D == 3) ? (int)std::floor((z - g->d.origin[2]) / dh) : 0;
        if (i < 0 || i >= N || j < 0 || j >= M || k < 0 || k >= K) continue;
        const size_t id = (size_t)k + (size_t)j * K + (size_t)i * K * M;
        u[0 + D * id] = (col[3][r] * dt) / dh;
        u[1 + D * id] = (col[4][r] * dt) / dh;
        if (D == 3) u[2 + D * id] = (col[5][r] * dt) / dh;
    }
    return 0;
}

// GridObj::_LBM_initGetInletProfileFromFile (GridObj_init_grids.cpp:1233-1312) with
// GridUtils::readVelocityFromFile (GridUtils.cpp:1467-1524): columns x y z ux uy uz, the
// profile taken along y and interpolated (or extrapolated from the end pair) at this grid's
// YPos.  The same stream handling as the reference: one istringstream re-pointed at each line,
// so a line holding no number (the file's final newline) appends the last value read to every
// column.  The values go to ux_in/uy_in/uz_in as the file gives them, without the ud2ulbm
// conversion the parabolic and uniform profiles get (:1336-1355).
int lumagpu_read_inlet_profile(lumagpu_grid *g, const char *path) {
    if (!g || !path) return fail(LUMAGPU_EINVAL, "null argument");
    if (g->ypos.size() != (size_t)g->d.M)
        return fail(LUMAGPU_ESTATE, "grid positions are not set (lumagpu_set_labels / lumagpu_set_positions)");
    std::ifstream datafile(path, std::ios::in);
    if (!datafile.is_open())
        return fail(LUMAGPU_EINVAL, std::string("Cannot open velocity data file named ") + path + ". Exiting.");
    std::vector<double> xb, yb, zb, uxb, uyb, uzb;
    std::string line_in;
    std::istringstream iss;
    double tmp = 0.0;
    while (!datafile.eof()) {
        std::getline(datafile, line_in, '\n');
        iss.str(line_in);
        iss.seekg(0);
        for (std::vector<double> *col : {&xb, &yb, &zb, &uxb, &uyb, &uzb}) {
            iss >> tmp;
            col->push_back(tmp);
        }
    }
    const size_t n = yb.size();
    if (n < 2) return fail(LUMAGPU_EINVAL, "Failed to read in inlet profile data. Exiting.");
    const int M = g->d.M;
    std::vector<double> ux(M, 0.0), uy(M, 0.0), uz(M, 0.0);
    for (int j = 0; j < M; j++) {
        const double y = g->ypos[j];
        for (size_t i = 0; i < n; i++) {
            if (i == 0 && yb[i] > y) {
                ux[j] = -(uxb[i + 1] - uxb[i]) + uxb[i];
                uy[j] = -(uyb[i + 1] - uyb[i]) + uyb[i];
                uz[j] = -(uzb[i + 1] - uzb[i]) + uzb[i];
                break;
            } else if (i == n - 1 && yb[i] < y) {
                ux[j] = (uxb[i] - uxb[i - 1]) + uxb[i];
                uy[j] = (uyb[i] - uyb[i - 1]) + uyb[i];
                uz[j] = (uzb[i] - uzb[i - 1]) + uzb[i];
                break;
            } else if (i + 1 < n && yb[i] < y && yb[i + 1] > y) {
                ux[j] = ((uxb[i + 1] - uxb[i]) / (yb[i + 1] - yb[i])) * (y - yb[i]) + uxb[i];
                uy[j] = ((uyb[i + 1] - uyb[i]) / (yb[i + 1] - yb[i])) * (y - yb[i]) + uyb[i];
                uz[j] = ((uzb[i + 1] - uzb[i]) / (yb[i + 1] - yb[i])) * (y - yb[i]) + uzb[i];
                break;
            } else if (yb[i] == y) {
                ux[j] = uxb[i];
                uy[j] = uyb[i];
                uz[j] = uzb[i];
                break;
            }
        }
    }
    return lumagpu_set_inlet(g, ux.data(), uy.data(), uz.data());
}

}  // extern "C"
