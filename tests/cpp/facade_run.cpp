// Drives the GPU step through the C++ facade (lumagpu_facade.hpp) the way LUMA's main loop
// drives GridObj and ObjectManager: a case written by the test (descriptor, labels, positions,
// inlet, initial rho and u, LUMA layouts), an optional geometry.config, a number of steps,
// then the post-collision f of the grid in LUMA's layout to an output file.
//   facade_run <case.bin> <geometry.config or -> <steps> <out.bin> [<restart path> or -] [split]
// With a restart path the state after the steps is also written as a restart file, read into
// a second grid of the same descriptor, and that grid's f is appended to the output.  With
// "split" the second half of the steps runs as LBM_stream_bc_macro, ObjectManager's four IB
// stages and LBM_force_collide_swap instead of LBM_multi_opt.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "lumagpu_facade.hpp"

namespace {

template <class T>
void get(std::ifstream &in, T *p, size_t n) {
    in.read(reinterpret_cast<char *>(p), sizeof(T) * n);
    if (!in) throw std::runtime_error("short case file");
}

struct Case {
    lumagpu_grid_desc d{};
    std::vector<int> lat;
    std::vector<double> x, y, z, ux, uy, uz, rho, u;
};

Case read_case(const char *path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open case file");
    Case c;
    int h[10];
    get(in, h, 10);
    c.d.dims = h[0];
    c.d.Q = h[1];
    c.d.N = h[2];
    c.d.M = h[3];
    c.d.K = h[4];
    c.d.precision = h[5];
    c.d.regularised = h[6];
    c.d.forced = h[7];
    c.d.fz_compat = h[8];
    c.d.level = h[9];
    double r[21];
    get(in, r, 21);
    c.d.omega = r[0];
    c.d.rho_in = r[1];
    c.d.rho0 = r[2];
    c.d.xmax = r[3];
    c.d.ymax = r[4];
    c.d.zmax = r[5];
    for (int i = 0; i < 6; ++i) c.d.wall_thk[i] = r[6 + i];
    c.d.dh = r[12];
    for (int i = 0; i < 3; ++i) c.d.origin[i] = r[13 + i];
    c.d.dt = r[16];
    c.d.dm = r[17];
    const size_t n = (size_t)c.d.N * c.d.M * c.d.K;
    c.lat.resize(n);
    get(in, c.lat.data(), n);
    c.x.resize(c.d.N);
    c.y.resize(c.d.M);
    c.z.resize(c.d.K);
    get(in, c.x.data(), c.x.size());
    get(in, c.y.data(), c.y.size());
    get(in, c.z.data(), c.z.size());
    c.ux.resize(c.d.M);
    c.uy.resize(c.d.M);
    c.uz.resize(c.d.M);
    get(in, c.ux.data(), c.d.M);
    get(in, c.uy.data(), c.d.M);
    get(in, c.uz.data(), c.d.M);
    c.rho.resize(n);
    c.u.resize(n * c.d.dims);
    get(in, c.rho.data(), n);
    get(in, c.u.data(), c.u.size());
    return c;
}

void fill(luma::GridObj &g, const Case &c) {
    g.LatTyp = c.lat;
    g.XPos = c.x;
    g.YPos = c.y;
    g.ZPos = c.z;
    g.ux_in = c.ux;
    g.uy_in = c.uy;
    g.uz_in = c.uz;
    g.rho = c.rho;
    g.u = c.u;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: facade_run case.bin geometry.config|- steps out.bin [restart]\n");
        return 2;
    }
    luma::set_error_handler([](const std::string &m) { throw std::runtime_error(m); });
    try {
        const Case c = read_case(argv[1]);
        const int steps = std::stoi(argv[3]);
        luma::GridObj g(c.d);
        fill(g, c);
        g.LBM_initGrid();
        luma::ObjectManager *om = luma::ObjectManager::getInstance(&g);
        if (std::strcmp(argv[2], "-") != 0) {
            const std::array<double, 3> shift = {0.0, 0.0, 0.0};
            om->io_readInGeomConfig(argv[2], shift);
            om->ibm_initialise();
        }
        const bool split = argc > 6 && std::strcmp(argv[6], "split") == 0;
        const int first = steps / 2;
        g.LBM_multi_run(first);
        for (int s = first; s < steps; ++s) {
            if (!split) {
                g.LBM_multi_opt();
                continue;
            }
            g.LBM_stream_bc_macro(0);
            om->ibm_interpolate(g.level);
            om->ibm_computeForce(g.level);
            om->ibm_spread(g.level);
            om->ibm_updateMacroscopic(g.level);
            g.LBM_force_collide_swap();
        }
        g.sync();
        std::ofstream out(argv[4], std::ios::binary);
        out.write(reinterpret_cast<const char *>(g.f.data()), sizeof(double) * g.f.size());
        if (argc > 5 && std::strcmp(argv[5], "-") != 0) {
            g.io_restart(luma::eWrite, argv[5]);
            luma::GridObj g2(c.d);
            fill(g2, c);
            g2.LBM_initGrid();
            g2.io_restart(luma::eRead, argv[5]);
            g2.sync();
            out.write(reinterpret_cast<const char *>(g2.f.data()), sizeof(double) * g2.f.size());
        }
        std::printf("facade_run: %d steps, t = %d\n", steps, g.t);
        luma::ObjectManager::destroyInstance();
    } catch (const std::exception &e) {
        std::fprintf(stderr, "facade_run failed: %s\n", e.what());
        return 1;
    }
    return 0;
}
