// ibm_system.hpp -- host-side state of a grid's immersed bodies (markers, supports, device
// copies), shared by the IB step (ibm.hip) and the discrete adjoint (adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "fem.hpp"
#include "ibm_kernels.hpp"
#include "internal.hpp"

namespace lumagpu {

constexpr int kSmax = 27;

struct Marker {
    double pos[3] = {0, 0, 0}, pos0[3] = {0, 0, 0}, vel[3] = {0, 0, 0}, vel_km1[3] = {0, 0, 0};
    double eps = 0.0, ds = 1.0, dilation = 1.0, area = 1.0;
    int ns = 0;
    int sijk[kSmax][3];
    double sxyz[kSmax][3];
    double delta[kSmax];
    // derivatives along the tangent direction (tangent-form grids only; zero otherwise)
    double dpos[3] = {0, 0, 0}, dvel[3] = {0, 0, 0}, dvel_km1[3] = {0, 0, 0};
    double deps = 0.0, dds = 0.0;
    double ddelta[kSmax] = {};
};

struct Body {
    std::vector<Marker> m;
    int flexible = 0;
    int first = 0;                            // index of the first marker in the grid-wide marker arrays
    std::unique_ptr<FemBeam> fem;             // flexible filaments only
    std::unique_ptr<FemBeamT<Dual>> femd;     // the same beam on a tangent-form grid
};

// Device buffer of n elements, reallocated only when n changes.  A fresh buffer is zeroed.
template <class T>
int dev_resize(T **d, size_t *cap, size_t n, hipStream_t s) {
    if (*d && *cap == n) return 0;
    if (*d) (void)hipFree(*d);
    *d = nullptr;
    *cap = 0;
    if (n == 0) return 0;
    if (hipMalloc(d, n * sizeof(T)) != hipSuccess) return fail(LUMAGPU_ENOMEM, "IB buffer allocation failed");
    *cap = n;
    if (hipMemsetAsync(*d, 0, n * sizeof(T), s) != hipSuccess) return fail(LUMAGPU_EHIP, "IB buffer clear failed");
    return 0;
}
template <class T>
int dev_upload(T **d, size_t *cap, const std::vector<T> &h, hipStream_t s) {
    if (int r = dev_resize(d, cap, h.size(), s)) return r;
    if (!h.empty() && hipMemcpyAsync(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s) != hipSuccess)
        return fail(LUMAGPU_EHIP, "IB buffer upload failed");
    return 0;
}

struct IbmSystem {
    std::vector<Body> bodies;
    int nm = 0;
    int S = 9;
    int initialised = 0;
    int universal = 0;          // L_UNIVERSAL_EPSILON_CALC: one epsilon system over all bodies
    // device copies: the marker and support arrays point into one arena, filled from a pinned
    // staging buffer by one copy per upload (a moving filament re-uploads every FSI
    // sub-iteration); the arena only grows
    int *d_scell = nullptr, *d_ns = nullptr, *d_sites = nullptr, *d_sptr = nullptr, *d_sent = nullptr;
    double *d_sdelta = nullptr, *d_vel = nullptr, *d_eps = nullptr, *d_ds = nullptr;
    void *d_pack = nullptr, *h_pack = nullptr;
    size_t pack_cap = 0;        // bytes of d_pack and h_pack
    hipEvent_t pack_ev = nullptr;   // recorded after the last copy out of h_pack
    double *d_interp = nullptr, *d_force = nullptr, *d_out = nullptr;
    size_t cap[11] = {};        // element counts of d_interp (9) and d_force (10)
    double *h_force = nullptr;  // pinned download buffer of the marker forces
    size_t h_force_cap = 0;
    int *d_gone = nullptr, *h_gone = nullptr;   // cells to clear (clear_gone), device and pinned host
    size_t gone_cap = 0;
    std::vector<int> site_of;   // per cell: index of the support site, -1 elsewhere (kept all -1 between uses)
    int *d_site_of = nullptr;   // device copy for the step kernel's spread (rigid systems), -1 elsewhere
    int site_map_ok = 0;        // d_site_of matches the current sites
    int nsites = 0;
    std::vector<int> sites;     // host copy of the distinct support cells
    std::vector<int> flagged;   // cells whose force and flag the device currently holds
    int sites_dirty = 0;        // the support sites changed since the last application
    int sub_it = 0;             // sub-iterations of the last step
    double sub_res = 0.0;
    std::vector<int> sub_log;   // sub-iterations of every FSI step since the bodies were initialised
    std::vector<int> pinned;    // prescribed sub-iteration counts of the next steps
    size_t pinned_pos = 0;

    ~IbmSystem() {
        for (void *p : {d_pack, (void *)d_interp, (void *)d_force, (void *)d_out, (void *)d_gone, (void *)d_site_of})
            if (p) (void)hipFree(p);
        for (void *p : {h_pack, (void *)h_force, (void *)h_gone})
            if (p) (void)hipHostFree(p);
        if (pack_ev) (void)hipEventDestroy(pack_ev);
    }
};

// true when no support of marker I can carry a nonzero delta weight of marker J
inline bool far_apart(const Marker &mI, const Marker &mJ, int D, double dh) {
    const double lim = (1.5 * mI.dilation + 1.5 * mJ.dilation + 0.01) * dh;
    for (int a = 0; a < D; ++a)
        if (std::fabs(mJ.pos[a] - mI.pos[a]) > lim) return true;
    return false;
}

// ibm_deltaKernel (ObjectManager_ops_ibm.cpp:255-272) and its derivative in the radius
double delta_kernel(double radius, double dilation);
double delta_kernel_deriv(double radius, double dilation);
// delta(x; s) = prod_a phi((x_a - s_a)/dh) and its gradient in x
double delta_at(const double *x, const double *s, double dh, double dil, int D);
void delta_grad(const double *x, const double *s, double dh, double dil, int D, double out[3]);
// the grid stores (value, derivative) pairs: its IB device arrays hold Duals
inline bool tangent_form(const lumagpu_grid *g) {
    return g->d.precision == LUMAGPU_FP64T || g->d.precision == LUMAGPU_FP32T;
}
IbmArgs ibm_args(lumagpu_grid *g);
// interpolate, force and spread at the current supports from the grid's current populations
int apply_core(lumagpu_grid *g);
// A x = rhs, row-major A, LAPACK dgetf2 + dgetrs('T') operation order
int lu_solve(int n, std::vector<double> &a, std::vector<double> &x);
int lu_solve(int n, std::vector<Dual> &a, std::vector<Dual> &x);
// the same solve for a matrix with kl sub- and ku super-diagonals (zeros elsewhere)
int lu_solve_band(int n, int kl, int ku, std::vector<double> &a, std::vector<double> &x,
                  const char *singular = "singular beam system");
int lu_solve_band(int n, int kl, int ku, std::vector<Dual> &a, std::vector<Dual> &x);

}  // namespace lumagpu
