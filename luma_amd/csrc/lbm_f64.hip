// lbm_f64.hip -- FP64 instantiation of the step kernels.  Built with -ffp-contract=off:
// the exact per-site arithmetic must round as LUMA's non-FMA CPU build does.
#include "launch.hpp"

namespace lumagpu {

hipError_t launch_step_f64(const LaunchSpec &ls, const StepArgs &a, hipStream_t s) {
    return launch_step_p<P64>(ls, a, s);
}
hipError_t launch_fused2_f64(int shape, const StepArgs &a, void *dst, int zc, bool store, hipStream_t s) {
    return launch_fused2_p<P64>(shape, a, dst, zc, store, s);
}

hipError_t launch_aux_f64(int op, int Q, const StepArgs &a, const double *p0, const double *p1,
                          double *out, int parity, int masked, hipStream_t s) {
    return launch_aux_p<P64>(op, Q, a, p0, p1, out, parity, masked, s);
}

hipError_t launch_ibm_interp_f64(int Q, const StepArgs &a, const IbmArgs &b, bool odd, int mode, hipStream_t s) {
    return launch_ibm_interp_p<P64>(Q, a, b, odd, mode, s);
}
hipError_t launch_sites_f64(int Q, const StepArgs &a, bool odd, int n, const int *cells, int ib_force, hipStream_t s) {
    return launch_sites_p<P64>(Q, a, odd, n, cells, ib_force, s);
}

hipError_t launch_edges_f64(int Q, const StepArgs &a, const EdgeArgs &e, bool odd, int phase, int store, hipStream_t s) {
    return launch_edges_p<P64>(Q, a, e, odd, phase, store, s);
}

hipError_t launch_zslab_f64(const StepArgs &a, int phase, hipStream_t s) { return launch_zslab_p<P64>(a, phase, s); }

}  // namespace lumagpu
