// lbm_dual.hip -- tangent-form instantiation of the step kernels: the FP64 exact arithmetic
// on Duals (value, derivative).  Built with -ffp-contract=off, as the FP64 form: the values
// then round exactly as the FP64 kernels do.
#include "launch.hpp"

namespace lumagpu {

hipError_t launch_step_dual(const LaunchSpec &ls, const StepArgs &a, hipStream_t s) {
    return launch_step_p<PDual>(ls, a, s);
}

hipError_t launch_aux_dual(int op, int Q, const StepArgs &a, const double *p0, const double *p1,
                           double *out, int parity, int masked, hipStream_t s) {
    return launch_aux_p<PDual>(op, Q, a, p0, p1, out, parity, masked, s);
}

hipError_t launch_ibm_interp_dual(int Q, const StepArgs &a, const IbmArgs &b, bool odd, int mode, hipStream_t s) {
    return launch_ibm_interp_p<PDual>(Q, a, b, odd, mode, s);
}
hipError_t launch_sites_dual(int Q, const StepArgs &a, bool odd, int n, const int *cells, int ib_force, hipStream_t s) {
    return launch_sites_p<PDual>(Q, a, odd, n, cells, ib_force, s);
}

hipError_t launch_edges_dual(int Q, const StepArgs &a, const EdgeArgs &e, bool odd, int phase, int store, hipStream_t s) {
    return launch_edges_p<PDual>(Q, a, e, odd, phase, store, s);
}

}  // namespace lumagpu
