// lbm_dual32.hip -- shifted tangent-form instantiation: the FP32-shifted storage and collision
// on float Duals (value, derivative), FP64 Duals for the macro arrays and force field.  Built
// with the shifted forms' flags.
#include "launch.hpp"

namespace lumagpu {

hipError_t launch_step_dual32(const LaunchSpec &ls, const StepArgs &a, hipStream_t s) {
    return launch_step_p<P32T>(ls, a, s);
}

hipError_t launch_aux_dual32(int op, int Q, const StepArgs &a, const double *p0, const double *p1,
                             double *out, int parity, int masked, hipStream_t s) {
    return launch_aux_p<P32T>(op, Q, a, p0, p1, out, parity, masked, s);
}

hipError_t launch_ibm_interp_dual32(int Q, const StepArgs &a, const IbmArgs &b, bool odd, int mode, hipStream_t s) {
    return launch_ibm_interp_p<P32T>(Q, a, b, odd, mode, s);
}
hipError_t launch_sites_dual32(int Q, const StepArgs &a, bool odd, int n, const int *cells, int ib_force, hipStream_t s) {
    return launch_sites_p<P32T>(Q, a, odd, n, cells, ib_force, s);
}

hipError_t launch_edges_dual32(int Q, const StepArgs &a, const EdgeArgs &e, bool odd, int phase, int store, hipStream_t s) {
    return launch_edges_p<P32T>(Q, a, e, odd, phase, store, s);
}

}  // namespace lumagpu
