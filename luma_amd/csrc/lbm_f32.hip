// lbm_f32.hip -- shifted-storage instantiations (FP32 and FP16 storage, FP32 collision).
#include "launch.hpp"

namespace lumagpu {

hipError_t launch_step_f32s(const LaunchSpec &ls, const StepArgs &a, hipStream_t s) {
    return launch_step_p<P32S>(ls, a, s);
}
hipError_t launch_step_f16s(const LaunchSpec &ls, const StepArgs &a, hipStream_t s) {
    return launch_step_p<P16S>(ls, a, s);
}
hipError_t launch_fused2_f32s(int shape, const StepArgs &a, void *dst, int zc, bool store, hipStream_t s) {
    return launch_fused2_p<P32S>(shape, a, dst, zc, store, s);
}
hipError_t launch_fused2_f16s(int shape, const StepArgs &a, void *dst, int zc, bool store, hipStream_t s) {
    return launch_fused2_p<P16S>(shape, a, dst, zc, store, s);
}
hipError_t launch_aux_f32s(int op, int Q, const StepArgs &a, const double *p0, const double *p1,
                           double *out, int parity, int masked, hipStream_t s) {
    return launch_aux_p<P32S>(op, Q, a, p0, p1, out, parity, masked, s);
}
hipError_t launch_aux_f16s(int op, int Q, const StepArgs &a, const double *p0, const double *p1,
                           double *out, int parity, int masked, hipStream_t s) {
    return launch_aux_p<P16S>(op, Q, a, p0, p1, out, parity, masked, s);
}

hipError_t launch_ibm_interp_f32s(int Q, const StepArgs &a, const IbmArgs &b, bool odd, int mode, hipStream_t s) {
    return launch_ibm_interp_p<P32S>(Q, a, b, odd, mode, s);
}
hipError_t launch_ibm_interp_f16s(int Q, const StepArgs &a, const IbmArgs &b, bool odd, int mode, hipStream_t s) {
    return launch_ibm_interp_p<P16S>(Q, a, b, odd, mode, s);
}
hipError_t launch_sites_f32s(int Q, const StepArgs &a, bool odd, int n, const int *cells, int ib_force, hipStream_t s) {
    return launch_sites_p<P32S>(Q, a, odd, n, cells, ib_force, s);
}
hipError_t launch_sites_f16s(int Q, const StepArgs &a, bool odd, int n, const int *cells, int ib_force, hipStream_t s) {
    return launch_sites_p<P16S>(Q, a, odd, n, cells, ib_force, s);
}
hipError_t launch_edges_f32s(int Q, const StepArgs &a, const EdgeArgs &e, bool odd, int phase, int store, hipStream_t s) {
    return launch_edges_p<P32S>(Q, a, e, odd, phase, store, s);
}
hipError_t launch_edges_f16s(int Q, const StepArgs &a, const EdgeArgs &e, bool odd, int phase, int store, hipStream_t s) {
    return launch_edges_p<P16S>(Q, a, e, odd, phase, store, s);
}

hipError_t launch_zslab_f32s(const StepArgs &a, int phase, hipStream_t s) { return launch_zslab_p<P32S>(a, phase, s); }
hipError_t launch_zslab_f16s(const StepArgs &a, int phase, hipStream_t s) { return launch_zslab_p<P16S>(a, phase, s); }

}  // namespace lumagpu
