// FP64 (complex128) instantiation of the fused sweeps.
#define WH_S double
#define WH_T double
#define WH_SUFFIX f64
#include "wh_sweeps_inst.inc"

namespace wh {
hipError_t launch_gpshift(const Member* members, const unsigned* work, int nwork, hipStream_t s) {
  hipLaunchKernelGGL((gpshift_kernel<16, 256>), dim3(nwork), dim3(256), 0, s, members, work);
  return hipGetLastError();
}
}  // namespace wh
