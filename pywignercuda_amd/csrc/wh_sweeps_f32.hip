// FP32 (complex64 state, FP32 arithmetic with FP64 phases and compensated twiddles)
// instantiation of the fused sweeps.
#define WH_S float
#define WH_T float
#define WH_SUFFIX f32
#include "wh_sweeps_inst.inc"
