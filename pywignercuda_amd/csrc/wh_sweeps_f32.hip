// FP32-bulk (complex64 state, FP64 phases) instantiation of the fused sweeps.
#define WH_T float
#define WH_SUFFIX f32
#include "wh_sweeps_inst.inc"
