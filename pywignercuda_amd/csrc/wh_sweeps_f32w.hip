// complex64 state with FP64 arithmetic: the sweeps widen on load and round once on store,
// so the state carries one FP32 rounding per sweep and no FP32 transform error.
#define WH_S float
#define WH_T double
#define WH_SUFFIX f32w
#include "wh_sweeps_inst.inc"
