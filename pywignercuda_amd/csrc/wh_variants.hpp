// Launch variants of the fused sweeps: (radix R, lanes per block).  The tuner
// (pywignercuda_amd/tune.py) times the valid variants per grid size on the GPU
// and stores its choice; the host falls back to the first valid entry.
#pragma once

namespace wh {

// tp: throughput form for batches of grids: twiddles read from the table instead of
// held in registers and the exchanges between transform passes held in the state's
// storage precision (complex64 for the complex64-state FP64-arithmetic form), so that
// three waves fit per SIMD (<= 168 VGPRs, ~36-41 KiB of LDS per block)
struct SweepVariant {
  int R, block, tp;
};

// index 0 is the default
constexpr SweepVariant kXVariants[] = {{16, 256, 0}, {8, 256, 0}, {8, 512, 0}, {4, 256, 0}, {16, 256, 1}};
constexpr SweepVariant kPVariants[] = {{16, 256, 0}, {8, 512, 0}, {8, 256, 0}, {16, 256, 1}};
constexpr SweepVariant kPCLVariants[] = {{16, 256, 0}, {8, 256, 0}, {16, 256, 1}};
constexpr int kNumX = sizeof(kXVariants) / sizeof(kXVariants[0]);
constexpr int kNumP = sizeof(kPVariants) / sizeof(kPVariants[0]);
constexpr int kNumPCL = sizeof(kPCLVariants) / sizeof(kPCLVariants[0]);

// largest grid side of the plain (x and p) throughput forms
constexpr int kTpMaxN = 1024;

// same rule as size_ok<N, R, BLOCK>() in wh_sweeps.hpp (and the throughput forms' size bound,
// which the Caldeira-Leggett one does not have; cl: a Caldeira-Leggett p variant)
constexpr bool variant_fits(int N, SweepVariant v, bool cl = false) {
  return N >= v.R * 4 && N / v.R <= v.block && v.block / (N / v.R) <= 64 && (!v.tp || cl || N <= kTpMaxN);
}
// rows per x-sweep tile / columns per p-sweep tile
constexpr int tile_lines(int N, SweepVariant v) { return v.block / (N / v.R); }

}  // namespace wh
