// Launch variants of the fused sweeps: (radix R, lanes per block).  The tuner
// (pywignercuda_amd/tune.py) times the valid variants per grid size on the GPU
// and stores its choice; the host falls back to the first valid entry.
#pragma once

namespace wh {

struct SweepVariant {
  int R, block;
};

// index 0 is the default
constexpr SweepVariant kXVariants[] = {{16, 256}, {8, 256}, {8, 512}, {4, 256}};
constexpr SweepVariant kPVariants[] = {{16, 256}, {8, 512}, {8, 256}};
constexpr SweepVariant kPCLVariants[] = {{16, 256}, {8, 256}};
constexpr int kNumX = sizeof(kXVariants) / sizeof(kXVariants[0]);
constexpr int kNumP = sizeof(kPVariants) / sizeof(kPVariants[0]);
constexpr int kNumPCL = sizeof(kPCLVariants) / sizeof(kPCLVariants[0]);

// same rule as size_ok<N, R, BLOCK>() in wh_sweeps.hpp
constexpr bool variant_fits(int N, SweepVariant v) {
  return N >= v.R * 4 && N / v.R <= v.block && v.block / (N / v.R) <= 64;
}
// rows per x-sweep tile / columns per p-sweep tile
constexpr int tile_lines(int N, SweepVariant v) { return v.block / (N / v.R); }

}  // namespace wh
