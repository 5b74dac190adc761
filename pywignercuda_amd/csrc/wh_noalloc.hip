// Scratch memory of wh_wigner2d_observe_noalloc: the observables of one Wigner function
// without any device allocation.  The member record, the column-sum work list and the
// column sums live in a file-scope device array (one copy per GPU, held until the process
// ends); the kernels are the column-sum and reduce kernels of wh_wigner2d_observe.
//
// Every call on a GPU shares that GPU's array: calls that overlap on one GPU (two
// streams, two host threads) race.  Nothing in the array needs to start at zero: the
// bind kernel rewrites the member record, the step index and the work list, and the
// column-sum kernel writes every column the reduce kernel reads.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>

#include "wh_kernels.hpp"

namespace wh {
hipError_t launch_colsum(int prec, const Member* members, const unsigned* work, int nwork, hipStream_t s);
hipError_t launch_reduce(const ReduceArgs& a, int nmembers, int* done, hipStream_t s);

constexpr int kNoAllocMaxX = 2048;  // X of the largest grid (column sums of 2048 columns)

struct NoAllocScratch {
  Member member;
  int step, pad[3];
  unsigned work[kNoAllocMaxX / kColsumCols];
  double colsums[size_t(kNoAllocMaxX) * kNCol];
};

static __device__ __align__(256) NoAllocScratch scratch;

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

__global__ void __launch_bounds__(256) noalloc_bind_kernel(NoAllocScratch* sc, Member m, int ntiles) {
  if (threadIdx.x == 0) {
    sc->member = m;
    sc->step = 0;
  }
  for (int i = threadIdx.x; i < ntiles; i += 256) sc->work[i] = unsigned(i);  // member 0, column tile i
}
}  // namespace

int noalloc_max_x() { return kNoAllocMaxX; }

// Device address of the current device's copy of `scratch`: looked up once per device
// (with that device current), then reused.  Returns hipErrorInvalidDevice for device
// indices beyond the table.
hipError_t noalloc_scratch(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!ptr) return hipErrorInvalidSymbol;
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

// Column sums and reduce of wh_wigner2d_observe with the member record in `scratch`.
// `m` carries the state, the record row and the host grids; its colsums pointer is set
// here.  X must not exceed kNoAllocMaxX (checked by the caller).
hipError_t launch_observe_noalloc(int prec, Member m, hipStream_t s) {
  void* base = nullptr;
  hipError_t e = noalloc_scratch(&base);
  if (e != hipSuccess) return e;
  char* b = static_cast<char*>(base);
  NoAllocScratch* sc = static_cast<NoAllocScratch*>(base);
  const Member* mem = reinterpret_cast<const Member*>(b + offsetof(NoAllocScratch, member));
  int* step = reinterpret_cast<int*>(b + offsetof(NoAllocScratch, step));
  unsigned* work = reinterpret_cast<unsigned*>(b + offsetof(NoAllocScratch, work));
  m.colsums = reinterpret_cast<double*>(b + offsetof(NoAllocScratch, colsums));
  m.obs_base = 0;
  const int ntiles = (m.X + kColsumCols - 1) / kColsumCols;
  hipLaunchKernelGGL(noalloc_bind_kernel, dim3(1), dim3(256), 0, s, sc, m, ntiles);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = launch_colsum(prec, mem, work, ntiles, s);
  if (e != hipSuccess) return e;
  ReduceArgs r{};
  r.members = mem;
  r.renorm = 0;
  r.record = 1;
  r.rho = 0;
  r.update_scale = 0;
  r.step = step;
  r.step_inc = nullptr;  // the done counter is only touched with step_inc
  r.next_step = -1;
  return launch_reduce(r, 1, nullptr, s);
}

}  // namespace wh
