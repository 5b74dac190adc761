// Reductions, small helpers and the plain axis transforms (cuFFT replacement).
//
//   colsum  column sums of a state (observables of a state not produced by a p sweep)
//   reduce  observables + complex norm of each member from its column sums,
//           renormalisation factor for the next load, Gross-Pitaevskii density
//           (GPU_Wigner2D_FFT.py:1778-1780, 1857-1878, 1912-1913; 435-455)
//   colpart_reduce  density after the kinetic step (GP.py:1933-1937)
//   fft_rows / fft_cols  in-place batched transforms for Plan_Z2Z_2D_Axis1 / Axis0
//           (cufft_wrapper.py:183-250, 396-403)
#include <hip/hip_runtime.h>

#include "wh_fft.hpp"
#include "wh_kernels.hpp"

namespace wh {

// ------------------------------------------------------------------ column sums
// One tile = kColsumCols consecutive columns of one member; 256 lanes = 64 row groups of
// kColsumCols lanes (each load instruction of a wave covers 16 rows x 4 columns), the 64
// group sums combined by a fixed tree in LDS.
template <typename T>
__device__ __forceinline__ void colsum_tile(const Member& m, int tile, double (*red)[kColsumCols][kNCol]) {
  constexpr int CC = kColsumCols, G = 256 / CC;
  const cplx<T>* W = reinterpret_cast<const cplx<T>*>(m.W);
  const int c = threadIdx.x % CC, g = threadIdx.x / CC;
  const int col = tile * CC + c;
  double acc[kNCol] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (col < m.X) {
#pragma unroll 4
    for (int i = g; i < m.P; i += G) {
      const cplx<T> v = W[size_t(i) * m.X + col];
      const double p = m.p_host[i], wr = double(v.x);
      acc[0] += wr;
      acc[1] += double(v.y);
      acc[2] += p * wr;
      acc[3] += p * p * wr;
      acc[4] += wr < 0.0 ? wr : 0.0;
    }
  }
#pragma unroll
  for (int k = 0; k < kNCol; ++k) red[g][c][k] = acc[k];
  __syncthreads();
  for (int h = G / 2; h > 0; h >>= 1) {
    if (g < h) {
#pragma unroll
      for (int k = 0; k < kNCol; ++k) red[g][c][k] += red[g + h][c][k];
    }
    __syncthreads();
  }
  if (g == 0 && col < m.X) {
#pragma unroll
    for (int k = 0; k < kNCol; ++k) m.colsums[size_t(col) * kNCol + k] = red[0][c][k];
  }
}

template <typename T>
__global__ void __launch_bounds__(256) colsum_kernel(const Member* members, const unsigned* work) {
  __shared__ double red[256 / kColsumCols][kColsumCols][kNCol];
  const unsigned w = work[blockIdx.x];
  colsum_tile<T>(members[w >> 16], int(w & 0xffffu), red);
}

// ------------------------------------------------------------------ reduce (one block per member)
// One member's record row, pending factor and (Gross-Pitaevskii) density from its column
// sums, by a block of 256 threads.
__device__ __forceinline__ void reduce_member(const ReduceArgs& a, const Member& m, int step, double (*red)[256],
                                              double* scale_sh) {
  constexpr int K = 12;
  const int tid = threadIdx.x;
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.0;
  for (int j = tid; j < m.X; j += 256) {
    const double* cs = m.colsums + size_t(j) * kNCol;
    const double s0 = cs[0], s0i = cs[1], s1 = cs[2], s2 = cs[3], sn = cs[4];
    const double x = m.x_host[j], dv = m.dv_host[j], v0 = m.v0_host[j];
    acc[0] += s0;
    acc[1] += s0i;
    acc[2] += x * s0;
    acc[3] += x * x * s0;
    acc[4] += s1;
    acc[5] += s2;
    acc[6] += x * s1;
    acc[7] += dv * s0;
    acc[8] += dv * s1;
    acc[9] += x * dv * s0;
    acc[10] += s2 * m.inv_2m + v0 * s0;
    acc[11] += sn;
  }
#pragma unroll
  for (int k = 0; k < K; ++k) red[k][tid] = acc[k];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (tid < h) {
#pragma unroll
      for (int k = 0; k < K; ++k) red[k][tid] += red[k][tid + h];
    }
    __syncthreads();
  }
  if (tid == 0) {
    // W_normalised = W / n, n = dXdP sum W (complex); observables are Re(dXdP sum W A / n)
    // = dXdP Re(sum W A) nr / |n|^2 up to Im(sum W A) ni / |n|^2 (below 1e-30 here)
    const double nr = m.dxdp * red[0][0], ni = m.dxdp * red[1][0];
    double f = 1.0, sr = 1.0, si = 0.0;
    if (a.renorm) {
      const double m2 = nr * nr + ni * ni;
      f = nr / m2;
      sr = nr / m2;
      si = -ni / m2;
    }
    scale_sh[0] = sr;
    scale_sh[1] = si;
    if (a.update_scale) {
      m.scale[0] = sr;
      m.scale[1] = si;
    }
    if (a.record) {
      double* o = m.obs + size_t(step - m.obs_base) * kNObs;
      const double d = m.dxdp * f;
      o[kObsX] = d * red[2][0];
      o[kObsX2] = d * red[3][0];
      o[kObsP] = d * red[4][0];
      o[kObsP2] = d * red[5][0];
      o[kObsXP] = d * red[6][0];
      o[kObsDV] = d * red[7][0];
      o[kObsPDV] = d * red[8][0];
      o[kObsXDV] = d * red[9][0];
      o[kObsH] = d * red[10][0];
      o[kObsNeg] = d * red[11][0];
      o[kObsNorm] = d * red[0][0];
      o[kObsRawNormRe] = nr;
      o[kObsRawNormIm] = ni;
    }
  }
  if (a.rho) {
    __syncthreads();
    const double sr = scale_sh[0], si = scale_sh[1];
    for (int j = tid; j < m.X; j += 256) {
      const double* cs = m.colsums + size_t(j) * kNCol;
      m.rho[j] = (cs[0] * sr - cs[1] * si) * m.rho_factor;
    }
  }
}

// ------------------------------------------------------------------ reduce (one block per member)
__global__ void __launch_bounds__(256) reduce_kernel(ReduceArgs a, int* done, int nblocks) {
  __shared__ double red[12][256];
  __shared__ double scale_sh[2];
  const int step = a.use_step_arg ? a.step_arg : *a.step;
  reduce_member(a, a.members[blockIdx.x], step, red, scale_sh);
  if (a.step_inc && threadIdx.x == 0) {
    // the last block to finish advances the step counter: every block has read it by then
    __threadfence();
    const int prev = atomicAdd(done, 1);
    if (prev == nblocks - 1) {
      *done = 0;
      *a.step_inc = a.next_step >= 0 ? a.next_step : step + 1;
      __threadfence();
    }
  }
}

// ------------------------------------------------------------------ density after K
__global__ void __launch_bounds__(256) colpart_reduce_kernel(const Member* members, int xradix, int xblock,
                                                             const int* step) {
  __shared__ double red[256];
  const Member& m = members[blockIdx.x];
  const int nblk = m.P / (xblock / (m.X / xradix));  // x-sweep tiles of this member
  double ps = 0.0;
  for (int x = threadIdx.x; x < m.X; x += 256) {
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += m.colpart[size_t(b) * m.X + x];
    s *= m.rho_factor;
    m.rho[x] = s;
    ps += s * s;
  }
  red[threadIdx.x] = ps;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0 && m.obs) m.obs[size_t(*step - m.obs_base) * kNObs + kObsProbSq] = m.dx_host * red[0];
}

// ------------------------------------------------------------------ scale by the pending factor
template <typename T>
__global__ void __launch_bounds__(256) scale_kernel(const Member* members, const unsigned* work) {
  const unsigned w = work[blockIdx.x];
  const Member& m = members[w >> 16];
  const size_t chunk = size_t(w & 0xffffu) * kScaleChunk;
  const size_t n = size_t(m.X) * size_t(m.P);
  cplx<T>* W = reinterpret_cast<cplx<T>*>(m.W);
  // all of this thread's loads in flight before the factor arrives and the stores start
  static_assert(kScaleChunk % 256 == 0, "whole elements per thread");
  constexpr int kPer = kScaleChunk / 256;
  cplx<T> v[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const size_t i = chunk + threadIdx.x + size_t(k) * 256;
    if (i < n) v[k] = W[i];
  }
  const double sr = m.scale[0], si = m.scale[1];
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const size_t i = chunk + threadIdx.x + size_t(k) * 256;
    if (i < n) {
      const double x = double(v[k].x), y = double(v[k].y);
      W[i] = {T(x * sr - y * si), T(x * si + y * sr)};
    }
  }
}

__global__ void set_int_kernel(int* p, int v) { *p = v; }
__global__ void bind_kernel(BindArgs a) {
  const int i = threadIdx.x;
  if (i < a.count) {
    Member& m = a.members[a.first + i];
    m.W = a.W[i];
    m.obs = a.obs[i];
    m.obs_base = a.obs_base;
  }
}
__global__ void set_scale_kernel(const Member* members, double re, double im) {
  const Member& m = members[blockIdx.x];
  if (threadIdx.x == 0) {
    m.scale[0] = re;
    m.scale[1] = im;
  }
}

// ------------------------------------------------------------------ plain axis transforms
template <typename T, int N, int R, int ROWS, int DIR>
__global__ void __launch_bounds__(ROWS*(N / R))
    fft_rows_kernel(void* data, int rows, int normalize, Twiddles<T> tw) {
  constexpr int TT = N / R;
  constexpr int LSZ = Padded<N, R>::size;
  __shared__ cplx<T> lds[ROWS * LSZ];
  const int tid = threadIdx.x;
  const int team = tid / TT, t = tid % TT;
  const int row = blockIdx.x * ROWS + team;
  const bool active = row < rows;
  cplx<T>* W = reinterpret_cast<cplx<T>*>(data) + size_t(active ? row : 0) * N;
  LdsRow<T> view{lds + team * LSZ};
  cplx<T> v[R];
#pragma unroll
  for (int j = 0; j < R; ++j) v[j] = active ? W[t + j * TT] : cplx<T>{T(0), T(0)};
  team_fft<T, N, R, DIR>(v, t, view, tw);
  const T f = normalize ? T(1.0 / N) : T(1);
  if (active) {
#pragma unroll
    for (int j = 0; j < R; ++j) W[t + j * TT] = cscale(v[j], f);
  }
}

template <typename T, int N, int R, int C, int DIR>
__global__ void __launch_bounds__(C*(N / R))
    fft_cols_kernel(void* data, int X, int normalize, Twiddles<T> tw) {
  constexpr int TT = N / R;
  constexpr int LSZ = Padded<N, R>::size;
  __shared__ cplx<T> lds[C * LSZ];
  const int tid = threadIdx.x;
  const int c = tid % C, t = tid / C;
  const int col = blockIdx.x * C + c;
  cplx<T>* W = reinterpret_cast<cplx<T>*>(data);
  LdsStrided<T, C> view{lds + c};
  const unsigned boff = (unsigned(t) * unsigned(X) + unsigned(col)) * unsigned(sizeof(cplx<T>));
  cplx<T> v[R];
#pragma unroll
  for (int j = 0; j < R; ++j) v[j] = *reinterpret_cast<cplx<T>*>(reinterpret_cast<char*>(W + size_t(j * TT) * X) + boff);
  team_fft<T, N, R, DIR>(v, t, view, tw);
  const T f = normalize ? T(1.0 / N) : T(1);
#pragma unroll
  for (int j = 0; j < R; ++j)
    *reinterpret_cast<cplx<T>*>(reinterpret_cast<char*>(W + size_t(j * TT) * X) + boff) = cscale(v[j], f);
}

// ------------------------------------------------------------------ launchers
namespace {
constexpr int kR = 16;
template <int N>
constexpr int fft_rows_per_block() {
  return (256 / (N / kR)) < 1 ? 1 : 256 / (N / kR);
}
template <typename T, int N>
constexpr int fft_cols_per_block() {
  constexpr int c = 256 / (N / kR);
  return c < 1 ? 1 : (c > 64 ? 64 : c);
}
template <typename T, int N, int DIR>
hipError_t fft_rows_n(void* d, int rows, int norm, Twiddles<T> tw, hipStream_t s) {
  constexpr int ROWS = fft_rows_per_block<N>();
  hipLaunchKernelGGL((fft_rows_kernel<T, N, kR, ROWS, DIR>), dim3((rows + ROWS - 1) / ROWS),
                     dim3(ROWS * (N / kR)), 0, s, d, rows, norm, tw);
  return hipGetLastError();
}
template <typename T, int N, int DIR>
hipError_t fft_cols_n(void* d, int X, int norm, Twiddles<T> tw, hipStream_t s) {
  constexpr int C = fft_cols_per_block<T, N>();
  hipLaunchKernelGGL((fft_cols_kernel<T, N, kR, C, DIR>), dim3(X / C), dim3(C * (N / kR)), 0, s, d, X,
                     norm, tw);
  return hipGetLastError();
}
#define WH_FFT_SWITCH(N, CALL)         \
  switch (N) {                         \
    case 64: return CALL(64);          \
    case 128: return CALL(128);        \
    case 256: return CALL(256);        \
    case 512: return CALL(512);        \
    case 1024: return CALL(1024);      \
    case 2048: return CALL(2048);      \
    default: return hipErrorInvalidValue; \
  }
template <typename T, int DIR>
hipError_t rows_t(int N, void* d, int rows, int norm, Twiddles<T> tw, hipStream_t s) {
#define CALL(n) fft_rows_n<T, n, DIR>(d, rows, norm, tw, s)
  WH_FFT_SWITCH(N, CALL)
#undef CALL
}
template <typename T, int DIR>
hipError_t cols_t(int N, void* d, int X, int norm, Twiddles<T> tw, hipStream_t s) {
#define CALL(n) fft_cols_n<T, n, DIR>(d, X, norm, tw, s)
  WH_FFT_SWITCH(N, CALL)
#undef CALL
}
}  // namespace

int fft_cols_batch_multiple(int prec, int N) {
  switch (N) {
    case 64: return prec == 0 ? fft_cols_per_block<double, 64>() : fft_cols_per_block<float, 64>();
    case 128: return prec == 0 ? fft_cols_per_block<double, 128>() : fft_cols_per_block<float, 128>();
    case 256: return prec == 0 ? fft_cols_per_block<double, 256>() : fft_cols_per_block<float, 256>();
    case 512: return prec == 0 ? fft_cols_per_block<double, 512>() : fft_cols_per_block<float, 512>();
    case 1024: return prec == 0 ? fft_cols_per_block<double, 1024>() : fft_cols_per_block<float, 1024>();
    case 2048: return prec == 0 ? fft_cols_per_block<double, 2048>() : fft_cols_per_block<float, 2048>();
    default: return 0;
  }
}

hipError_t launch_fft_axis(int prec, int columns, int N, int dir, void* d, int batch, int norm,
                           const void* tw_hi, const void* tw_lo, hipStream_t s) {
  if (prec == 0) {
    Twiddles<double> tw{static_cast<const cplx<double>*>(tw_hi), nullptr};
    if (columns) return dir < 0 ? cols_t<double, -1>(N, d, batch, norm, tw, s) : cols_t<double, 1>(N, d, batch, norm, tw, s);
    return dir < 0 ? rows_t<double, -1>(N, d, batch, norm, tw, s) : rows_t<double, 1>(N, d, batch, norm, tw, s);
  }
  Twiddles<float> tw{static_cast<const cplx<float>*>(tw_hi), static_cast<const cplx<float>*>(tw_lo)};
  if (columns) return dir < 0 ? cols_t<float, -1>(N, d, batch, norm, tw, s) : cols_t<float, 1>(N, d, batch, norm, tw, s);
  return dir < 0 ? rows_t<float, -1>(N, d, batch, norm, tw, s) : rows_t<float, 1>(N, d, batch, norm, tw, s);
}

hipError_t launch_colsum(int prec, const Member* members, const unsigned* work, int nwork, hipStream_t s) {
  if (prec == 0)
    hipLaunchKernelGGL(colsum_kernel<double>, dim3(nwork), dim3(256), 0, s, members, work);
  else
    hipLaunchKernelGGL(colsum_kernel<float>, dim3(nwork), dim3(256), 0, s, members, work);
  return hipGetLastError();
}
hipError_t launch_reduce(const ReduceArgs& a, int nmembers, int* done, hipStream_t s) {
  hipLaunchKernelGGL(reduce_kernel, dim3(nmembers), dim3(256), 0, s, a, done, nmembers);
  return hipGetLastError();
}
hipError_t launch_colpart_reduce(const Member* members, int nmembers, int xradix, int xblock, const int* step,
                                 hipStream_t s) {
  hipLaunchKernelGGL(colpart_reduce_kernel, dim3(nmembers), dim3(256), 0, s, members, xradix, xblock, step);
  return hipGetLastError();
}
hipError_t launch_scale(int prec, const Member* members, const unsigned* work, int nwork, hipStream_t s) {
  if (prec == 0)
    hipLaunchKernelGGL(scale_kernel<double>, dim3(nwork), dim3(256), 0, s, members, work);
  else
    hipLaunchKernelGGL(scale_kernel<float>, dim3(nwork), dim3(256), 0, s, members, work);
  return hipGetLastError();
}
hipError_t launch_set_int(int* p, int v, hipStream_t s) {
  hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, s, p, v);
  return hipGetLastError();
}
hipError_t launch_bind(const BindArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(bind_kernel, dim3(1), dim3(64), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_set_scale(const Member* members, int nmembers, double re, double im, hipStream_t s) {
  hipLaunchKernelGGL(set_scale_kernel, dim3(nmembers), dim3(64), 0, s, members, re, im);
  return hipGetLastError();
}

}  // namespace wh
