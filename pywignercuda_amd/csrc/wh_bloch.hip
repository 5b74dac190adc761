// Imaginary-time (Bloch) Gross-Pitaevskii ground-state solver
// (GPU_Wigner2D_GPitaevskii_Bloch, GPU_Wigner2D_GPitaevskii.py:2064-2543), complex128.
//
// One Bloch step (Run, GP.py:2183-2245), every factor real:
//   [rho = dP sum_p W, G = B- + B+]  p sweep: F_p, exp(-dt/4 [V(x - h th/2) + V(x + h th/2) + a G]), F_p^-1
//   x sweep: F_x, exp(-dt/2 [K(p + h l/2) + K(p - h l/2)]), F_x^-1
//   [rho, G]  p sweep again, with the column sums of the result
//   finish: norm, observables, energies, purity, the acceptance test and the next dt (device)
//   commit: kept -> W = W/n, W_step = W/n; refused -> W = W_step
// (Kernel_Bloch of expPotential :88-110, of expPLambdaKinetic :160-175, of
//  expPotential_GrossPitaevskii :498-517; MakeGrossPitaevskiiTerms :1717-1735.)
// Run_ExitedState1 (:2343-2535) also removes the ground state, W -= c W_0 with
// c = 2 pi hbar dXdP sum W W_0; the column sums of W_0 are formed once, so the sums of the
// projected state follow from those of W without another pass over the grid.
//
// dt lives in device memory (halved by the finish kernel on a refused step), so the whole
// loop runs without host synchronisation.  Not a hot path (the reference runs a few hundred
// steps at 512^2): the sweeps use plain block-synchronised teams.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "wh_bloch.hpp"
#include "wh_fft.hpp"

namespace wh {
namespace {

constexpr int kR = 16;
constexpr int kBlock = 256;

__device__ __forceinline__ int fidx_b(int k, int n) { return ((k + (n >> 1)) & (n - 1)) - (n >> 1); }

__device__ __forceinline__ double hpoly(const double* c, int deg, double x) {
  double r = c[deg];
  for (int n = deg - 1; n >= 0; --n) r = fma(r, x, c[n]);
  return r;
}

// ------------------------------------------------------------------ p sweep (column tiles)
// C = kBlock / (N / kR) columns per block; lane (c = tid % C, t = tid / C) holds the strided
// elements t + j N/kR of column c (the layout of the plain column transform).
template <int N>
__global__ void __launch_bounds__(kBlock) bloch_psweep_kernel(BlochArgs a, int colsums) {
  constexpr int TT = N / kR, C = kBlock / TT, LSZ = Padded<N, kR>::size;
  __shared__ cplx<double> lds[C * LSZ];
  const int tid = threadIdx.x, c = tid % C, t = tid / C;
  const int col = blockIdx.x * C + c;
  const int X = a.X;
  cplx<double>* W = reinterpret_cast<cplx<double>*>(a.W);
  const Twiddles<double> tw{reinterpret_cast<const cplx<double>*>(a.twP), nullptr};
  LdsStrided<double, C> view{lds + c};
  cplx<double> v[kR];
#pragma unroll
  for (int j = 0; j < kR; ++j) v[j] = W[size_t(t + j * TT) * X + col];
  team_fft<double, N, kR, -1>(v, t, view, tw);
  const double dt = a.ctrl[kBDt];
  const int step = int(a.ctrl[kBIter]);
  const double x = a.kdx * double(fidx_b(col, X));
  const double* cf = a.pot + size_t(step < a.pot_rows ? step : a.pot_rows - 1) * (kMaxDegB + 1);
#pragma unroll
  for (int j = 0; j < kR; ++j) {
    const int i = t + j * TT;
    double e;
    if (a.vsum) {
      e = a.vsum[size_t(i) * X + col];
    } else {
      const double u = 0.5 * a.khbar * a.kdtheta * double(fidx_b(i, N));
      e = hpoly(cf, a.pot_deg, x - u) + hpoly(cf, a.pot_deg, x + u);
    }
    if (a.gp) e = fma(a.ka, a.G[size_t(i) * X + col], e);
    v[j] = cscale(v[j], exp(-0.25 * dt * e) * (1.0 / N));
  }
  team_fft<double, N, kR, +1>(v, t, view, tw);
#pragma unroll
  for (int j = 0; j < kR; ++j) W[size_t(t + j * TT) * X + col] = v[j];
  if (!colsums) return;
  // column sums of the result (observables, norm, density, purity, overlap with W_0)
  double acc[kBCol];
#pragma unroll
  for (int k = 0; k < kBCol; ++k) acc[k] = 0.0;
#pragma unroll
  for (int j = 0; j < kR; ++j) {
    const int i = t + j * TT;
    const double p = a.p_host[i], kp = a.k_host[i];
    const double wr = v[j].x, wi = v[j].y;
    acc[0] += wr;
    acc[1] += wi;
    acc[2] += p * wr;
    acc[3] += p * p * wr;
    acc[4] += kp * wr;
    acc[5] += wr * wr - wi * wi;
    acc[6] += 2.0 * wr * wi;
    if (a.Wground) {
      const cplx<double> g = reinterpret_cast<const cplx<double>*>(a.Wground)[size_t(i) * X + col];
      acc[7] += wr * g.x - wi * g.y;
      acc[8] += wr * g.y + wi * g.x;
    }
  }
  __syncthreads();  // the transform's LDS is free
  double* red = reinterpret_cast<double*>(lds);
#pragma unroll
  for (int k = 0; k < kBCol; ++k) red[(t * C + c) * kBCol + k] = acc[k];
  __syncthreads();
  if (tid < C) {
    for (int k = 0; k < kBCol; ++k) {
      double s = 0.0;
      for (int q = 0; q < TT; ++q) s += red[(q * C + tid) * kBCol + k];
      a.colsums[size_t(blockIdx.x * C + tid) * kBCol + k] = s;
    }
  }
}

// ------------------------------------------------------------------ x sweep (row teams)
template <int N>
__global__ void __launch_bounds__(kBlock) bloch_xsweep_kernel(BlochArgs a) {
  constexpr int TT = N / kR, ROWS = kBlock / TT, LSZ = Padded<N, kR>::size;
  __shared__ cplx<double> lds[ROWS * LSZ];
  const int tid = threadIdx.x, team = tid / TT, t = tid % TT;
  const int row = blockIdx.x * ROWS + team;
  cplx<double>* W = reinterpret_cast<cplx<double>*>(a.W) + size_t(row) * N;
  const Twiddles<double> tw{reinterpret_cast<const cplx<double>*>(a.twX), nullptr};
  LdsRow<double> view{lds + team * LSZ};
  cplx<double> v[kR];
#pragma unroll
  for (int j = 0; j < kR; ++j) v[j] = W[t + j * TT];
  team_fft<double, N, kR, -1>(v, t, view, tw);
  const double dt = a.ctrl[kBDt];
  const double p = a.kdp * double(fidx_b(row, a.P));
#pragma unroll
  for (int j = 0; j < kR; ++j) {
    const double u = 0.5 * a.khbar * a.kdlambda * double(fidx_b(t + j * TT, N));
    const double e = hpoly(a.kcoef, a.kdeg, p + u) + hpoly(a.kcoef, a.kdeg, p - u);
    v[j] = cscale(v[j], exp(-0.5 * dt * e) * (1.0 / N));
  }
  team_fft<double, N, kR, +1>(v, t, view, tw);
#pragma unroll
  for (int j = 0; j < kR; ++j) W[t + j * TT] = v[j];
}

// ------------------------------------------------------------------ column sums of a state
// 32 columns per block, 8 row groups.  gp_density: also rho(x) = dP Re sum_p W
// (ProbabilityX, GP.py:1743-1748).  colsums == nullptr: the density only.
// 4-column tiles over 64 row groups of 4 lanes (one load instruction covers 16 rows of the
// tile's columns), group sums combined by a fixed tree in LDS
constexpr int kBColsumCols = 4;
__global__ void __launch_bounds__(kBlock) bloch_colsum_kernel(BlochArgs a, const void* state, double* colsums,
                                                             int density) {
  constexpr int CC = kBColsumCols, G = kBlock / CC;
  __shared__ double red[G][CC][kBCol];
  const int c = threadIdx.x % CC, g = threadIdx.x / CC;
  const int col = blockIdx.x * CC + c;
  const int X = a.X;
  const cplx<double>* W = reinterpret_cast<const cplx<double>*>(state);
  double acc[kBCol];
#pragma unroll
  for (int k = 0; k < kBCol; ++k) acc[k] = 0.0;
  if (col < X) {
#pragma unroll 4
    for (int i = g; i < a.P; i += G) {
      const cplx<double> v = W[size_t(i) * X + col];
      const double p = a.p_host[i];
      acc[0] += v.x;
      acc[1] += v.y;
      acc[2] += p * v.x;
      acc[3] += p * p * v.x;
      acc[4] += a.k_host[i] * v.x;
      acc[5] += v.x * v.x - v.y * v.y;
      acc[6] += 2.0 * v.x * v.y;
    }
  }
#pragma unroll
  for (int k = 0; k < kBCol; ++k) red[g][c][k] = acc[k];
  __syncthreads();
  for (int h = G / 2; h > 0; h >>= 1) {
    if (g < h) {
#pragma unroll
      for (int k = 0; k < kBCol; ++k) red[g][c][k] += red[g + h][c][k];
    }
    __syncthreads();
  }
  if (g == 0 && col < X) {
    if (colsums) {
#pragma unroll
      for (int k = 0; k < kBCol; ++k) colsums[size_t(col) * kBCol + k] = red[0][c][k];
    }
    if (density) a.rho[col] = a.dp_host * red[0][c][0];
  }
}

// ------------------------------------------------------------------ B- + B+ (one theta row per team)
// G(theta, x) = Re F_x^-1[ F_x(rho) 2 cos(lambda theta / 2) ] = rho(x - theta/2) + rho(x + theta/2)
// (the host Lambda and Theta grids, e^{i lambda theta / 2} without hbar: FFT.py:594-595)
template <int N>
__global__ void __launch_bounds__(kBlock) bloch_gshift_kernel(BlochArgs a) {
  constexpr int TT = N / kR, ROWS = kBlock / TT, LSZ = Padded<N, kR>::size;
  __shared__ cplx<double> lds[ROWS * LSZ];
  const int tid = threadIdx.x, team = tid / TT, t = tid % TT;
  const int row = blockIdx.x * ROWS + team;
  const Twiddles<double> tw{reinterpret_cast<const cplx<double>*>(a.twX), nullptr};
  LdsRow<double> view{lds + team * LSZ};
  cplx<double> v[kR];
#pragma unroll
  for (int j = 0; j < kR; ++j) v[j] = {a.rho[t + j * TT], 0.0};
  team_fft<double, N, kR, -1>(v, t, view, tw);
  const double half_theta = 0.5 * a.theta_host[row];
#pragma unroll
  for (int j = 0; j < kR; ++j) v[j] = cscale(v[j], 2.0 * cos(a.lambda_host[t + j * TT] * half_theta) * (1.0 / N));
  team_fft<double, N, kR, +1>(v, t, view, tw);
#pragma unroll
  for (int j = 0; j < kR; ++j) a.G[size_t(row) * N + t + j * TT] = v[j].x;
}

// ------------------------------------------------------------------ finish (one block)
// mode kBInit: the state before the loop (norm, E); kBStep: one step (record, acceptance from
// the second iteration on).  Complex quantities as the reference forms them: n = dXdP sum W,
// observables Re(dXdP sum W A / n), purity 2 pi hbar dXdP sum W^2 / n^2, rho = dP sum_p W / n,
// E = dXdP sum W H / n + dX a/2 sum rho^2.
__global__ void __launch_bounds__(kBlock) bloch_finish_kernel(BlochArgs a, int mode) {
  constexpr int K = 16;
  __shared__ double red[K][kBlock];
  __shared__ double cproj[2];
  const int tid = threadIdx.x;
  const int X = a.X;
  const bool excited = a.Wground != nullptr && mode == kBStep;
  // projection coefficient c = 2 pi hbar dXdP sum W W_0 (first pass), then everything else
  if (excited) {
    double s0 = 0.0, s1 = 0.0;
    for (int j = tid; j < X; j += kBlock) {
      s0 += a.colsums[size_t(j) * kBCol + 7];
      s1 += a.colsums[size_t(j) * kBCol + 8];
    }
    red[0][tid] = s0;
    red[1][tid] = s1;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
      if (tid < h) {
        red[0][tid] += red[0][tid + h];
        red[1][tid] += red[1][tid + h];
      }
      __syncthreads();
    }
    if (tid == 0) {
      const double f = 2.0 * kPiB * a.hbar_host * a.dxdp;
      cproj[0] = f * red[0][0];
      cproj[1] = f * red[1][0];
    }
    __syncthreads();
  }
  const double cr = excited ? cproj[0] : 0.0, ci = excited ? cproj[1] : 0.0;
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.0;
  for (int j = tid; j < X; j += kBlock) {
    const double* s = a.colsums + size_t(j) * kBCol;
    double S0 = s[0], S1 = s[1], Sp = s[2], Sp2 = s[3], SK = s[4], Q0 = s[5], Q1 = s[6];
    if (excited) {
      // W - c W_0 (W_0 real up to rounding; its imaginary part enters S1 and Q only)
      const double* g = a.g_colsums + size_t(j) * kBCol;
      const double o0 = s[7], o1 = s[8];  // sum W W_0 of this column
      S0 -= cr * g[0] - ci * g[1];
      S1 -= cr * g[1] + ci * g[0];
      Sp -= cr * g[2];
      Sp2 -= cr * g[3];
      SK -= cr * g[4];
      // sum (W - c W0)^2 = sum W^2 - 2 c sum W W0 + c^2 sum W0^2
      const double c2r = cr * cr - ci * ci, c2i = 2.0 * cr * ci;
      Q0 += -2.0 * (cr * o0 - ci * o1) + (c2r * g[5] - c2i * g[6]);
      Q1 += -2.0 * (cr * o1 + ci * o0) + (c2r * g[6] + c2i * g[5]);
    }
    const double x = a.x_host[j], v0 = a.v0_host[j];
    acc[0] += S0;
    acc[1] += S1;
    acc[2] += x * S0;
    acc[3] += x * x * S0;
    acc[4] += Sp;
    acc[5] += Sp2;
    acc[6] += SK + v0 * S0;
    acc[8] += Q0;
    acc[9] += Q1;
    // sum_x (sum_p W)^2 (complex), for the non-linear energy
    acc[10] += S0 * S0 - S1 * S1;
    acc[11] += 2.0 * S0 * S1;
  }
#pragma unroll
  for (int k = 0; k < 12; ++k) red[k][tid] = acc[k];
  __syncthreads();
  for (int h = kBlock / 2; h > 0; h >>= 1) {
    if (tid < h) {
#pragma unroll
      for (int k = 0; k < 12; ++k) red[k][tid] += red[k][tid + h];
    }
    __syncthreads();
  }
  if (tid != 0) return;
  double r[12];
  for (int k = 0; k < 12; ++k) r[k] = red[k][0];
  // n = dXdP sum W (complex); 1/n
  const double nr = a.dxdp * r[0], ni = a.dxdp * r[1];
  const double m2 = nr * nr + ni * ni;
  const double ir = nr / m2, ii = -ni / m2;
  // Re(dXdP z / n) for a real sum z (imaginary parts of x, p sums are below rounding)
  const double d = a.dxdp * ir;
  // energy: dXdP sum W H / n (complex) + dX a/2 dP^2 sum S^2 / n^2
  // (the imaginary part of sum W H is of rounding size and enters only through 1/n)
  const double hr = a.dxdp * r[6];
  double er = hr * ir, ei = hr * ii;
  const double i2r = ir * ir - ii * ii, i2i = 2.0 * ir * ii;
  double nler = 0.0, nlei = 0.0;
  if (a.gp) {
    const double f = a.dx_host * 0.5 * a.gp_host * a.dp_host * a.dp_host;
    nler = f * (r[10] * i2r - r[11] * i2i);
    nlei = f * (r[10] * i2i + r[11] * i2r);
    er += nler;
    ei += nlei;
  }
  const double pf = 2.0 * kPiB * a.hbar_host * a.dxdp;
  const double pr = pf * (r[8] * i2r - r[9] * i2i), pim = pf * (r[8] * i2i + r[9] * i2r);
  double* c = a.ctrl;
  if (mode == kBInit) {
    c[kBNr] = nr;
    c[kBNi] = ni;
    c[kBEr] = er;
    c[kBEi] = ei;
    c[kBCr] = 0.0;
    c[kBCi] = 0.0;
    c[kBDecision] = kBKeepNoCopy;
    return;
  }
  const int it = int(c[kBIter]);
  double* h = a.hist + size_t(it - a.hist_base) * kBHist;
  h[0] = c[kBDt];
  h[1] = d * r[2];
  h[2] = d * r[3];
  h[3] = d * r[4];
  h[4] = d * r[5];
  h[5] = er - nler;  // Hamiltonian_average: dXdP sum W H
  h[6] = er;
  h[7] = nler;
  h[8] = pr;
  h[9] = pim;
  c[kBNr] = nr;
  c[kBNi] = ni;
  c[kBCr] = cr;
  c[kBCi] = ci;
  double decision = kBKeepNoCopy;  // the first iteration is never tested (GP.py:2276)
  if (it > a.first_test) {
    const bool lower = er < c[kBEr] || (er == c[kBEr] && ei < c[kBEi]);
    if (lower && sqrt(pr * pr + pim * pim) < 1.0) {
      decision = kBKeep;
      c[kBEr] = er;
      c[kBEi] = ei;
    } else {
      decision = kBRefuse;
    }
  }
  h[10] = decision == kBKeep ? 1.0 : decision == kBRefuse ? 0.0 : -1.0;
  h[11] = ei;
  c[kBDecision] = decision;
}

// ------------------------------------------------------------------ commit
// kept: W = (W - c W_0) / n (and W_step = W when the step was tested); refused: W = W_step and
// dt halves; last block to finish advances the iteration counter (separate kernel below).
__global__ void __launch_bounds__(kBlock) bloch_commit_kernel(BlochArgs a) {
  const double decision = a.ctrl[kBDecision];
  const size_t n = size_t(a.X) * size_t(a.P);
  cplx<double>* W = reinterpret_cast<cplx<double>*>(a.W);
  cplx<double>* S = reinterpret_cast<cplx<double>*>(a.Wstep);
  const cplx<double>* G = reinterpret_cast<const cplx<double>*>(a.Wground);
  const double nr = a.ctrl[kBNr], ni = a.ctrl[kBNi], m2 = nr * nr + ni * ni;
  const cplx<double> inv{nr / m2, -ni / m2};
  const cplx<double> cp{a.ctrl[kBCr], a.ctrl[kBCi]};
  for (size_t e = size_t(blockIdx.x) * kBlock + threadIdx.x; e < n; e += size_t(gridDim.x) * kBlock) {
    if (decision == kBRefuse) {
      W[e] = S[e];
    } else {
      cplx<double> w = W[e];
      if (G) w = csub(w, cmul(cp, G[e]));
      w = cmul(w, inv);
      W[e] = w;
      if (decision == kBKeep) S[e] = w;
    }
  }
}

__global__ void bloch_advance_kernel(double* ctrl) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    if (ctrl[kBDecision] == kBRefuse) ctrl[kBDt] *= 0.5;
    ctrl[kBIter] += 1.0;
  }
}

__global__ void bloch_set_kernel(double* ctrl, double dt, double iter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    ctrl[kBDt] = dt;
    ctrl[kBIter] = iter;
  }
}

#define WH_BLOCH_SWITCH(N, CALL)   \
  switch (N) {                     \
    case 128: CALL(128); break;    \
    case 256: CALL(256); break;    \
    case 512: CALL(512); break;    \
    case 1024: CALL(1024); break;  \
    case 2048: CALL(2048); break;  \
    default: return hipErrorInvalidValue; \
  }

}  // namespace

hipError_t launch_bloch_psweep(const BlochArgs& a, int colsums, hipStream_t s) {
#define CALL(n) hipLaunchKernelGGL(bloch_psweep_kernel<n>, dim3(a.X / (kBlock / (n / kR))), dim3(kBlock), 0, s, a, colsums)
  WH_BLOCH_SWITCH(a.P, CALL)
#undef CALL
  return hipGetLastError();
}
hipError_t launch_bloch_xsweep(const BlochArgs& a, hipStream_t s) {
#define CALL(n) hipLaunchKernelGGL(bloch_xsweep_kernel<n>, dim3(a.P / (kBlock / (n / kR))), dim3(kBlock), 0, s, a)
  WH_BLOCH_SWITCH(a.X, CALL)
#undef CALL
  return hipGetLastError();
}
hipError_t launch_bloch_gshift(const BlochArgs& a, hipStream_t s) {
#define CALL(n) hipLaunchKernelGGL(bloch_gshift_kernel<n>, dim3(a.P / (kBlock / (n / kR))), dim3(kBlock), 0, s, a)
  WH_BLOCH_SWITCH(a.X, CALL)
#undef CALL
  return hipGetLastError();
}
hipError_t launch_bloch_colsum(const BlochArgs& a, const void* state, double* colsums, int density, hipStream_t s) {
  hipLaunchKernelGGL(bloch_colsum_kernel, dim3((a.X + kBColsumCols - 1) / kBColsumCols), dim3(kBlock), 0, s, a, state,
                     colsums, density);
  return hipGetLastError();
}
hipError_t launch_bloch_finish(const BlochArgs& a, int mode, hipStream_t s) {
  hipLaunchKernelGGL(bloch_finish_kernel, dim3(1), dim3(kBlock), 0, s, a, mode);
  return hipGetLastError();
}
hipError_t launch_bloch_commit(const BlochArgs& a, hipStream_t s) {
  const size_t n = size_t(a.X) * size_t(a.P);
  const int blocks = int(std::min<size_t>((n + kBlock - 1) / kBlock, 4096));
  hipLaunchKernelGGL(bloch_commit_kernel, dim3(blocks), dim3(kBlock), 0, s, a);
  hipLaunchKernelGGL(bloch_advance_kernel, dim3(1), dim3(64), 0, s, a.ctrl);
  return hipGetLastError();
}
hipError_t launch_bloch_set(double* ctrl, double dt, int iter, hipStream_t s) {
  hipLaunchKernelGGL(bloch_set_kernel, dim3(1), dim3(64), 0, s, ctrl, dt, double(iter));
  return hipGetLastError();
}

}  // namespace wh
