// Adjoint (reverse-time) step of the Lie split operator, for gradients of a loss
// of the final state with respect to the polynomial coefficients a_k(n) of V.
//
// Forward step n (GPU_Wigner2D_FFT.py:1884-1899): W_{n+1} = F_p^-1 M_n F_p K W_n,
// M_n = r exp(-i phi_n), phi_n = vphase (V_n(x - u) - V_n(x + u)), u = hbar theta / 2.
// Adjoint state: Lambda_n = U_n^dagger Lambda_{n+1}
//                         = K^dagger F_p^-1 conj(M_n) F_p Lambda_{n+1}
// (p sweep with the conjugated potential multiplier, then the x sweep with the
// conjugated kinetic multiplier).  Gradient (Parseval along p, numpy conventions):
//   dL/da_k(n) = Re sum conj(Lambda_{n+1}) dU_n/da_k W_n
//              = (vphase / P) sum_{theta,x} g_k Im( conj(F_p Lambda_{n+1}) F_p W_{n+1} ),
//   g_k = (x - u)^k - (x + u)^k,
// so the adjoint p sweep reads Lambda_{n+1} and the forward state W_{n+1} and
// accumulates the k <= 8 column sums of g_k Im(...) in its epilogue.
#include <hip/hip_runtime.h>

#include "wh_sweeps.hpp"

namespace wh {

template <typename T, int N, int R, int BLOCK>
__device__ __forceinline__ void adj_psweep_body(const SweepArgs& a, const Member& m, int tile, cplx<T>* lds) {
  constexpr int TT = N / R, C = BLOCK / TT;
  const int tid = threadIdx.x;
  const int c = tid % C, t = tid / C;
  const int X = m.X;
  const int col = tile * C + c;
  cplx<T>* __restrict__ L = reinterpret_cast<cplx<T>*>(m.W);
  cplx<T>* __restrict__ F = reinterpret_cast<cplx<T>*>(const_cast<void*>(m.W2));
  const Twiddles<T> tw{reinterpret_cast<const cplx<T>*>(m.twP_hi), reinterpret_cast<const cplx<T>*>(m.twP_lo)};
  LdsStrided<T, C> view{lds + c};
  const int step = *a.step + a.step_add;
  const unsigned boff = (unsigned(t) * unsigned(X) + unsigned(col)) * unsigned(sizeof(cplx<T>));

  cplx<T> v[R], w[R];
#pragma unroll
  for (int j = 0; j < R; ++j) {
    v[j] = at_off(L + size_t(j * TT) * X, boff);
    w[j] = at_off(F + size_t(j * TT) * X, boff);
  }
  team_fft<T, N, R, -1>(v, t, view, tw);
  team_fft<T, N, R, -1>(w, t, view, tw);

  const double x = m.dx * double(fidx(col, X));
  const int row = step < m.pot_rows ? step : m.pot_rows - 1;
  const double* __restrict__ cf = m.pot + size_t(row) * (kMaxDeg + 1);
  const double mask = m.xmask ? m.xmask[col] : 1.0;
  const int deg = m.pot_deg;
  double acc[kMaxDeg + 1];
#pragma unroll
  for (int k = 0; k <= kMaxDeg; ++k) acc[k] = 0.0;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int i = t + j * TT;
    const double th = m.dtheta * double(fidx(i, N));
    const double u = a.hbar_half * th;
    const double xm = x - u, xp = x + u;
    // Im(conj(v) w)
    const double im = double(v[j].x) * double(w[j].y) - double(v[j].y) * double(w[j].x);
    // every coefficient up to kMaxDeg: dL/da_k is not zero where a_k currently is
    double pm = 1.0, pp = 1.0;
#pragma unroll
    for (int k = 0; k <= kMaxDeg; ++k) {
      acc[k] += (pm - pp) * im;
      pm *= xm;
      pp *= xp;
    }
    const double phi = a.vphase * (poly(cf, deg, xm) - poly(cf, deg, xp));
    const double r = (m.theta_decay ? m.theta_decay[i] : 1.0) * mask;
    v[j] = apply(v[j], expmi(r, -phi));  // conj(M_n)
  }
  team_fft<T, N, R, +1>(v, t, view, tw);
  const T inv = T(1.0 / N);
#pragma unroll
  for (int j = 0; j < R; ++j) at_off(L + size_t(j * TT) * X, boff) = cscale(v[j], inv);

  // column sums of the gradient terms (lanes c + C*t of a wave hold column c)
#pragma unroll
  for (int off = C; off < 64; off <<= 1) {
#pragma unroll
    for (int k = 0; k <= kMaxDeg; ++k) acc[k] += __shfl_xor(acc[k], off, 64);
  }
  constexpr int kWaves = (BLOCK + 63) / 64;
  constexpr int kLanes = C < 64 ? C : 64;
  __syncthreads();
  double* red = reinterpret_cast<double*>(lds);
  const int wave = tid >> 6, lane = tid & 63;
  if (lane < kLanes) {
#pragma unroll
    for (int k = 0; k <= kMaxDeg; ++k) red[(wave * kLanes + lane) * (kMaxDeg + 1) + k] = acc[k];
  }
  __syncthreads();
  if (tid < C) {
    for (int k = 0; k <= kMaxDeg; ++k) {
      double s = 0.0;
      for (int q = 0; q < kWaves; ++q) s += red[(q * kLanes + tid) * (kMaxDeg + 1) + k];
      m.gcols[size_t(col - c + tid) * (kMaxDeg + 1) + k] = s;
    }
  }
}

template <typename T, int R, int BLOCK>
__global__ void __launch_bounds__(BLOCK) adj_psweep_kernel(SweepArgs a) {
  __shared__ cplx<T> lds[lds_elems<BLOCK, R>()];
  const unsigned w = a.work[blockIdx.x];
  const Member& m = a.members[w >> 16];
  const int tile = int(w & 0xffffu);
#define BODY(n) adj_psweep_body<T, n, R, BLOCK>(a, m, tile, lds)
  WH_DISPATCH(m.P, BODY)
#undef BODY
}

// grad[k] = scale * sum_x gcols[x][k]   (one block per member; member 0 for single plans)
__global__ void __launch_bounds__(256) grad_reduce_kernel(const Member* members, double* grad, double scale) {
  __shared__ double red[kMaxDeg + 1][256];
  const Member& m = members[blockIdx.x];
  double acc[kMaxDeg + 1];
#pragma unroll
  for (int k = 0; k <= kMaxDeg; ++k) acc[k] = 0.0;
  for (int x = threadIdx.x; x < m.X; x += 256) {
#pragma unroll
    for (int k = 0; k <= kMaxDeg; ++k) acc[k] += m.gcols[size_t(x) * (kMaxDeg + 1) + k];
  }
#pragma unroll
  for (int k = 0; k <= kMaxDeg; ++k) red[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
#pragma unroll
      for (int k = 0; k <= kMaxDeg; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x <= kMaxDeg)
    grad[size_t(blockIdx.x) * (kMaxDeg + 1) + threadIdx.x] = scale * red[threadIdx.x][0];
}

// Adjoint tile geometry: radix 8, 256 lanes (two register arrays per lane).
constexpr int kAdjR = 8, kAdjBlock = 256;
int adj_cols(int P) { return kAdjBlock / (P / kAdjR); }

hipError_t launch_adj_psweep(int prec, const SweepArgs& a, int nwork, hipStream_t s) {
  if (prec == 0)
    hipLaunchKernelGGL((adj_psweep_kernel<double, kAdjR, kAdjBlock>), dim3(nwork), dim3(kAdjBlock), 0, s, a);
  else
    hipLaunchKernelGGL((adj_psweep_kernel<float, kAdjR, kAdjBlock>), dim3(nwork), dim3(kAdjBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_grad_reduce(const Member* members, int nmembers, double* grad, double scale, hipStream_t s) {
  hipLaunchKernelGGL(grad_reduce_kernel, dim3(nmembers), dim3(256), 0, s, members, grad, scale);
  return hipGetLastError();
}

}  // namespace wh
